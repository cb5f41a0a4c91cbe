// Device scratch region of the allocation-free solve (lbfgsb_minimize_noalloc_*,
// capi.cpp).  One file-scope array per GPU the library runs on, held until
// the process ends; its device address is looked up once per device with
// hipGetSymbolAddress, and only that address is handed to host code and
// kernels (the array's name is a host-side stand-in in host code).
//
// Every allocation-free call on a GPU shares that GPU's region: a call
// overlapping another on the same GPU races with it.  The region keeps the
// last call's contents; the engine clears what must start at zero (the S/Y
// ring) with the fill kernel below and writes everything else before it
// reads it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "engine.hpp"

namespace lbfgsb {

// 64 MiB per GPU: an fp64 solve of up to 291,584 variables, fp32 485,888
// (m = 8; engine_region_bytes gives the exact need of a size).
constexpr size_t kNoallocBytes = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_region_mu;
void* g_region[kMaxDevices] = {};
}  // namespace

int noalloc_region(void** ptr, size_t* bytes) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_region_mu);
  if (!g_region[dev]) {
    // with this device current: its own copy of the array
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    g_region[dev] = p;
  }
  *ptr = g_region[dev];
  *bytes = kNoallocBytes;
  return 0;
}

__global__ void __launch_bounds__(kBlock) k_fill_zero(unsigned long long* p, size_t words) {
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < words;
       i += (size_t)gridDim.x * kBlock)
    p[i] = 0ull;
}

// Zeroes `bytes` (a multiple of 8) at p with a kernel, not a memset node.
void launch_fill_zero(hipStream_t s, void* p, size_t bytes) {
  const size_t words = bytes / 8;
  if (!words) return;
  size_t blocks = (words + kBlock - 1) / kBlock;
  if (blocks > (size_t)kMaxBlocks) blocks = kMaxBlocks;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(kBlock), 0, s,
                     (unsigned long long*)p, words);
}

}  // namespace lbfgsb
