// Device generator of the benchmark problems (the same arrays as
// lbfgsb_amd/problems.py builds on the host): a request of the serving
// workload names its objective, size and seed, and its data are written
// straight into its HBM slot instead of crossing PCIe.
//
// Every draw is splitmix64 of (seed ^ stream * phi64 ^ counter), so a value
// depends on (seed, stream, counter) only.  The uniform draws and what is
// computed from them with + - * are bit-identical to the host's; the normal
// draws (Box-Muller: a log and a cosine) may differ from the host's libm in
// the last bit.  The file is compiled with -ffp-contract=off, so a * b + c
// rounds twice, as numpy does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "generate.hpp"
#include "objectives.hpp"

namespace lbfgsb {

namespace {

constexpr uint64_t kPhi64 = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kMix1 = 0xBF58476D1CE4E5B9ull;
constexpr uint64_t kMix2 = 0x94D049BB133111EBull;

// streams of the draws (problems.py)
constexpr uint64_t kStreamB = 1, kStreamC = 2, kStreamX0 = 3, kStreamBound = 4, kStreamNbd = 5,
                   kStreamA = 6;

__device__ __forceinline__ double uniform_at(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t z = (seed ^ (stream * kPhi64)) ^ i;
  z += kPhi64;
  z = (z ^ (z >> 30)) * kMix1;
  z = (z ^ (z >> 27)) * kMix2;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Box-Muller on the counter pair (2i, 2i+1), u1 taken as 1 - u.
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t stream, uint64_t i) {
  const double u1 = 1.0 - uniform_at(seed, stream, 2 * i);
  const double u2 = uniform_at(seed, stream, 2 * i + 1);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

template <typename T>
__global__ void k_generate(GenSpec sp, T* x0, int* nbd, T* l, T* u, T* b, T* c, T* a0, T* a1,
                           T* a2) {
  const long long n = sp.n;
  const uint64_t seed = sp.seed;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const uint64_t ui = (uint64_t)i;
    switch (sp.kind) {
      case OBJK_GRID_QUAD:
        x0[i] = (T)0.0;
        nbd[i] = 2;
        l[i] = (T)-0.05;
        u[i] = (T)0.05;
        b[i] = (T)uniform_at(seed, kStreamB, ui);
        break;
      case OBJK_GRID_COS: {
        const double x = uniform_at(seed, kStreamX0, ui) - 0.5;
        const double o0 = uniform_at(seed, kStreamBound, 2 * ui);
        const double o1 = uniform_at(seed, kStreamBound, 2 * ui + 1);
        const double draw = uniform_at(seed, kStreamNbd, ui);
        x0[i] = (T)x;
        nbd[i] = draw < 0.2 ? 0 : (draw < 0.5 ? 1 : 2);
        l[i] = (T)(x - (0.1 + 0.9 * o0));
        u[i] = (T)(x + (0.1 + 0.9 * o1));
        b[i] = (T)normal_at(seed, kStreamB, ui);
        c[i] = (T)(0.5 + 1.5 * uniform_at(seed, kStreamC, ui));
        break;
      }
      case OBJK_ROSEN: {
        const bool odd = (i % 2) == 0;  // 1-based odd index
        x0[i] = (T)(-1.2 + 2.4 * uniform_at(seed, kStreamX0, ui));
        nbd[i] = 2;
        l[i] = (T)(odd ? -1.5 : -0.5);
        u[i] = (T)(odd ? 2.0 : 1.5);
        break;
      }
      default: {  // OBJK_PENTA
        const double d0 = uniform_at(seed, kStreamA, 3 * ui);
        const double d1 = uniform_at(seed, kStreamA, 3 * ui + 1);
        const double d2 = uniform_at(seed, kStreamA, 3 * ui + 2);
        const double o0 = uniform_at(seed, kStreamBound, 2 * ui);
        const double o1 = uniform_at(seed, kStreamBound, 2 * ui + 1);
        x0[i] = (T)0.0;
        nbd[i] = 2;
        l[i] = (T)(-(0.1 + 0.9 * o0));
        u[i] = (T)(0.1 + 0.9 * o1);
        b[i] = (T)normal_at(seed, kStreamB, ui);
        a0[i] = (T)(4.0 + 4.0 * d0);
        a1[i] = (T)(i < n - 1 ? (2.0 * d1 - 1.0) * 0.9 : 0.0);
        a2[i] = (T)(i < n - 2 ? (2.0 * d2 - 1.0) * 0.9 : 0.0);
        break;
      }
    }
  }
}

}  // namespace

int gen_check(const GenSpec& sp) {
  if (sp.n <= 0) return -1;
  switch (sp.kind) {
    case OBJK_GRID_QUAD:
    case OBJK_GRID_COS:
      return (sp.nx > 0 && sp.ny > 0 && sp.nx * sp.ny == sp.n) ? 0 : -1;
    case OBJK_ROSEN:
      return sp.n % 2 == 0 ? 0 : -1;
    case OBJK_PENTA:
      return 0;
    default:
      return -1;
  }
}

template <typename T>
int launch_generate(hipStream_t s, const GenSpec& sp, T* x0, int* nbd, T* l, T* u, T* b, T* c,
                    T* a0, T* a1, T* a2) {
  if (gen_check(sp) != 0) return -1;
  const bool grid = sp.kind == OBJK_GRID_QUAD || sp.kind == OBJK_GRID_COS;
  if (!x0 || !nbd || !l || !u) return -1;
  if ((grid || sp.kind == OBJK_PENTA) && !b) return -1;
  if (sp.kind == OBJK_GRID_COS && !c) return -1;
  if (sp.kind == OBJK_PENTA && (!a0 || !a1 || !a2)) return -1;
  const long long blocks = (sp.n + 255) / 256;
  const int g = (int)(blocks < 4096 ? blocks : 4096);
  hipLaunchKernelGGL(k_generate<T>, dim3(g), dim3(256), 0, s, sp, x0, nbd, l, u, b, c, a0, a1,
                     a2);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template int launch_generate<float>(hipStream_t, const GenSpec&, float*, int*, float*, float*,
                                    float*, float*, float*, float*, float*);
template int launch_generate<double>(hipStream_t, const GenSpec&, double*, int*, double*,
                                     double*, double*, double*, double*, double*, double*);

template <typename S, typename D>
__global__ void __launch_bounds__(kBlock)
k_convert_copy(const S* __restrict__ src, D* __restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    dst[i] = (D)src[i];
}

int launch_convert_copy(hipStream_t s, const void* src, int src_dtype, void* dst, int dst_dtype,
                        long long n) {
  if (n <= 0) return 0;
  long long blocks = (n + kBlock - 1) / kBlock;
  if (blocks > kMaxBlocks) blocks = kMaxBlocks;
  const dim3 g((unsigned)blocks), b(kBlock);
  if (src_dtype == 0 && dst_dtype == 1)
    hipLaunchKernelGGL((k_convert_copy<float, double>), g, b, 0, s, (const float*)src,
                       (double*)dst, n);
  else if (src_dtype == 1 && dst_dtype == 0)
    hipLaunchKernelGGL((k_convert_copy<double, float>), g, b, 0, s, (const double*)src,
                       (float*)dst, n);
  else if (src_dtype == 1)
    hipLaunchKernelGGL((k_convert_copy<double, double>), g, b, 0, s, (const double*)src,
                       (double*)dst, n);
  else
    hipLaunchKernelGGL((k_convert_copy<float, float>), g, b, 0, s, (const float*)src,
                       (float*)dst, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace lbfgsb
