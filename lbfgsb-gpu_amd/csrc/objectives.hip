// Objective kernels for the benchmark workloads (the callback side of the
// boundary).  They run on the callback's stream, write g in the bulk type and
// per-block FP64 partials of f, which one workgroup finishes on the device.
// The per-term formulas are in obj_device.hpp.
// The shape follows the reference's example callback
// (examples/dsscfg/dsscfg_cuda.cu:43-133): a device kernel for g and an
// energy reduction whose value is handed back as the host f.
#include "common.hpp"
#include "host_mailbox.hpp"
#include "launch.hpp"
#include "obj_device.hpp"
#include "objectives.hpp"

namespace lbfgsb {

// Terms per thread per step of k_obj, every term's loads issued before the
// first one is evaluated: four Rosenbrock pairs (a streaming term: 1.196 ->
// 1.185 ms per 16.8M iteration); one term of the other kinds, whose
// neighbours' loads already hit the cache (four made the 2048^2 cos grid
// slower, 0.584 -> 0.593 ms per iteration).
template <int KIND>
constexpr int obj_unroll() {
  return KIND == OBJK_ROSEN ? 4 : 1;
}

// The terms of an objective of kind KIND, U per thread per step (terms k,
// k + S, .., k + (U - 1) S with S the grid's thread count), added to the
// thread's share of f in the order a one-term loop adds them.
template <int KIND, typename T>
__global__ void __launch_bounds__(kBlock)
k_obj(ObjDesc o, long long n, const T* __restrict__ x, T* __restrict__ g, double* part) {
  constexpr int kObjU = obj_unroll<KIND>();
  __shared__ double sh[kWavesPerBlock];
  double acc[1] = {0.0};
  const long long count = obj_terms(o, n);
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long kb = blockIdx.x * (long long)kBlock + threadIdx.x; kb < count;
       kb += kObjU * stride) {
    ObjOut<T> r[kObjU];
    double fk[kObjU];
#pragma unroll
    for (int q = 0; q < kObjU; q++) {
      const long long k = kb + q * stride;
      fk[q] = obj_eval_kind<KIND>(o, n, x, k < count ? k : count - 1, r[q]);
    }
#pragma unroll
    for (int q = 0; q < kObjU; q++) {
      const long long k = kb + q * stride;
      if (k < count) {
        obj_store_kind<KIND>(g, k, r[q]);
        acc[0] += fk[q];
      }
    }
  }
  block_sum_store<1>(acc, sh, part + blockIdx.x, gridDim.x);
}

// f into f_out on the device and, when hf is given, into the host's pinned
// mailbox followed by its sequence number.  One wave.
__global__ void __launch_bounds__(kWave)
s_obj_finish(const double* part, int G, double* f_out, HostF* hf, unsigned seq) {
  const int lane = threadIdx.x & 63;
  // every partial of the lane (b = lane + 64 k < G <= kMaxBlocks) loaded
  // before the first two-sum, then summed in order of b
  constexpr int kPer = kMaxBlocks / kWave;
  double pv[kPer];
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int b = lane + k * kWave;
    pv[k] = b < G ? part[b] : 0.0;
  }
  double s = 0.0, c = 0.0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    if (lane + k * kWave >= G) break;
    double e;
    two_sum(s, pv[k], s, e);
    c += e;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double os = __shfl_xor(s, o, kWave), oc = __shfl_xor(c, o, kWave), e;
    two_sum(s, os, s, e);
    c += oc + e;
  }
  if (lane == 0) {
    *f_out = s + c;
    if (hf) hf->f = s + c;
  }
  if (hf) mail_publish(&hf->s, seq);
}

template <typename T>
void launch_objective(hipStream_t s, const ObjDesc& o, long long n, const T* x, T* g,
                      double* part, int G, double* f_dev, HostF* hf, unsigned seq) {
  switch (o.kind) {
    case OBJK_GRID_QUAD:
      hipLaunchKernelGGL((k_obj<OBJK_GRID_QUAD, T>), dim3(G), dim3(kBlock), 0, s, o, n, x, g, part);
      break;
    case OBJK_GRID_COS:
      hipLaunchKernelGGL((k_obj<OBJK_GRID_COS, T>), dim3(G), dim3(kBlock), 0, s, o, n, x, g, part);
      break;
    case OBJK_DSSCFG:
      hipLaunchKernelGGL((k_obj<OBJK_DSSCFG, T>), dim3(G), dim3(kBlock), 0, s, o, n, x, g, part);
      break;
    case OBJK_ROSEN:
      hipLaunchKernelGGL((k_obj<OBJK_ROSEN, T>), dim3(G), dim3(kBlock), 0, s, o, n, x, g, part);
      break;
    default:
      hipLaunchKernelGGL((k_obj<OBJK_PENTA, T>), dim3(G), dim3(kBlock), 0, s, o, n, x, g, part);
      break;
  }
  hipLaunchKernelGGL(s_obj_finish, dim3(1), dim3(kWave), 0, s, part, G, f_dev, hf, seq);
}

template void launch_objective<float>(hipStream_t, const ObjDesc&, long long, const float*,
                                      float*, double*, int, double*, HostF*, unsigned);
template void launch_objective<double>(hipStream_t, const ObjDesc&, long long, const double*,
                                       double*, double*, int, double*, HostF*, unsigned);

}  // namespace lbfgsb
