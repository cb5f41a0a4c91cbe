// Single-workgroup FP64 control kernels: they finish the per-block partial
// sums of the bulk passes and run the scalar logic of the reference driver
// on the device (lbfgsbcpu.cpp:279-462), so that a whole iteration can be
// enqueued without a host round trip.  One launch = one block of 256
// threads.  Each kernel copies the device state block into LDS, runs the
// scalar logic there and writes it back: the decisions on thread 0, the
// small dense algebra shared out over the workgroup where its entries are
// independent and kept in registers where it is a dependent chain.
#include "common.hpp"
#include "dense.hpp"
#include "control.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

#define A1(a, ld, i, j) (a)[((i)-1) * (ld) + ((j)-1)]

static_assert(sizeof(DevScalars) % 8 == 0, "DevScalars is copied as 8-byte words");
constexpr int kDsWords = sizeof(DevScalars) / 8;

__device__ void ds_load(DevScalars* sds, const DevScalars* ds) {
  const unsigned long long* src = (const unsigned long long*)ds;
  unsigned long long* dst = (unsigned long long*)sds;
  for (int i = threadIdx.x; i < kDsWords; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}
__device__ void ds_store(DevScalars* ds, const DevScalars* sds) {
  __syncthreads();
  const unsigned long long* src = (const unsigned long long*)sds;
  unsigned long long* dst = (unsigned long long*)ds;
  for (int i = threadIdx.x; i < kDsWords; i += blockDim.x) dst[i] = src[i];
}

// Compensated reduction of value rows v0..v0+nv (each G partials long) into
// out[0..nv): one wave per row, four independent accumulators per lane,
// two-sum combination across accumulators and lanes.
__device__ void reduce_rows(const double* part, int G, int v0, int nv, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int v = w; v < nv; v += kWavesPerBlock) {
    const double* row = part + (size_t)(v0 + v) * G;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = lane;
    for (; b + 3 * kWave < G; b += 4 * kWave) {
      a0 += row[b];
      a1 += row[b + kWave];
      a2 += row[b + 2 * kWave];
      a3 += row[b + 3 * kWave];
    }
    for (; b < G; b += kWave) a0 += row[b];
    double s, c = 0.0, e;
    two_sum(a0, a1, s, e);
    c += e;
    two_sum(s, a2, s, e);
    c += e;
    two_sum(s, a3, s, e);
    c += e;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double os = __shfl_xor(s, o, kWave), oc = __shfl_xor(c, o, kWave);
      two_sum(s, os, s, e);
      c += oc + e;
    }
    if (lane == 0) out[v] = s + c;
  }
}

__device__ double reduce_max(const double* pm, int G, double* sh) {
  double m = -__builtin_inf();
  for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmax(m, pm[b]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < kWavesPerBlock; w++) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}

__device__ void reduce_argmin(const double* pv, const long long* pi, int G, double* shv,
                              long long* shi, double& v, long long& idx) {
  double bv = __builtin_inf();
  long long bi = 0x7fffffffffffffffLL;
  for (int b = threadIdx.x; b < G; b += blockDim.x) {
    const double x = pv[b];
    const long long k = pi[b];
    if (x < bv || (x == bv && k < bi)) {
      bv = x;
      bi = k;
    }
  }
  wave_argmin(bv, bi);
  if ((threadIdx.x & 63) == 0) {
    shv[threadIdx.x >> 6] = bv;
    shi[threadIdx.x >> 6] = bi;
  }
  __syncthreads();
  v = shv[0];
  idx = shi[0];
  for (int w = 1; w < kWavesPerBlock; w++)
    if (shv[w] < v || (shv[w] == v && shi[w] < idx)) {
      v = shv[w];
      idx = shi[w];
    }
  __syncthreads();
}

__device__ void reduce_min(const double* pv, int G, double* sh, double& v) {
  double m = __builtin_inf();
  for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmin(m, pv[b]);
  m = -wave_max(-m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  v = sh[0];
  for (int w = 1; w < kWavesPerBlock; w++) v = fmin(v, sh[w]);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Multi-block finish of per-block partials: block b sums row v0+b (G
// entries) into ds->red[b] with a compensated block tree.  p4_rows skips the
// rows of k_p4 that the iteration does not use (entering/leaving Gram when
// nothing entered or left, the new formk column when no update happened).
__global__ void __launch_bounds__(kBlock)
k_rowsum(const double* part, int G, int v0, DevScalars* ds, int p4_rows) {
  __shared__ double ss[kWavesPerBlock], sc[kWavesPerBlock];
  if (ds->halt) return;
  const int v = blockIdx.x;
  if (p4_rows) {
    const bool skip = (v >= P4_CORR && !ds->p4_flagged) ||
                      (v >= P4_NC && v < P4_NC + 32 && !ds->updatd);
    if (skip) {
      if (threadIdx.x == 0) ds->red[v] = 0.0;
      return;
    }
  }
  const double* row = part + (size_t)(v0 + v) * G;
  double a0 = 0.0, a1 = 0.0;
  int b = threadIdx.x;
  for (; b + kBlock < G; b += 2 * kBlock) {
    a0 += row[b];
    a1 += row[b + kBlock];
  }
  for (; b < G; b += kBlock) a0 += row[b];
  double s, c, e;
  two_sum(a0, a1, s, c);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double os = __shfl_xor(s, o, kWave), oc = __shfl_xor(c, o, kWave);
    two_sum(s, os, s, e);
    c += oc + e;
  }
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s;
    sc[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = ss[0], tc = sc[0];
    for (int w = 1; w < kWavesPerBlock; w++) {
      two_sum(t, ss[w], t, e);
      tc += sc[w] + e;
    }
    ds->red[v] = t + tc;
  }
}

__global__ void __launch_bounds__(kBlock)
s_resume(DevScalars* ds, int reset_bp) {
  if (threadIdx.x == 0) logic_resume(ds, reset_bp);
}

__global__ void __launch_bounds__(kBlock)
s_setup_errclb(DevScalars* ds, const double* part, int G) {
  __shared__ double r[2];
  reduce_rows(part, G, 0, 2, r);
  __syncthreads();
  if (threadIdx.x == 0) logic_setup_errclb(ds, r[0], r[1]);
}

__global__ void __launch_bounds__(kBlock)
s_setup_active(DevScalars* ds, const double* part, int G) {
  __shared__ double r[3];
  reduce_rows(part, G, 0, 3, r);
  __syncthreads();
  if (threadIdx.x == 0) logic_setup_active(ds, r);
}

// First projected-gradient test after the initial evaluation.
__global__ void __launch_bounds__(kBlock)
s_setup_projgr(DevScalars* ds, const double* pmax, int G, double f) {
  __shared__ double sh[kWavesPerBlock];
  double m = reduce_max(pmax, G, sh);
  if (threadIdx.x == 0) logic_setup_projgr(ds, m, f);
}

// Stopping tests after an accepted step (lbfgsbcpu.cpp:394-429); without a
// customised stopping callback this kernel also runs the memory update and
// the next Cauchy controller.
__global__ void __launch_bounds__(kBlock)
s_stop_tests(DevScalars* ds, const double* part, const double* pmax,
                             const double* pmin, const long long* pmini, int G) {
  __shared__ DevScalars sds;
  __shared__ double sh[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if (ds->halt) {
    if (threadIdx.x == 0) ds->write_trial = 0;
    return;
  }
  const double sbg = reduce_max(pmax, G, sh);
  double bk;
  long long ibk;
  reduce_argmin(pmin, pmini, G, sh, shi, bk, ibk);
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  (void)part;
  // s->red holds the row sums of the pass-A partials (k_rowsum)
  if (threadIdx.x == 0) logic_stop_tests(s, s->red, sbg, bk, ibk);
  __syncthreads();
  post_update_block(s);
  build_middle(s);
  ds_store(ds, s);
}

// Memory update + Cauchy controller.  `reduce` selects the standalone path
// (iteration start / restart) that reduces the classification partials here.
__global__ void __launch_bounds__(kBlock)
s_update_gcp(DevScalars* ds, const double* part, const double* pmin,
                             const long long* pmini, int G, int reduce) {
  __shared__ DevScalars sds;
  __shared__ double sh[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if (ds->halt) {
    if (threadIdx.x == 0) ds->write_trial = 0;
    return;
  }
  if (!reduce && ds->upd_done) return;
  double bk = 0.0;
  long long ibk = 0;
  if (reduce) reduce_argmin(pmin, pmini, G, sh, shi, bk, ibk);
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  (void)part;
  if (threadIdx.x == 0) logic_update_gcp(s, reduce, s->red, bk, ibk);
  __syncthreads();
  post_update_block(s);
  build_middle(s);
  ds_store(ds, s);
}

// After k_p4: free/active bookkeeping (lbfgsbcpu.cpp:1143-1161), formk,
// and the two triangular solves of subsm (lbfgsbcpu.cpp:1459-1471).
__global__ void __launch_bounds__(kBlock)
s_formk_subsm(DevScalars* ds, const double* part, int G) {
  __shared__ DevScalars sds;
  __shared__ int flags[1];
  if (ds->halt) return;
  (void)part;
  (void)G;
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  const double* r = s->red;  // row sums of the k_p4 partials (k_rowsum)
  if (threadIdx.x == 0) flags[0] = s->p4_flagged;
  __syncthreads();
  logic_formk_subsm(s, r, flags[0] != 0);
  ds_store(ds, s);
}

// Backtrack length of subsm (lbfgsbcpu.cpp:1484-1538).
__global__ void __launch_bounds__(kBlock)
s_alpha(DevScalars* ds, const double* pmin, const long long* pmini, int G) {
  __shared__ double sh[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if (ds->halt || ds->skip_subsm) return;
  double a;
  long long i;
  reduce_argmin(pmin, pmini, G, sh, shi, a, i);
  if (threadIdx.x == 0) logic_alpha(ds, a, i);
}

// First call of lnsrlb for an iteration (lbfgsbcpu.cpp:1263-1339).
__global__ void __launch_bounds__(kBlock)
s_ls_init(DevScalars* ds, const double* part, const double* pmin, int G) {
  __shared__ DevScalars sds;
  __shared__ double r[P5_COUNT];
  __shared__ double sh[kWavesPerBlock];
  if (ds->halt) return;
  reduce_rows(part, G, 0, P5_COUNT, r);
  double smin;
  reduce_min(pmin, G, sh, smin);
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  if (threadIdx.x == 0) logic_ls_init(s, r[P5_DTD], r[P5_GD], smin);
  ds_store(ds, s);
}

// Later calls of lnsrlb, after an evaluation with value f
// (lbfgsbcpu.cpp:1312-1339 and the loop exit at :352).
__global__ void __launch_bounds__(kBlock)
s_ls_step(DevScalars* ds, const double* part, int G, double f) {
  __shared__ DevScalars sds;
  __shared__ double r[1];
  reduce_rows(part, G, 0, 1, r);
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  if (threadIdx.x == 0) logic_ls_step(s, f, r[0]);
  ds_store(ds, s);
}

// ---------------------------------------------------------------------------
void launch_rowsum(hipStream_t s, const double* part, int G, int v0, int nrows, DevScalars* ds,
                   int p4_rows) {
  hipLaunchKernelGGL(k_rowsum, dim3(nrows), dim3(kBlock), 0, s, part, G, v0, ds, p4_rows);
}
void launch_resume(hipStream_t s, DevScalars* ds, int reset_bp) {
  hipLaunchKernelGGL(s_resume, dim3(1), dim3(64), 0, s, ds, reset_bp);
}
void launch_setup_errclb(hipStream_t s, DevScalars* ds, const double* part, int G) {
  hipLaunchKernelGGL(s_setup_errclb, dim3(1), dim3(kBlock), 0, s, ds, part, G);
}
void launch_setup_active(hipStream_t s, DevScalars* ds, const double* part, int G) {
  hipLaunchKernelGGL(s_setup_active, dim3(1), dim3(kBlock), 0, s, ds, part, G);
}
void launch_setup_projgr(hipStream_t s, DevScalars* ds, const double* pmax, int G, double f) {
  hipLaunchKernelGGL(s_setup_projgr, dim3(1), dim3(kBlock), 0, s, ds, pmax, G, f);
}
void launch_stop_tests(hipStream_t s, DevScalars* ds, const double* part, const double* pmax,
                       const double* pmin, const long long* pmini, int G) {
  hipLaunchKernelGGL(s_stop_tests, dim3(1), dim3(kBlock), 0, s, ds, part, pmax, pmin, pmini, G);
}
void launch_update_gcp(hipStream_t s, DevScalars* ds, const double* part, const double* pmin,
                       const long long* pmini, int G, int reduce) {
  hipLaunchKernelGGL(s_update_gcp, dim3(1), dim3(kBlock), 0, s, ds, part, pmin, pmini, G,
                     reduce);
}
void launch_formk_subsm(hipStream_t s, DevScalars* ds, const double* part, int G) {
  hipLaunchKernelGGL(s_formk_subsm, dim3(1), dim3(kBlock), 0, s, ds, part, G);
}
void launch_alpha(hipStream_t s, DevScalars* ds, const double* pmin, const long long* pmini,
                  int G) {
  hipLaunchKernelGGL(s_alpha, dim3(1), dim3(kBlock), 0, s, ds, pmin, pmini, G);
}
void launch_ls_init(hipStream_t s, DevScalars* ds, const double* part, const double* pmin,
                    int G) {
  hipLaunchKernelGGL(s_ls_init, dim3(1), dim3(kBlock), 0, s, ds, part, pmin, G);
}
void launch_ls_step(hipStream_t s, DevScalars* ds, const double* part, int G, double f) {
  hipLaunchKernelGGL(s_ls_step, dim3(1), dim3(kBlock), 0, s, ds, part, G, f);
}

#undef A1

}  // namespace lbfgsb
