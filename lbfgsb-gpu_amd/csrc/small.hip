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
#include "host_mailbox.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

#define A1(a, ld, i, j) (a)[((i)-1) * (ld) + ((j)-1)]


// The bulk passes leave at most kMaxBlocks partials per value.  The
// finishing reductions load every partial of a thread before the first
// operation on them, so a thread waits on its loads once instead of once per
// loop trip, and then combine them in the order of the plain loop.
constexpr int kPartPerLane = kMaxBlocks / kWave;     // one wave per row
constexpr int kPartPerThread = kMaxBlocks / kBlock;  // the whole block

// Compensated reduction of value rows v0..v0+nv (each G <= kMaxBlocks
// partials long) into out[0..nv): one wave per row, four independent
// accumulators per lane (partials lane, lane + 64, lane + 128, lane + 192,
// then the next 256; the tail into the first), two-sum combination across
// accumulators and lanes.
__device__ void reduce_rows(const double* part, int G, int v0, int nv, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int v = w; v < nv; v += kWavesPerBlock) {
    const double* row = part + (size_t)(v0 + v) * G;
    double pv[kPartPerLane];
#pragma unroll
    for (int k = 0; k < kPartPerLane; k++) {
      const int b = lane + k * kWave;
      pv[k] = b < G ? row[b] : 0.0;
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    bool tail = false;
#pragma unroll
    for (int q = 0; q < kPartPerLane / 4; q++) {
      const int b = lane + q * 4 * kWave;
      if (tail) continue;
      if (b + 3 * kWave < G) {
        a0 += pv[4 * q];
        a1 += pv[4 * q + 1];
        a2 += pv[4 * q + 2];
        a3 += pv[4 * q + 3];
      } else {
        tail = true;  // b + 4 * kWave > G: nothing after this group
#pragma unroll
        for (int r = 0; r < 3; r++)
          if (b + r * kWave < G) a0 += pv[4 * q + r];
      }
    }
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

// The partials of this thread (b = threadIdx.x + k * kBlock < G), padded.
__device__ __forceinline__ void load_partials(const double* p, int G, double pad,
                                              double (&v)[kPartPerThread]) {
#pragma unroll
  for (int k = 0; k < kPartPerThread; k++) {
    const int b = threadIdx.x + k * kBlock;
    v[k] = b < G ? p[b] : pad;
  }
}

__device__ double reduce_max(const double* pm, int G, double* sh) {
  double pv[kPartPerThread];
  load_partials(pm, G, -__builtin_inf(), pv);
  double m = -__builtin_inf();
#pragma unroll
  for (int k = 0; k < kPartPerThread; k++) m = fmax(m, pv[k]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < kWavesPerBlock; w++) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}

// Smallest (value, index) pair, index breaking ties.
__device__ void reduce_argmin(const double* pv, const long long* pi, int G, double* shv,
                              long long* shi, double& v, long long& idx) {
  double xv[kPartPerThread];
  long long xi[kPartPerThread];
  load_partials(pv, G, __builtin_inf(), xv);
#pragma unroll
  for (int k = 0; k < kPartPerThread; k++) {
    const int b = threadIdx.x + k * kBlock;
    xi[k] = b < G ? pi[b] : 0x7fffffffffffffffLL;
  }
  double bv = __builtin_inf();
  long long bi = 0x7fffffffffffffffLL;
#pragma unroll
  for (int k = 0; k < kPartPerThread; k++) {
    if (xv[k] < bv || (xv[k] == bv && xi[k] < bi)) {
      bv = xv[k];
      bi = xi[k];
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
  double xv[kPartPerThread];
  load_partials(pv, G, __builtin_inf(), xv);
  double m = __builtin_inf();
#pragma unroll
  for (int k = 0; k < kPartPerThread; k++) m = fmin(m, xv[k]);
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
// spec: the rows of a speculative pass A, summed on a halted chain.  With
// pmax, one block more finishes pass A's projected-gradient norm and smallest
// breakpoint into ds->red_sbg, red_bk, red_ibk.
__global__ void __launch_bounds__(kBlock)
k_rowsum(const double* part, int G, int v0, DevScalars* ds, int p4_rows, int spec,
         const double* pmax, const double* pmin, const long long* pmini) {
  __shared__ double ss[kWavesPerBlock], sc[kWavesPerBlock];
  __shared__ long long si[kWavesPerBlock];
  if (ds->halt && !spec) return;
  if (pmax && blockIdx.x == gridDim.x - 1) {
    const double sbg = reduce_max(pmax, G, ss);
    double bk;
    long long ibk;
    reduce_argmin(pmin, pmini, G, ss, si, bk, ibk);
    if (threadIdx.x == 0) {
      ds->red_sbg = sbg;
      ds->red_bk = bk;
      ds->red_ibk = ibk;
    }
    return;
  }
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
  // partials t, t + 256 into a0, a1, then the next 512; the tail into a0
  double pv[kPartPerThread];
  load_partials(row, G, 0.0, pv);
  double a0 = 0.0, a1 = 0.0;
  bool tail = false;
#pragma unroll
  for (int q = 0; q < kPartPerThread / 2; q++) {
    const int b = threadIdx.x + q * 2 * kBlock;
    if (tail) continue;
    if (b + kBlock < G) {
      a0 += pv[2 * q];
      a1 += pv[2 * q + 1];
    } else {
      tail = true;  // b + 2 * kBlock > G: nothing after this pair
      if (b < G) a0 += pv[2 * q];
    }
  }
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
  if (ds->halt) {
    if (threadIdx.x == 0) ds->write_trial = 0;
    return;
  }
  (void)part;
  (void)pmax;
  (void)pmin;
  (void)pmini;
  (void)G;
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  // s->red holds the row sums of the pass-A partials, red_sbg / red_bk /
  // red_ibk the norm and smallest breakpoint (k_rowsum)
  if (threadIdx.x == 0) logic_stop_tests(s, s->red, s->red_sbg, s->red_bk, s->red_ibk);
  __syncthreads();
  post_update_middle(s);
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
  post_update_middle(s);
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

// The evaluated value: f, or when fdev is given the objective's FP64 value
// there, rounded to fp32 for an fp32 solve as a callback returns it.
__device__ __forceinline__ double eval_value(double f, const double* fdev, int single) {
  if (!fdev) return f;
  return single ? (double)(float)*fdev : *fdev;
}

// Later calls of lnsrlb, after an evaluation (lbfgsbcpu.cpp:1312-1339 and
// the loop exit at :352).
__global__ void __launch_bounds__(kBlock)
s_ls_step(DevScalars* ds, const double* part, int G, double f, const double* fdev,
          int single) {
  __shared__ DevScalars sds;
  __shared__ double r[1];
  reduce_rows(part, G, 0, 1, r);
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  if (threadIdx.x == 0) logic_ls_step(s, eval_value(f, fdev, single), r[0]);
  ds_store(ds, s);
}

// Line-search step after a speculative pass A on the evaluated trial point
// (k_pass_a with spec), with g'd from that pass (ds->red[PA_GD], summed by
// k_rowsum).  On acceptance with the memory update pass A assumed (updatd,
// the evaluated point itself, no custom trial write), the spare ring slot
// holding the new pair takes the newest logical slot and the stopping tests
// follow as in s_stop_tests; on any other acceptance the status is ST_REDO
// and the host runs the regular tail pass.
__global__ void __launch_bounds__(kBlock)
s_ls_stop(DevScalars* ds, const double* pmax, const double* pmin, const long long* pmini,
          int G, double f, const double* fdev, int single) {
  __shared__ DevScalars sds;
  (void)pmax;
  (void)pmin;
  (void)pmini;
  (void)G;
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  if (threadIdx.x == 0) {
    const int iter0 = s->iter;
    const double stp0 = s->stp;
    logic_ls_step(s, eval_value(f, fdev, single), s->red[PA_GD]);
    if (s->iter != iter0 && s->status == ST_RUNNING) {
      if (s->updatd && !s->write_trial && s->stp == stp0) {
        ring_take_spare(s, s->itail);
        logic_stop_tests(s, s->red, s->red_sbg, s->red_bk, s->red_ibk);
      } else {
        s->status = ST_REDO;
        s->halt = 1;
      }
    }
  }
  __syncthreads();
  post_update_middle(s);
  ds_store(ds, s);
}

// The device state into the host's pinned copy, then the sequence number
// (the end of every chain the host waits on).
__global__ void __launch_bounds__(kBlock)
s_publish(const DevScalars* ds, DevScalars* host, MailSeq* m, unsigned seq) {
  const unsigned long long* src = (const unsigned long long*)ds;
  unsigned long long* dst = (unsigned long long*)host;
  for (int i = threadIdx.x; i < kDsWords; i += blockDim.x) dst[i] = src[i];
  mail_publish(m, seq);
}

// ---------------------------------------------------------------------------
void launch_publish(hipStream_t s, const DevScalars* ds, DevScalars* host, MailSeq* m,
                    unsigned seq) {
  hipLaunchKernelGGL(s_publish, dim3(1), dim3(kBlock), 0, s, ds, host, m, seq);
}
void launch_rowsum(hipStream_t s, const double* part, int G, int v0, int nrows, DevScalars* ds,
                   int p4_rows, int spec, const double* pmax, const double* pmin,
                   const long long* pmini) {
  hipLaunchKernelGGL(k_rowsum, dim3(nrows + (pmax ? 1 : 0)), dim3(kBlock), 0, s, part, G, v0,
                     ds, p4_rows, spec, pmax, pmin, pmini);
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
void launch_ls_step(hipStream_t s, DevScalars* ds, const double* part, int G, double f,
                    const double* fdev, int single) {
  hipLaunchKernelGGL(s_ls_step, dim3(1), dim3(kBlock), 0, s, ds, part, G, f, fdev, single);
}
void launch_ls_stop(hipStream_t s, DevScalars* ds, const double* pmax, const double* pmin,
                    const long long* pmini, int G, double f, const double* fdev, int single) {
  hipLaunchKernelGGL(s_ls_stop, dim3(1), dim3(kBlock), 0, s, ds, pmax, pmin, pmini, G, f,
                     fdev, single);
}

#undef A1

}  // namespace lbfgsb
