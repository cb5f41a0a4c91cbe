// Bulk n-length passes of one L-BFGS-B iteration (HBM-streaming kernels).
//
// Pass map (one launch each, every launch reads each n-vector at most once):
//   k_active    projection of x0 and iwhere set-up     (lbfgsbcpu.cpp:467-515)
//   k_errclb    input validation counts                (lbfgsbcpu.cpp:871-905)
//   k_projgr    projected-gradient norm                (lbfgsbcpu.cpp:1389-1414)
//   k_pass_a    [iteration tail: projgr, ||dx||, y=g-g_old, y'y, ring write,
//                new Gram row]  + Cauchy classification, f1, p = W'd and the
//                breakpoint list                        (lbfgsbcpu.cpp:394-449,
//                                                        567-701, 1341-1387);
//                speculative form on the evaluated trial point, with g'd
//                                                      (lbfgsbcpu.cpp:1312)
//   k_p4        Cauchy point, free/active counts, new formk column, signed
//               entering/leaving Gram, reduced gradient r and W'r
//                                                      (lbfgsbcpu.cpp:705-817,
//                                                        819-869, 907-1054,
//                                                        1135-1173, 1446-1458)
//   k_p5a       subspace direction and step-to-box ratio (lbfgsbcpu.cpp:1472-1516)
//   k_p5b       backtrack, d = z - x, d'd, g'd, stpmx     (lbfgsbcpu.cpp:1517-1533,
//                                                        343-345, 1263-1299)
//   k_trial     line-search trial point                  (lbfgsbcpu.cpp:1329-1335)
//   k_gd        g'd after an evaluation in the regular program (chains with
//               an after-iteration callback, partial runs) (lbfgsbcpu.cpp:1312)
//
// Every kernel is a grid-stride loop over 256-thread blocks; per-block
// partial sums go to part[v * G + block] and are finished by the
// single-workgroup FP64 kernels in small.hip.  Products of the bulk type are
// formed and accumulated in FP64 (an FP32 x FP32 product is exact in FP64),
// and the cross-block sums are compensated (two-sum), which is the
// compensated-dot scheme used for the FP32 bulk path.
#include "common.hpp"
#include "launch.hpp"
#include "pass_device.hpp"
#include "pass_ops.hpp"
#include "state.hpp"

namespace lbfgsb {

__device__ inline int grid_blocks() { return gridDim.x; }

// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_errclb(long long n, const int* __restrict__ nbd, const T* __restrict__ l,
         const T* __restrict__ u, double* part) {
  __shared__ double sh[kWavesPerBlock * 2];
  double v[2] = {0.0, 0.0};
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    int nb = nbd[i];
    if (nb < 0 || nb > 3) v[0] += 1.0;
    if (nb == 2 && l[i] > u[i]) v[1] += 1.0;
  }
  block_sum_store<2>(v, sh, part + blockIdx.x, gridDim.x);
}

// Projection of x onto the box and iwhere initialisation
// (lbfgsbcpu.cpp:467-515).  Partials: [projected, constrained, not-boxed].
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_active(long long n, T* __restrict__ x, const T* __restrict__ l,
         const T* __restrict__ u, const int* __restrict__ nbd,
         uint8_t* __restrict__ iwhere, double* part) {
  __shared__ double sh[kWavesPerBlock * 3];
  double v[3] = {0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int nb = nbd[i];
    T xi = x[i];
    const T li = l[i], ui = u[i];
    if (nb > 0) {
      if (nb <= 2 && xi <= li) {
        if (xi < li) {
          v[0] += 1.0;
          x[i] = li;
        }
      } else if (nb >= 2 && xi >= ui) {
        if (xi > ui) {
          v[0] += 1.0;
          x[i] = ui;
        }
      }
    }
    int iw;
    if (nb != 2) v[2] += 1.0;
    if (nb == 0) {
      iw = -1;
    } else {
      v[1] += 1.0;
      iw = (nb == 2 && ui - li <= 0) ? 3 : 0;
    }
    // every variable starts in the free set of the (empty) previous
    // partition; the byte carries nbd from here on (validated by k_errclb)
    iwhere[i] = iw_make(iw, true, true, nb & 3);
  }
  block_sum_store<3>(v, sh, part + blockIdx.x, gridDim.x);
}

__device__ inline void block_max_store(double v, double* out) {
  __shared__ double shm[kWavesPerBlock];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = shm[0];
    for (int w = 1; w < kWavesPerBlock; w++) m = fmax(m, shm[w]);
    *out = m;
  }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_projgr(long long n, const T* __restrict__ x, const T* __restrict__ g,
         const T* __restrict__ l, const T* __restrict__ u,
         const int* __restrict__ nbd, double* pmax) {
  double m = 0.0;
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    T gi = projected_gi(x[i], g[i], l[i], u[i], nbd[i]);
    m = fmax(m, (double)fabs(gi));
  }
  block_max_store(m, pmax + blockIdx.x);
}

// ---------------------------------------------------------------------------
// k_pass_a: Cauchy classification (+ the iteration tail when FUSED).
//   FUSED: x = X[cur] (new iterate), xold = X[1-cur], g = G[cur],
//          gold = G[1-cur]; when ds->updatd the new slot s = stp*d,
//          y = g - gold is written at ds->itail and the Gram row
//          s'y_j, s_j's over the retained slots is formed.
//   spec (FUSED): the same on the evaluated trial point before the line
//          search has accepted it, on the ring an accepted step with an
//          update leaves, with the new pair in the spare physical slot and
//          g'd summed as well; it runs whatever ds->halt says (the chain is
//          halted after an evaluation request).
template <typename T, bool FUSED>
__global__ void __launch_bounds__(kBlock)
k_pass_a(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
         const T* __restrict__ u, uint8_t* __restrict__ iwhere, T* __restrict__ W, const T* __restrict__ d,
         Cand* __restrict__ nearc, unsigned* __restrict__ nearn, int near_cap, DevScalars* ds,
         double* part, double* pmax, double* pmin, long long* pmini, int rev, int spec) {
  __shared__ double sh[kWavesPerBlock * PA_COUNT];
  __shared__ unsigned s_near;
  if (ds->halt && !spec) return;
  if (threadIdx.x == 0) s_near = 0u;
  __syncthreads();
  using Op = PassAOp<T, FUSED>;
  Op op;
  op.init(ds, X, Gv, l, u, iwhere, W, d, nearc + (size_t)blockIdx.x * near_cap, &s_near,
          (unsigned)near_cap, spec != 0);
  if constexpr (sizeof(T) == 4) {
    // FP32: two consecutive runs of 256 elements per block step, software
    // pipelined: the loads of one run are in flight while the other run is
    // processed
    const long long nc = (n + 2 * kBlock - 1) / (2 * kBlock);
    long long c = blockIdx.x;
    typename Op::In ea, eb;
    if (c < nc) {
      const long long ia = chunk_start(c, nc, 2 * kBlock, rev) + threadIdx.x;
      op.load(ea, ia < n ? ia : n - 1);
    }
    for (; c < nc; c += gridDim.x) {
      const long long i0 = chunk_start(c, nc, 2 * kBlock, rev);
      const long long ia = i0 + threadIdx.x, ib = ia + kBlock;
      const long long cn = c + gridDim.x;
      const long long na = (cn < nc ? chunk_start(cn, nc, 2 * kBlock, rev) : i0) + threadIdx.x;
      op.load(eb, ib < n ? ib : n - 1);
      op.keep(ea);
      if (ia < n) op.element(ia, ea);
      // unconditional (the next chunk, or this one again after the last):
      // a branch around it would make the wait for eb's loads cover these too
      op.load(ea, na < n ? na : n - 1);
      op.keep(eb);
      if (ib < n) op.element(ib, eb);
    }
  } else {
    const long long nc = (n + kBlock - 1) / kBlock;
    for (long long c = blockIdx.x; c < nc; c += gridDim.x) {
      const long long i = chunk_start(c, nc, kBlock, rev) + threadIdx.x;
      if (i >= n) continue;
      typename Op::In e;
      op.load(e, i);
      op.keep(e);
      op.element(i, e);
    }
  }
  op.finish();
  block_sum_store<PA_COUNT>(op.acc, sh, part + blockIdx.x, gridDim.x);
  // the barriers of block_sum_store order every list append before this read
  if (threadIdx.x == 0) nearn[blockIdx.x] = s_near;
  if (FUSED) block_max_store(op.sbg, pmax + blockIdx.x);
  double bkmin = op.bkmin;
  long long ibk = op.ibk;
  wave_argmin(bkmin, ibk);
  __shared__ double shv[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if ((threadIdx.x & 63) == 0) {
    shv[threadIdx.x >> 6] = bkmin;
    shi[threadIdx.x >> 6] = ibk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv = shv[0];
    long long bi = shi[0];
    for (int w = 1; w < kWavesPerBlock; w++) {
      if (shv[w] < bv || (shv[w] == bv && shi[w] < bi)) {
        bv = shv[w];
        bi = shi[w];
      }
    }
    pmin[blockIdx.x] = bv;
    pmini[blockIdx.x] = bi;
  }
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_p4(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
     const T* __restrict__ u, uint8_t* __restrict__ iwhere, const T* __restrict__ W, DevScalars* ds,
     double* part, int rev) {
  __shared__ double sh[kWavesPerBlock * P4_CORR];
  __shared__ double s_gram[256];
  __shared__ double s_row[kWavesPerBlock][kRowBuf];
  if (ds->halt) return;
  double* rowbuf = s_row[threadIdx.x >> 6];
  P4Op<T> op;
  op.init(ds, X, Gv, l, u, iwhere, W);
  if constexpr (sizeof(T) == 4) {
    // FP32: each block takes two consecutive runs of 256 elements per step,
    // software pipelined: one run's loads are in flight while the other run
    // is processed (the register footprint leaves two waves per SIMD, so a
    // wave cannot rely on others to cover its arithmetic)
    const long long nc = (n + 2 * kBlock - 1) / (2 * kBlock);
    long long c = blockIdx.x;
    typename P4Op<T>::In ea, eb;
    if (c < nc) {
      const long long ia = chunk_start(c, nc, 2 * kBlock, rev) + threadIdx.x;
      op.load(ea, ia < n ? ia : n - 1);
    }
    for (; c < nc; c += gridDim.x) {
      const long long i0 = chunk_start(c, nc, 2 * kBlock, rev);
      const long long ia = i0 + threadIdx.x, ib = ia + kBlock;
      const long long cn = c + gridDim.x;
      const long long na = (cn < nc ? chunk_start(cn, nc, 2 * kBlock, rev) : i0) + threadIdx.x;
      op.load(eb, ib < n ? ib : n - 1);
      op.keep(ea);
      op.corr_step(ia, ia < n ? op.element(ia, ea) : 0, ea, rowbuf);
      op.load(ea, na < n ? na : n - 1);  // unconditional, as in k_pass_a
      op.keep(eb);
      op.corr_step(ib, ib < n ? op.element(ib, eb) : 0, eb, rowbuf);
    }
  } else {
    // FP64: one element per step (two would not fit two waves per SIMD)
    const long long nc = (n + kBlock - 1) / kBlock;
    for (long long c = blockIdx.x; c < nc; c += gridDim.x) {
      const long long ia = chunk_start(c, nc, kBlock, rev) + threadIdx.x;
      typename P4Op<T>::In ea;
      op.load(ea, ia < n ? ia : n - 1);
      op.keep(ea);
      op.corr_step(ia, ia < n ? op.element(ia, ea) : 0, ea, rowbuf);
    }
  }
  block_sum_store<P4_CORR>(op.acc, sh, part + blockIdx.x, gridDim.x);
  if (op.track && op.col > 0) {
    op.corr_block_sum(s_gram, kWavesPerBlock);
    const int upcl = ds->updatd ? op.col - 1 : op.col;
    for (int t = threadIdx.x; t < 136; t += blockDim.x)
      part[(size_t)(P4_CORR + t) * gridDim.x + blockIdx.x] = corr_pair_value(t, s_gram, upcl);
  }
}

// x_cp of every variable from the current x, g, class and search outcome:
// what k_p4, k_p5a and k_p5b compute per element (an inspection aid; the
// iteration stores no x_cp).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cauchy_point(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
               const T* __restrict__ u, const uint8_t* __restrict__ iwhere, const DevScalars* ds, T* __restrict__ xcp) {
  CauchyOutcome<T> co;
  co.init(ds);
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    int iw = iw_get(iwhere[i]);
    xcp[i] = co.point(i, x[i], g[i], l[i], u[i], iw_nbd(iwhere[i]), iw);
  }
}

// k_p5a: subspace Newton direction over the free variables and the
// step-to-box ratio (lbfgsbcpu.cpp:1472-1516).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_p5a(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l, const T* __restrict__ u,
      const uint8_t* __restrict__ iwhere,
      const T* __restrict__ W, T* __restrict__ dsub, DevScalars* ds, double* pmin,
      long long* pmini, int rev) {
  if (ds->halt || ds->skip_subsm) return;
  P5aOp<T> op;
  op.init(ds, X, Gv, l, u, iwhere, W, dsub);
  const long long nc = (n + kBlock - 1) / kBlock;
  for (long long c = blockIdx.x; c < nc; c += gridDim.x) {
    const long long i = chunk_start(c, nc, kBlock, rev) + threadIdx.x;
    if (i >= n) continue;
    typename P5aOp<T>::In e;
    op.load(e, i);
    op.keep(e);
    op.element(i, e);
  }
  double amin = op.amin;
  long long ai = op.ai;
  wave_argmin(amin, ai);
  __shared__ double shv[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if ((threadIdx.x & 63) == 0) {
    shv[threadIdx.x >> 6] = amin;
    shi[threadIdx.x >> 6] = ai;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv = shv[0];
    long long bi = shi[0];
    for (int w = 1; w < kWavesPerBlock; w++)
      if (shv[w] < bv || (shv[w] == bv && shi[w] < bi)) {
        bv = shv[w];
        bi = shi[w];
      }
    pmin[blockIdx.x] = bv;
    pmini[blockIdx.x] = bi;
  }
}

// k_p5b: apply the backtrack (pinning the limiting variable), z into the
// trial buffer, and the line-search set-up sums d'd, g'd (d = z - x) and the
// nbd-aware max step.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_p5b(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
      const T* __restrict__ u, const uint8_t* __restrict__ iwhere, const T* __restrict__ dsub, DevScalars* ds,
      double* part, double* pmin, int rev) {
  __shared__ double sh[kWavesPerBlock * P5_COUNT];
  if (ds->halt) return;
  P5bOp<T> op;
  op.init(ds, X, Gv, l, u, iwhere, dsub);
  const long long nc = (n + kBlock - 1) / kBlock;
  for (long long c = blockIdx.x; c < nc; c += gridDim.x) {
    const long long i = chunk_start(c, nc, kBlock, rev) + threadIdx.x;
    if (i >= n) continue;
    typename P5bOp<T>::In e;
    op.load(e, i);
    op.keep(e);
    op.element(i, e);
  }
  block_sum_store<P5_COUNT>(op.acc, sh, part + blockIdx.x, gridDim.x);
  double smin = -wave_max(-op.smin);
  __shared__ double shm[kWavesPerBlock];
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = smin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = shm[0];
    for (int w = 1; w < kWavesPerBlock; w++) v = fmin(v, shm[w]);
    pmin[blockIdx.x] = v;
  }
}

// The subspace point z of every variable, as k_p5b forms it, into out.
template <typename T>
__device__ __forceinline__ void subspace_points(long long n, Pair<T> X, Pair<T> Gv,
                                                const T* __restrict__ l,
                                                const T* __restrict__ u,
                                                const uint8_t* __restrict__ iwhere,
                                                const T* __restrict__ dsub,
                                                const DevScalars* ds, T* __restrict__ out) {
  P5bOp<T> op;
  op.init(ds, X, Gv, l, u, iwhere, dsub);
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    typename P5bOp<T>::In e;
    op.load(e, i);
    out[i] = op.subspace_point(i, e);
  }
}

// Trial point x_b = stp*d + x_a, or the subspace point z when stp == 1
// (lbfgsbcpu.cpp:1329-1335).  k_p5b leaves z in the trial buffer, so a unit
// step only gets here (write_trial) when a trial since has overwritten it;
// z is then formed again from the state it came from.  The first trial off
// z stores d = z - x (trial_element).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_trial(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l, const T* __restrict__ u,
        const uint8_t* __restrict__ iwhere,
        const T* __restrict__ dsub, T* __restrict__ d, const DevScalars* ds) {
  if (!ds->write_trial) return;
  const T* __restrict__ xa = X[ds->trial_src];
  T* __restrict__ xb = X[ds->trial_dst];
  const double stp = ds->stp;
  const bool trial_d = ds->trial_d;
  P5bOp<T> sub;
  sub.init(ds, X, Gv, l, u, iwhere, dsub, ds->trial_src);
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    trial_element(i, ds, sub, xa, xb, d, stp, trial_d);
}

// z of every variable into out (an inspection aid: the iteration keeps z
// only in the trial buffer).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_subspace_point(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
                 const T* __restrict__ u, const uint8_t* __restrict__ iwhere, const T* __restrict__ dsub,
                 const DevScalars* ds, T* __restrict__ out) {
  subspace_points(n, X, Gv, l, u, iwhere, dsub, ds, out);
}

// g_b . d after an evaluation (lbfgsbcpu.cpp:1312).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gd(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ d, const DevScalars* ds,
     double* part, int rev) {
  __shared__ double sh[kWavesPerBlock];
  const int cur = ds->phase & 1;
  const T* __restrict__ gb = Gv[cur ^ 1];
  LsDirection<T> dir;
  dir.init(ds, X, d);
  double acc[1] = {0.0};
  const long long nc = (n + kBlock - 1) / kBlock;
  for (long long c = blockIdx.x; c < nc; c += gridDim.x) {
    const long long i = chunk_start(c, nc, kBlock, rev) + threadIdx.x;
    if (i < n) acc[0] += (double)gb[i] * (double)dir.at(i);
  }
  block_sum_store<1>(acc, sh, part + blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// host launch wrappers
template <typename T>
void launch_errclb(hipStream_t s, int G, long long n, const int* nbd, const T* l,
                   const T* u, double* part) {
  hipLaunchKernelGGL(k_errclb<T>, dim3(G), dim3(kBlock), 0, s, n, nbd, l, u, part);
}
template <typename T>
void launch_active(hipStream_t s, int G, long long n, T* x, const T* l, const T* u,
                   const int* nbd, uint8_t* iwhere, double* part) {
  hipLaunchKernelGGL(k_active<T>, dim3(G), dim3(kBlock), 0, s, n, x, l, u, nbd,
                     iwhere, part);
}
template <typename T>
void launch_projgr(hipStream_t s, int G, long long n, const T* x, const T* g,
                   const T* l, const T* u, const int* nbd, double* pmax) {
  hipLaunchKernelGGL(k_projgr<T>, dim3(G), dim3(kBlock), 0, s, n, x, g, l, u, nbd,
                     pmax);
}
template <typename T>
void launch_pass_a(hipStream_t s, int G, bool fused, long long n, Pair<T> X,
                   Pair<T> Gv, const T* l, const T* u, uint8_t* iwhere, T* W, const T* d,
                   Cand* nearc, unsigned* nearn, int near_cap, DevScalars* ds,
                   double* part, double* pmax, double* pmin, long long* pmini, int rev,
                   int spec) {
  if (fused)
    hipLaunchKernelGGL((k_pass_a<T, true>), dim3(G), dim3(kBlock), 0, s, n, X, Gv,
                       l, u, iwhere, W, d, nearc, nearn, near_cap, ds, part, pmax, pmin, pmini,
                       rev, spec);
  else
    hipLaunchKernelGGL((k_pass_a<T, false>), dim3(G), dim3(kBlock), 0, s, n, X,
                       Gv, l, u, iwhere, W, d, nearc, nearn, near_cap, ds, part, pmax, pmin,
                       pmini, rev, 0);
}
template <typename T>
void launch_p4(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
               const T* l, const T* u, uint8_t* iwhere,
               const T* W, DevScalars* ds, double* part, int rev) {
  hipLaunchKernelGGL(k_p4<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere, W, ds, part, rev);
}
template <typename T>
void launch_p5a(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                const T* u, const uint8_t* iwhere, const T* W, T* dsub,
                DevScalars* ds, double* pmin, long long* pmini, int rev) {
  hipLaunchKernelGGL(k_p5a<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere,
                     W, dsub, ds, pmin, pmini, rev);
}
template <typename T>
void launch_cauchy_point(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                         const T* u, const uint8_t* iwhere,
                         const DevScalars* ds, T* xcp) {
  hipLaunchKernelGGL(k_cauchy_point<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere,
                     ds, xcp);
}
template <typename T>
void launch_p5b(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
                const T* l, const T* u, const uint8_t* iwhere,
                const T* dsub, DevScalars* ds, double* part, double* pmin, int rev) {
  hipLaunchKernelGGL(k_p5b<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere, dsub, ds, part, pmin, rev);
}
template <typename T>
void launch_trial(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                  const T* u, const uint8_t* iwhere, const T* dsub,
                  T* d, const DevScalars* ds) {
  hipLaunchKernelGGL(k_trial<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere,
                     dsub, d, ds);
}
template <typename T>
void launch_subspace_point(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
                           const T* l, const T* u, const uint8_t* iwhere,
                           const T* dsub, const DevScalars* ds, T* out) {
  hipLaunchKernelGGL(k_subspace_point<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere, dsub, ds, out);
}
template <typename T>
void launch_gd(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* d,
               const DevScalars* ds, double* part, int rev) {
  hipLaunchKernelGGL(k_gd<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, d, ds, part, rev);
}
// d = z - x of every variable during the line search into out (inspection).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_direction(long long n, Pair<T> X, const T* __restrict__ d, const DevScalars* ds,
            T* __restrict__ out) {
  LsDirection<T> dir;
  dir.init(ds, X, d);
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    out[i] = dir.at(i);
}
template <typename T>
void launch_direction(hipStream_t s, int G, long long n, Pair<T> X, const T* d,
                      const DevScalars* ds, T* out) {
  hipLaunchKernelGGL(k_direction<T>, dim3(G), dim3(kBlock), 0, s, n, X, d, ds, out);
}

#define LBFGSB_INST_BULK(T)                                                         \
  template void launch_errclb<T>(hipStream_t, int, long long, const int*, const T*, \
                                 const T*, double*);                                \
  template void launch_active<T>(hipStream_t, int, long long, T*, const T*,         \
                                 const T*, const int*, uint8_t*, double*);          \
  template void launch_projgr<T>(hipStream_t, int, long long, const T*, const T*,   \
                                 const T*, const T*, const int*, double*);          \
  template void launch_pass_a<T>(hipStream_t, int, bool, long long, Pair<T>,      \
                                 Pair<T>, const T*, const T*,         \
                                 uint8_t*, T*, const T*, Cand*, unsigned*, int, DevScalars*, \
                                 double*, double*, double*, long long*, int, int);       \
  template void launch_p4<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,     \
                             const T*, const T*, uint8_t*, const T*,    \
                             DevScalars*, double*, int);                                 \
  template void launch_p5a<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,      \
                              const T*, const T*, const uint8_t*,       \
                              const T*, T*, DevScalars*, double*, long long*, int);      \
  template void launch_cauchy_point<T>(hipStream_t, int, long long, Pair<T>,      \
                                       Pair<T>, const T*, const T*,    \
                                       const uint8_t*, const DevScalars*, T*);      \
  template void launch_p5b<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,    \
                              const T*, const T*, const uint8_t*,       \
                              const T*, DevScalars*, double*, double*, int);             \
  template void launch_trial<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,    \
                                const T*, const T*, const uint8_t*,     \
                                const T*, T*, const DevScalars*);                   \
  template void launch_subspace_point<T>(hipStream_t, int, long long, Pair<T>,      \
                                         Pair<T>, const T*, const T*,  \
                                         const uint8_t*, const T*,                  \
                                         const DevScalars*, T*);                    \
  template void launch_gd<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,      \
                             const T*, const DevScalars*, double*, int);                 \
  template void launch_direction<T>(hipStream_t, int, long long, Pair<T>, const T*, \
                                    const DevScalars*, T*);

LBFGSB_INST_BULK(float)
LBFGSB_INST_BULK(double)

}  // namespace lbfgsb
