// Batched form: one persistent kernel solves many independent problems.
//
// Every workgroup takes problems from a queue (largest first, as ordered by
// the host) and runs each one to completion -- or for an iteration budget,
// for the serving loop -- with all of its state in HBM slices of
// concatenated arrays and its control state (DevScalars) in LDS.  The
// iteration runs the same step programs as the plain engine (engine.cpp):
// the same pass operators (pass_ops.hpp) over the problem with block-level
// reductions instead of grid-wide ones, the same control logic
// (control.hpp), and the Cauchy rounds of gcp_device.hpp inside the
// workgroup, sorting in LDS (or in HBM for rounds larger than one LDS sort).
// The objective is one of the built-in kinds, evaluated in the workgroup
// with the value kept on the device.  No host round trip happens inside a
// launch.
#include "batch.hpp"
#include "common.hpp"
#include "control.hpp"
#include "gcp_device.hpp"
#include "obj_device.hpp"
#include "pass_device.hpp"
#include "pass_ops.hpp"

namespace lbfgsb {

namespace {

constexpr int kBDsWords = sizeof(DevScalars) / 8;

// LDS of one workgroup.
struct BatchLds {
  DevScalars s;
  double red[P4_COUNT + 1];  // sums of the last pass (PartA / Part4 / Part5 layouts)
  double f;                  // objective value of the last evaluation
  double rv[4];              // [sbgnrm, bkmin, amin, smin] of the last passes
  long long ri[2];           // [ibkmin, ibd]
  union {
    RoundLds round;
    struct {
      double sh[kMaxRoundWaves * P4_CORR];
      double corr[kMaxRoundWaves][136];
      long long shi[kMaxRoundWaves];
    } r;
  } u;
};

template <typename T>
struct Prob {
  long long n;
  Pair<T> X, G;
  const T* l;
  const T* u;
  const int* nbd;
  uint8_t* iwhere;
  T* W;
  T* d;
  T* z;
  T* r;
  T* dsub;
  Cand* ca;
  Cand* cb;
  ObjDesc obj;
};

__device__ __forceinline__ int nwaves() { return blockDim.x >> 6; }

// Sums of NV per-thread values over the workgroup into out[0 .. NV).
template <int NV>
__device__ __forceinline__ void bsum(const double (&v)[NV], double* sh, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = nwaves();
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double s = wave_sum(v[q]);
    if (lane == 0) sh[wid * NV + q] = s;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < NV; q += blockDim.x) {
    double t = 0.0;
    for (int w = 0; w < nw; w++) t += sh[w * NV + q];
    out[q] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ double bmax(double v, double* sh) {
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < nwaves(); w++) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ double bmin(double v, double* sh) { return -bmax(-v, sh); }

__device__ __forceinline__ void bargmin(double& v, long long& i, double* sh, long long* shi) {
  wave_argmin(v, i);
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = v;
    shi[threadIdx.x >> 6] = i;
  }
  __syncthreads();
  v = sh[0];
  i = shi[0];
  for (int w = 1; w < nwaves(); w++)
    if (sh[w] < v || (sh[w] == v && shi[w] < i)) {
      v = sh[w];
      i = shi[w];
    }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// set-up passes (lbfgsbcpu.cpp:258-278)
template <typename T>
__device__ __forceinline__ void b_errclb(const Prob<T>& P, BatchLds& L) {
  double v[2] = {0.0, 0.0};
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x) {
    const int nb = P.nbd[i];
    if (nb < 0 || nb > 3) v[0] += 1.0;
    if (nb == 2 && P.l[i] > P.u[i]) v[1] += 1.0;
  }
  bsum<2>(v, L.u.r.sh, L.red);
  if (threadIdx.x == 0) logic_setup_errclb(&L.s, L.red[0], L.red[1]);
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void b_active(const Prob<T>& P, BatchLds& L) {
  T* x = P.X[0];
  double v[3] = {0.0, 0.0, 0.0};
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x) {
    const int nb = P.nbd[i];
    const T xi = x[i], li = P.l[i], ui = P.u[i];
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
    P.iwhere[i] = iw_make(iw, true, true);
  }
  bsum<3>(v, L.u.r.sh, L.red);
  if (threadIdx.x == 0) {
    const double r[3] = {L.red[0], L.red[1], L.red[2]};
    logic_setup_active(&L.s, r);
  }
  __syncthreads();
}

// Objective at X[which] into G[which]; L.f = f rounded to the bulk type, as
// a callback returning `real& f` would hand it over.
template <typename T>
__device__ __forceinline__ void b_objective(const Prob<T>& P, BatchLds& L, int which) {
  const T* x = P.X[which];
  T* g = P.G[which];
  double acc[1] = {0.0};
  const long long cnt = obj_terms(P.obj, P.n);
  for (long long k = threadIdx.x; k < cnt; k += blockDim.x) acc[0] += obj_term(P.obj, P.n, x, g, k);
  bsum<1>(acc, L.u.r.sh, &L.f);
  if (threadIdx.x == 0) L.f = (double)(T)L.f;
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void b_projgr(const Prob<T>& P, BatchLds& L) {
  const T* x = P.X[0];
  const T* g = P.G[0];
  double m = 0.0;
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x)
    m = fmax(m, (double)fabs(projected_gi(x[i], g[i], P.l[i], P.u[i], P.nbd[i])));
  m = bmax(m, L.u.r.sh);
  if (threadIdx.x == 0) logic_setup_projgr(&L.s, m, L.f);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// iteration passes
template <typename T, bool FUSED>
__device__ __forceinline__ void b_pass_a(const Prob<T>& P, BatchLds& L) {
  if (L.s.halt) return;
  using Op = PassAOp<T, FUSED>;
  Op op;
  op.init(&L.s, P.X, P.G, P.l, P.u, P.nbd, P.iwhere, P.W, P.d);
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x) {
    typename Op::In e;
    op.load(e, i);
    op.keep(e);
    op.element(i, e);
  }
  op.finish();
  bsum<PA_COUNT>(op.acc, L.u.r.sh, L.red);
  const double sbg = bmax(op.sbg, L.u.r.sh);
  double bk = op.bkmin;
  long long ibk = op.ibk;
  bargmin(bk, ibk, L.u.r.sh, L.u.r.shi);
  if (threadIdx.x == 0) {
    L.rv[0] = sbg;
    L.rv[1] = bk;
    L.ri[0] = ibk;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void b_p4(const Prob<T>& P, BatchLds& L) {
  if (L.s.halt) return;
  P4Op<T> op;
  op.init(&L.s, P.X, P.G, P.l, P.u, P.nbd, P.iwhere, P.W, P.z, P.r);
  const long long n = P.n;
  for (long long i0 = 0; i0 < n; i0 += blockDim.x) {
    const long long i = i0 + threadIdx.x;
    typename P4Op<T>::In e;
    op.load(e, i < n ? i : n - 1);
    op.keep(e);
    op.corr_step(i, i < n ? op.element(i, e) : 0);
  }
  bsum<P4_CORR>(op.acc, L.u.r.sh, L.red);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const bool tracked = op.track && op.col > 0;
  if (tracked) {
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (lane + 64 * q < 136) L.u.r.corr[wid][lane + 64 * q] = op.corr[q];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 136; t += blockDim.x) {
    double s = 0.0;
    if (tracked)
      for (int w = 0; w < nwaves(); w++) s += L.u.r.corr[w][t];
    L.red[P4_CORR + t] = s;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void b_p5a(const Prob<T>& P, BatchLds& L) {
  if (L.s.halt || L.s.skip_subsm) return;
  P5aOp<T> op;
  op.init(&L.s, P.l, P.u, P.nbd, P.iwhere, P.W, P.z, P.r, P.dsub);
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x) {
    typename P5aOp<T>::In e;
    op.load(e, i);
    op.keep(e);
    op.element(i, e);
  }
  double a = op.amin;
  long long ai = op.ai;
  bargmin(a, ai, L.u.r.sh, L.u.r.shi);
  if (threadIdx.x == 0) {
    L.rv[2] = a;
    L.ri[1] = ai;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void b_p5b(const Prob<T>& P, BatchLds& L) {
  if (L.s.halt) return;
  P5bOp<T> op;
  op.init(&L.s, P.X, P.G, P.l, P.u, P.nbd, P.iwhere, P.z, P.dsub, P.d);
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x) {
    typename P5bOp<T>::In e;
    op.load(e, i);
    op.keep(e);
    op.element(i, e);
  }
  bsum<P5_COUNT>(op.acc, L.u.r.sh, L.red);
  const double smin = bmin(op.smin, L.u.r.sh);
  if (threadIdx.x == 0) L.rv[3] = smin;
  __syncthreads();
}

// Trial point x_b = stp d + x_a, or the subspace point when stp == 1
// (lbfgsbcpu.cpp:1329-1335).
template <typename T>
__device__ __forceinline__ void b_trial(const Prob<T>& P, BatchLds& L) {
  if (!L.s.write_trial) return;
  const T* xa = P.X[L.s.trial_src];
  T* xb = P.X[L.s.trial_dst];
  const double stp = L.s.stp;
  const bool one = stp == 1.0;
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x)
    xb[i] = one ? P.z[i] : (T)(stp * (double)P.d[i] + (double)xa[i]);
  __syncthreads();
}

// g_b . d after an evaluation (lbfgsbcpu.cpp:1312).
template <typename T>
__device__ __forceinline__ void b_gd(const Prob<T>& P, BatchLds& L) {
  const T* gb = P.G[(L.s.phase & 1) ^ 1];
  double acc[1] = {0.0};
  for (long long i = threadIdx.x; i < P.n; i += blockDim.x)
    acc[0] += (double)gb[i] * (double)P.d[i];
  bsum<1>(acc, L.u.r.sh, L.red);
}

// Breakpoint candidates with t in [tau_lo, tau_hi) into P.ca.
template <typename T>
__device__ __forceinline__ void b_select(const Prob<T>& P, BatchLds& L) {
  DevScalars* s = &L.s;
  const int cur = s->phase & 1;
  const T* __restrict__ x = P.X[cur];
  const T* __restrict__ g = P.G[cur];
  const double lo = s->gcp_tau_lo, hi = s->gcp_tau_hi;
  const int lane = threadIdx.x & 63;
  __syncthreads();
  if (threadIdx.x == 0) s->gcp_need_select = 0;
  for (long long i0 = 0; i0 < P.n; i0 += blockDim.x) {
    const long long i = i0 + threadIdx.x;
    bool take = false;
    T tb = 0;
    if (i < P.n && breakpoint_of(x[i], g[i], P.l[i], P.u[i], P.nbd[i], iw_get(P.iwhere[i]), tb)) {
      const double t = (double)tb;
      take = t >= lo && t < hi;
    }
    const unsigned long long mask = __ballot(take);
    if (mask) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(&s->cand_count, (unsigned long long)__popcll(mask));
      base = __shfl(base, 0, kWave);
      const unsigned long long slot = base + __popcll(mask & ((1ull << lane) - 1ull));
      if (take && slot < (unsigned long long)P.n) {
        Cand c;
        c.t = (double)tb;
        c.idx = (int)i;
        c.pad = 0;
        P.ca[slot] = c;
      }
    }
  }
  __syncthreads();
}

// The breakpoint rounds of the Cauchy search, all inside the workgroup.
template <typename T>
__device__ __forceinline__ void b_gcp(const Prob<T>& P, BatchLds& L) {
  DevScalars* s = &L.s;
  if (!s->gcp_pending || s->halt) return;
  for (int guard = 0; guard < 64 && s->gcp_pending; guard++) {
    b_select(P, L);
    const long long cnt = (long long)s->cand_count;
    const long long K = cnt < P.n ? cnt : P.n;
    stage_gcp_const(s, L.u.round.gc);
    const T* g = P.G[s->phase & 1];
    if (K <= kSmallRound) {
      block_sort_small(P.ca, K, L.u.round.sc);
      block_round_scan(s, L.u.round, L.u.round.sc, K, g, P.W);
    } else {
      const Cand* sorted = block_sort_large(P.ca, P.cb, K, L.u.round.sc);
      block_round_scan(s, L.u.round, sorted, K, g, P.W);
    }
    if (s->gcp_pending) {
      // every candidate of the t-range passed: the next range
      if (threadIdx.x == 0) logic_resume(s, 0);
      __syncthreads();
    }
  }
}

// One step of the chain (the block form of the plain engine's kernels,
// engine.cpp: launch_step).
template <typename T>
__device__ __forceinline__ void b_step(int step, double f, const Prob<T>& P, BatchLds& L) {
  DevScalars* s = &L.s;
  const bool t0 = threadIdx.x == 0;
  switch (step) {
    case K_RESUME_RESET:
      if (t0) logic_resume(s, 1);
      break;
    case K_RESUME:
      if (t0) logic_resume(s, 0);
      break;
    case K_PASSA_SA:
      b_pass_a<T, false>(P, L);
      break;
    case K_PASSA_FUSED:
      b_pass_a<T, true>(P, L);
      break;
    case K_UPD_REDUCE:
      if (s->halt) {
        if (t0) s->write_trial = 0;
        break;
      }
      if (t0) logic_update_gcp(s, 1, L.red, L.rv[1], L.ri[0]);
      __syncthreads();
      build_middle(s);
      break;
    case K_STOP:
      if (s->halt) {
        if (t0) s->write_trial = 0;
        break;
      }
      if (t0) logic_stop_tests(s, L.red, L.rv[0], L.rv[1], L.ri[0]);
      __syncthreads();
      build_middle(s);
      break;
    case K_UPD:
      if (s->halt) {
        if (t0) s->write_trial = 0;
        break;
      }
      if (s->upd_done) break;
      if (t0) logic_update_gcp(s, 0, L.red, 0.0, 0);
      __syncthreads();
      build_middle(s);
      break;
    case K_GCP:
      b_gcp(P, L);
      break;
    case K_P4:
      b_p4(P, L);
      break;
    case K_FORMK:
      if (s->halt) break;
      if (t0) logic_formk_subsm(s, L.red, s->p4_flagged != 0);
      break;
    case K_P5A:
      b_p5a(P, L);
      break;
    case K_ALPHA:
      if (s->halt || s->skip_subsm) break;
      if (t0) logic_alpha(s, L.rv[2], L.ri[1]);
      break;
    case K_P5B:
      b_p5b(P, L);
      break;
    case K_LSINIT:
      if (s->halt) break;
      if (t0) logic_ls_init(s, L.red[P5_DTD], L.red[P5_GD], L.rv[3]);
      break;
    case K_TRIAL:
      b_trial(P, L);
      break;
    case K_GD:
      b_gd(P, L);
      break;
    case K_LSSTEP:
      if (t0) logic_ls_step(s, f, L.red[0]);
      break;
    default:
      break;
  }
  __syncthreads();
}

// The programs of the chain (engine.cpp: program_steps), -1 terminated; the
// loop keeps one copy of every step in the code.
__constant__ int kProgSteps[3][16] = {
    {K_RESUME_RESET, K_PASSA_SA, K_UPD_REDUCE, K_GCP, K_P4, K_FORMK, K_P5A, K_ALPHA, K_P5B,
     K_LSINIT, K_TRIAL, -1, -1, -1, -1, -1},
    {K_GD, K_LSSTEP, K_TRIAL, K_PASSA_FUSED, K_STOP, K_UPD, K_GCP, K_P4, K_FORMK, K_P5A, K_ALPHA,
     K_P5B, K_LSINIT, K_TRIAL, -1, -1},
    {K_RESUME_RESET, K_PASSA_FUSED, K_STOP, K_UPD, K_GCP, K_P4, K_FORMK, K_P5A, K_ALPHA, K_P5B,
     K_LSINIT, K_TRIAL, -1, -1, -1, -1}};

template <typename T>
__device__ __forceinline__ void b_program(int prog, double f, const Prob<T>& P, BatchLds& L) {
  if (prog < PROG_ITER_START || prog > PROG_AFTER_ACCEPT) return;
  for (int k = 0; k < 16; k++) {
    const int step = kProgSteps[prog][k];
    if (step < 0) break;
    b_step(step, f, P, L);
  }
}

// One problem: set-up (first launch only) and the chain loop of the host
// driver (capi.cpp: drive), with the objective evaluated in the workgroup.
template <typename T>
__device__ __forceinline__ void solve_one(const BatchProb& bp, const BatchBufs<T>& B, DevScalars* dsg,
                          BatchResult* res, int iter_budget, BatchLds& L) {
  Prob<T> P;
  P.n = bp.n;
  P.X = Pair<T>{B.x + bp.off, B.xb + bp.off};
  P.G = Pair<T>{B.g + bp.off, B.gb + bp.off};
  P.l = B.l + bp.off;
  P.u = B.u + bp.off;
  P.nbd = B.nbd + bp.off;
  P.iwhere = B.iwhere + bp.off;
  P.W = B.W + bp.woff;
  P.d = B.d + bp.off;
  P.z = B.z + bp.off;
  P.r = B.r + bp.off;
  P.dsub = B.dsub + bp.off;
  P.ca = B.ca + bp.off;
  P.cb = B.cb + bp.off;
  P.obj = bp.obj;

  {
    const unsigned long long* src = (const unsigned long long*)dsg;
    unsigned long long* dst = (unsigned long long*)&L.s;
    for (int i = threadIdx.x; i < kBDsWords; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  DevScalars* s = &L.s;
  if (!s->b_started) {
    if (threadIdx.x == 0) res->clk_start = (long long)wall_clock64();
    b_errclb(P, L);
    if (s->status != ST_DONE) {
      b_active(P, L);
      b_objective(P, L, 0);
      b_projgr(P, L);
    }
    if (threadIdx.x == 0) {
      s->b_started = 1;
      s->b_next_prog = PROG_ITER_START;
      s->b_next_f = 0.0;
    }
    __syncthreads();
  }
  const int iter0 = s->iter;
  // every chain program either evaluates f (at most 20 per iteration) or
  // restarts; this bound is never reached by a correct chain
  long long guard = 64LL * ((long long)s->max_iteration + 2) + 4096;
  while (s->status != ST_DONE) {
    if (iter_budget > 0 && s->iter - iter0 >= iter_budget) break;
    if (--guard < 0) {
      if (threadIdx.x == 0) set_done(s, -2);
      __syncthreads();
      break;
    }
    const int prog = s->b_next_prog;
    const double fv = s->b_next_f;
    __syncthreads();
    b_program(prog, fv, P, L);
    const int st = s->status;
    if (st == ST_NEED_EVAL) {
      b_objective(P, L, s->trial_dst);
      if (threadIdx.x == 0) {
        s->b_next_f = L.f;
        s->b_next_prog = PROG_AFTER_EVAL;
      }
    } else if (st == ST_RESTART) {
      if (threadIdx.x == 0) s->b_next_prog = PROG_ITER_START;
    } else if (st != ST_DONE) {
      // a chain that ends without a request is an internal error
      if (threadIdx.x == 0) set_done(s, -2);
    }
    __syncthreads();
  }
  const bool done = s->status == ST_DONE;
  if (done && (s->phase & 1)) {
    // leave the accepted iterate in the caller's buffer
    for (long long i = threadIdx.x; i < P.n; i += blockDim.x) P.X[0][i] = P.X[1][i];
  }
  if (threadIdx.x == 0) {
    res->f = s->f;
    res->residual_g = s->residual_g;
    res->residual_f = s->residual_f;
    res->residual_x = s->residual_x;
    res->iter = s->iter;
    res->info = s->info;
    res->nfgv = s->nfgv;
    res->done = done ? 1 : 0;
    if (done) res->clk_end = (long long)wall_clock64();
  }
  __syncthreads();
  {
    const unsigned long long* src = (const unsigned long long*)&L.s;
    unsigned long long* dst = (unsigned long long*)dsg;
    for (int i = threadIdx.x; i < kBDsWords; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT)
k_batch(const BatchProb* __restrict__ probs, const int* __restrict__ order, int count,
        unsigned* next, BatchBufs<T> B, DevScalars* dsa, BatchResult* res, int iter_budget) {
  __shared__ BatchLds L;
  __shared__ int s_k;
  for (;;) {
    if (threadIdx.x == 0) s_k = (int)atomicAdd(next, 1u);
    __syncthreads();
    const int k = s_k;
    __syncthreads();
    if (k >= count) return;
    const int p = order[k];
    const BatchProb bp = probs[p];
    solve_one<T>(bp, B, dsa + p, res + p, iter_budget, L);
  }
}

}  // namespace

template <typename T>
int launch_batch(hipStream_t s, int threads, int blocks, const BatchProb* probs, const int* order,
                 int count, unsigned* next, BatchBufs<T> B, DevScalars* dsa, BatchResult* res,
                 int iter_budget) {
  switch (threads) {
    case 256:
      hipLaunchKernelGGL((k_batch<T, 256>), dim3(blocks), dim3(256), 0, s, probs, order, count,
                         next, B, dsa, res, iter_budget);
      return 0;
    case 512:
      hipLaunchKernelGGL((k_batch<T, 512>), dim3(blocks), dim3(512), 0, s, probs, order, count,
                         next, B, dsa, res, iter_budget);
      return 0;
    default:
      return -1;
  }
}

template int launch_batch<float>(hipStream_t, int, int, const BatchProb*, const int*, int,
                                 unsigned*, BatchBufs<float>, DevScalars*, BatchResult*, int);
template int launch_batch<double>(hipStream_t, int, int, const BatchProb*, const int*, int,
                                  unsigned*, BatchBufs<double>, DevScalars*, BatchResult*, int);

}  // namespace lbfgsb
