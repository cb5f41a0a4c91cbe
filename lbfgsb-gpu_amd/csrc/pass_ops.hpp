// Element operators of the bulk passes.  Each operator holds the uniform
// parameters of one pass (read from the device state), the per-thread
// accumulators, and three steps per element: load (every global load of the
// element, issued together), keep (see keep_loads) and element (the
// arithmetic and the stores).  The grid-wide kernels of bulk.hip and the
// per-problem batched kernel drive the same operators with different loops
// and reductions, so both forms compute the same per-element values.
#pragma once

#include "common.hpp"
#include "pass_device.hpp"
#include "state.hpp"

namespace lbfgsb {

// FP64 MFMA accumulator (four doubles per lane).
typedef double GramAcc __attribute__((ext_vector_type(4)));
// rows of 16 the Cauchy-point pass stages per wave for its Gram, and the
// doubles of LDS per wave that holds them (the rows, then their signs)
constexpr int kRowRows = 8;
constexpr int kRowBuf = kRowRows * 16 + kRowRows;

// Value of signed Gram pair t (< 136: YY lower, SS lower, SY; corr_pair)
// from the summed 16 x 16 Gram g (row-major), over the memory positions
// below upcl (0 for the others).
__device__ __forceinline__ double corr_pair_value(int t, const double* g, int upcl) {
  int cat = 0, a = 0, b = 0;
  corr_pair(t, cat, a, b);
  if (a >= upcl || b >= upcl) return 0.0;
  // YY (cat 0), SS (cat 1), SY (cat 2: a the S position, b the Y position)
  const int ra = cat == 0 ? a : kMaxM + a;
  const int rb = cat == 1 ? kMaxM + b : b;
  return g[ra * 16 + rb];
}

// ---------------------------------------------------------------------------
// Pass A: the iteration tail (projected gradient norm, ||x - x_old||^2,
// y = g - g_old, y'y, the new S/Y pair written at ds->itail and its Gram row
// s'y_j, s_j's) when FUSED, and the Cauchy classification sums (f1, p = W'd
// over the GCP-free set, breakpoint counts, smallest breakpoint)
// (lbfgsbcpu.cpp:394-449, 567-701, 1341-1387).
template <typename T, bool FUSED>
struct PassAOp {
  const T* __restrict__ x;
  const T* __restrict__ g;
  const T* __restrict__ xold;
  const T* __restrict__ gold;
  const T* __restrict__ l;
  const T* __restrict__ u;
  uint8_t* __restrict__ iwhere;
  T* __restrict__ W;
  const T* __restrict__ d;
  // plain form: the breakpoints with t below ntau are listed in nout (at most
  // ncap, the count in the block's LDS counter ncnt) and the class is not
  // stored; null: the class is stored (batched form)
  Cand* __restrict__ nout;
  unsigned* ncnt;
  unsigned ncap;
  double ntau;
  int col, itail, wnew;
  bool upd, stp1, dstored, spec;
  T stp;
  RingReads rr;
  double acc[PA_COUNT];
  double sbg;
  double pny, pns;  // new pair's share of p (position col - 1)
  double bkmin;
  long long ibk;

  struct In {
    T x, g, l, u, xo, go, dd;
    uint8_t iwb;
    T yv[kMaxM], sv[kMaxM];
  };

  // spec: run on the evaluated trial point before the line search has
  // accepted it (FUSED only): the trial buffers are the new iterate, the
  // ring is the one an accepted step with a memory update leaves, the new
  // pair goes to the spare physical slot, and g'd is summed as well.
  __device__ __forceinline__ void init(const DevScalars* ds, Pair<T> X, Pair<T> Gv, const T* l_,
                              const T* u_, uint8_t* iwhere_, T* W_,
                              const T* d_, Cand* nout_ = nullptr, unsigned* ncnt_ = nullptr,
                              unsigned ncap_ = 0, bool spec_ = false) {
    spec = FUSED && spec_;
    const int cur = (ds->phase & 1) ^ (spec ? 1 : 0);
    nout = nout_;
    ncnt = ncnt_;
    ncap = ncap_;
    ntau = ds->gcp_near_tau;
    x = X[cur];
    g = Gv[cur];
    xold = X[cur ^ 1];
    gold = Gv[cur ^ 1];
    l = l_;
    u = u_;
    iwhere = iwhere_;
    W = W_;
    d = d_;
    stp = (T)ds->stp;
    stp1 = ds->stp == 1.0;
    dstored = ds->d_stored;
    if (spec) {
      // the ring after accept_step with updatd = 1 (lbfgsbcpu.cpp:437-449)
      const int m = ds->m, iu = ds->iupdat + 1;
      int head = ds->head;
      if (iu <= m) {
        col = iu;
        itail = (head + iu - 1) % m;
      } else {
        col = ds->col;
        itail = (ds->itail + 1) % m;
        head = (head + 1) % m;
      }
      upd = true;
      rr = ring_reads(head, m, col, itail, ds->pslot_x);
      wnew = ring_spare(ds);
    } else {
      col = ds->col;
      itail = ds->itail;
      upd = FUSED && ds->updatd;
      // the slot being replaced by the new pair is not read
      rr = ring_reads(ds->head, ds->m, col, upd ? itail : -1, ds->pslot_x);
      wnew = ring_phys(ds, itail);  // physical slot the new pair goes to
    }
#pragma unroll
    for (int v = 0; v < PA_COUNT; v++) acc[v] = 0.0;
    sbg = 0.0;
    pny = pns = 0.0;
    bkmin = __builtin_inf();
    ibk = 0x7fffffffffffffffLL;
  }
  __device__ __forceinline__ void load(In& e, long long i) const {
    e.x = x[i];
    e.g = g[i];
    e.l = l[i];
    e.u = u[i];
    e.iwb = iwhere[i];
    e.xo = FUSED ? xold[i] : (T)0;
    e.go = upd ? gold[i] : (T)0;
    e.dd = upd && dstored ? d[i] : (T)0;
    load_ring_raw(W, i, rr, e.yv, e.sv);
  }
  __device__ __forceinline__ void keep(const In& e) const {
    keep_loads(e.x, e.g, e.l, e.u, e.iwb, e.xo, e.go, e.dd, e.yv, e.sv);
  }
  __device__ __forceinline__ void element(long long i, In& e) {
    mask_ring(rr, e.yv, e.sv);
    const T xi = e.x, gi = e.g, li = e.l, ui = e.u;
    const int nb = iw_nbd(e.iwb);
    const uint8_t iwb = e.iwb;
    T ys = 0, ss = 0;
    if (FUSED) {
      const T gp = projected_gi(xi, gi, li, ui, nb);
      sbg = fmax(sbg, (double)fabs(gp));
      const T xd = e.xo - xi;
      acc_prod(acc[PA_TF], xd, xd);
      if (upd) {
        ys = gi - e.go;
        // d = z - x_old; unless a trial stored it, the accepted point is z
        const T dd = dstored ? e.dd : xi - e.xo;
        ss = stp1 ? dd : dd * stp;
        acc_prod(acc[PA_RR], ys, ys);
        if (spec) acc_prod(acc[PA_GD], gi, dd);  // g'd of the line search
        st_nt(&W[widx(i, wnew)], ys);
        st_nt(&W[widx(i, kWPhys + wnew)], ss);
      }
    }
    // Unused positions hold zeros (mask_ring), and with an update the
    // newest position col - 1 (slot itail, not read) is zero too: the sums
    // run over all positions without guards, and the new pair's share of p
    // goes to pny / pns (added at position col - 1 by finish()).
    if (upd) {
#pragma unroll
      for (int j = 0; j < kMaxM - 1; j++) {
        acc_prod(acc[PA_SY + j], ss, e.yv[j]);
        acc_prod(acc[PA_SS + j], e.sv[j], ss);
      }
    }
    int iw;
    T neggi, tbrk;
    bool isfree, isbrk;
    classify(xi, gi, li, ui, nb, iw_get(iwb), iw, neggi, isfree, isbrk, tbrk);
    // GCP-free variables feed f1 and p; the others add an exact zero
    const T nf = isfree ? neggi : (T)0;
    acc_prod(acc[PA_F1], nf, nf);
#pragma unroll
    for (int j = 0; j < kMaxM; j++) {
      acc_prod(acc[PA_P + j], e.yv[j], nf);
      acc_prod(acc[PA_P + 8 + j], e.sv[j], nf);
    }
    if (upd) {
      acc_prod(pny, ys, nf);
      acc_prod(pns, ss, nf);
    }
    if (isfree) {
      if (isbrk) {
        acc[PA_NBREAK] += 1.0;
        const double tt = (double)tbrk;
        if (tt < bkmin || (tt == bkmin && i < ibk)) {
          bkmin = tt;
          ibk = i;
        }
      } else {
        acc[PA_NFREEGCP] += 1.0;
        if (fabs(neggi) > 0) acc[PA_UNBND] = 1.0;  // an OR, kept 0/1 through the sum
      }
    }
    if (nout) {
      // the Cauchy search selects its candidates from these when its t-range
      // lies below ntau (the same t the selection pass computes from x, g,
      // l, u and the stored class); the class itself is not stored: the
      // Cauchy-point pass reclassifies from the stored one (classification
      // is idempotent), so this pass writes no per-variable stream for the
      // search
      // (one LDS atomic per wave: the wave's listed lanes take consecutive
      // slots; lanes of the wave that left the loop take no part)
      const bool lst = isbrk && (double)tbrk < ntau;
      const unsigned long long lm = __ballot(lst);
      if (lm) {
        const int lane = threadIdx.x & 63, lead = __ffsll((long long)lm) - 1;
        unsigned base = 0u;
        if (lane == lead) base = atomicAdd(ncnt, (unsigned)__popcll(lm));
        base = __shfl(base, lead, kWave);
        const unsigned k = base + (unsigned)__popcll(lm & ((1ull << lane) - 1ull));
        if (lst && k < ncap) {
          Cand c;
          c.t = (double)tbrk;
          c.idx = (int)i;
          c.pad = 0;
          nout[k] = c;
        }
      }
    } else {
      iwhere[i] = iw_make(iw, iw_wasfree(iwb), iw_prevfree(iwb), nb);
    }
  }
  __device__ __forceinline__ void finish() {
    if (upd) {
#pragma unroll
      for (int j = 0; j < kMaxM; j++) {
        acc[PA_P + j] += j == col - 1 ? pny : 0.0;
        acc[PA_P + 8 + j] += j == col - 1 ? pns : 0.0;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// The Cauchy point of one variable from the search's outcome
// (lbfgsbcpu.cpp:705-817) and the reduced gradient there
// (lbfgsbcpu.cpp:1446-1458).  The classification is idempotent and the
// outcome (mode, tsum, last passed breakpoint, M c, theta) is fixed from the
// end of the search to the line search, so the Cauchy-point pass and the
// subspace passes after it compute the same x_cp and r from x, g, l, u, the
// stored class and W, bit for bit: neither vector is stored.
template <typename T>
struct CauchyOutcome {
  int mode;
  bool shortcut;
  T tsum;
  double t_last;
  long long i_last;
  double theta;
  double mcy[kMaxM], mcs[kMaxM];

  __device__ __forceinline__ void init(const DevScalars* ds) {
    shortcut = ds->shortcut;
    mode = ds->gcp_mode;
    tsum = (T)ds->gcp_tsum;
    t_last = ds->gcp_last_t;
    i_last = ds->gcp_last_idx;
    theta = ds->theta;
    const int col = ds->col;
#pragma unroll
    for (int j = 0; j < kMaxM; j++) {
      mcy[j] = j < col ? ds->mc[j] : 0.0;
      mcs[j] = j < col ? theta * ds->mc[col + j] : 0.0;
    }
  }
  // x_cp of variable i; iw: the stored class in, the class after the
  // search out (unchanged under the unconstrained shortcut).
  __device__ __forceinline__ T point(long long i, T xi, T gi, T li, T ui, int nb,
                                     int& iw) const {
    T zi = xi;
    if (!shortcut) {
      T neggi, tb;
      bool isfree, isbrk;
      classify(xi, gi, li, ui, nb, iw, iw, neggi, isfree, isbrk, tb);
      const T dcp = isfree ? neggi : (T)0;
      if (mode == 3) {
        if (isbrk) {
          iw = dcp > 0 ? 2 : 1;
          zi = dcp > 0 ? ui : li;
        }
      } else if (mode != 2) {
        bool passed = false;
        if (mode == 1 && isbrk) {
          const double tt = (double)tb;
          passed = tt < t_last || (tt == t_last && i <= i_last);
        }
        if (passed) {
          iw = dcp > 0 ? 2 : 1;
          zi = dcp > 0 ? ui : li;
        } else {
          zi = xi + tsum * dcp;
        }
      }
    }
    return zi;
  }
  // r_i = -theta (x_cp - x) - g + (W M c)_i in FP64 (-g under the shortcut),
  // over a masked row (unused positions zero); callers round it to T.
  __device__ __forceinline__ double reduced(T xi, T gi, T zi, const T (&yv)[kMaxM],
                                            const T (&sv)[kMaxM]) const {
    double ri = -(double)gi;
    if (!shortcut) {
      ri = -theta * (double)(zi - xi) - (double)gi;
#pragma unroll
      for (int j = 0; j < kMaxM; j++) {
        ri = __builtin_fma((double)yv[j], mcy[j], ri);
        ri = __builtin_fma((double)sv[j], mcs[j], ri);
      }
    }
    return ri;
  }
};

// ---------------------------------------------------------------------------
// Cauchy-point pass: x_cp from the search's outcome, free / entering /
// leaving sets, the new formk column, the signed entering/leaving Gram
// pairs, r = -theta (z - x) - g + W M c and W'r over the free set
// (lbfgsbcpu.cpp:705-817, 819-869, 907-1054, 1135-1173, 1446-1458).
template <typename T>
struct P4Op {
  static constexpr int kAcc = P4_CORR;  // 3 counts, 32 new-column sums, 16 W'r
  const T* __restrict__ x;
  const T* __restrict__ g;
  const T* __restrict__ l;
  const T* __restrict__ u;
  uint8_t* __restrict__ iwhere;
  const T* __restrict__ W;
  DevScalars* ds;
  int col;
  bool shortcut, track, newcol;
  CauchyOutcome<T> co;
  RingReads rr;
  double acc[kAcc];
  // signed Gram of this wave's entering (+1) / leaving (-1) variables,
  // G = sum_f sign_f w_f w_f' over their rows w_f = [y_0..7, s_0..7]
  // (16 x 16, an FP64 MFMA accumulator: G[i][j] in lane j + 16 (i % 4),
  // element i / 4), fed four rows at a time
  GramAcc gacc;

  struct In {
    T x, g, l, u;
    uint8_t iwb;
    T yv[kMaxM], sv[kMaxM];
    T yn, sn;  // the newest pair again (a cache hit): no select over positions
  };

  __device__ __forceinline__ void init(DevScalars* ds_, Pair<T> X, Pair<T> Gv, const T* l_, const T* u_,
                              uint8_t* iwhere_, const T* W_) {
    ds = ds_;
    const int cur = ds->phase & 1;
    x = X[cur];
    g = Gv[cur];
    l = l_;
    u = u_;
    iwhere = iwhere_;
    W = W_;
    const int m = ds->m, head = ds->head;
    col = ds->col;
    shortcut = ds->shortcut;
    co.init(ds);
    track = ds->iter > 0 && ds->cnstnd && !shortcut;
    newcol = ds->updatd && col > 0;
    rr = ring_reads(head, m, col, -1, ds->pslot_x);
#pragma unroll
    for (int v = 0; v < kAcc; v++) acc[v] = 0.0;
    gacc = GramAcc{0.0, 0.0, 0.0, 0.0};
  }
  __device__ __forceinline__ void load(In& e, long long i) const {
    e.x = x[i];
    e.g = g[i];
    e.l = l[i];
    e.u = u[i];
    e.iwb = iwhere[i];
    load_ring_raw(W, i, rr, e.yv, e.sv);
    const T* __restrict__ row = W + widx((size_t)i, 0);
    e.yn = row[rr.newest * kTile];
    e.sn = row[(kWPhys + rr.newest) * kTile];
  }
  __device__ __forceinline__ void keep(const In& e) const {
    keep_loads(e.x, e.g, e.l, e.u, e.iwb, e.yv, e.sv, e.yn, e.sn);
  }
  // Returns +1 / -1 for an entering / leaving variable, 0 otherwise.
  __device__ __forceinline__ int element(long long i, In& e) {
    mask_ring(rr, e.yv, e.sv);
    int flag = 0;
    const T xi = e.x, gi = e.g, li = e.l, ui = e.u;
    const int nb = iw_nbd(e.iwb);
    const uint8_t iwb = e.iwb;
    const bool wasfree = iw_wasfree(iwb);
    int iw = iw_get(iwb);
    const T zi = co.point(i, xi, gi, li, ui, nb, iw);
    const bool freenow = iw <= 0;
    if (!shortcut) {
      acc[P4_NFREE] += freenow ? 1.0 : 0.0;
      if (track) {
        if (!wasfree && freenow) {
          acc[P4_NENTER] += 1.0;
          flag = 1;
        }
        if (wasfree && !freenow) {
          acc[P4_NLEAVE] += 1.0;
          flag = -1;
        }
      }
      st_nt(&iwhere[i], iw_make(iw, freenow, wasfree, nb));
    }
    // Positions j >= col hold zeros (mask_ring), so the sums below run over
    // all kMaxM positions without guards: unused accumulators collect exact
    // zeros and are never read.
    if (col > 0) {
      if (newcol) {
        // free variables feed the Y'Y / S'Y sums, active ones the S'S / S'Y
        // sums; the other side adds an exact zero (no divergent stores,
        // which the compiler would merge into a dynamically indexed one)
        const T yf = freenow ? e.yn : (T)0, sa = freenow ? (T)0 : e.sn;
#pragma unroll
        for (int j = 0; j < kMaxM; j++) {
          acc_prod(acc[P4_NC + j], yf, e.yv[j]);
          acc_prod(acc[P4_NC + 24 + j], e.sv[j], yf);
          acc_prod(acc[P4_NC + 8 + j], sa, e.sv[j]);
          acc_prod(acc[P4_NC + 16 + j], sa, e.yv[j]);
        }
      }
      // r over the free set; 0 elsewhere, so W'r needs no branch
      const T rt = freenow ? (T)co.reduced(xi, gi, zi, e.yv, e.sv) : (T)0;
#pragma unroll
      for (int j = 0; j < kMaxM; j++) {
        acc_prod(acc[P4_WV + j], e.yv[j], rt);
        acc_prod(acc[P4_WV + 8 + j], e.sv[j], rt);
      }
    }
    return flag;
  }
  // Signed Gram over this wave's entering / leaving variables (all lanes of
  // the wave take part).  The flagged variables' lanes have their W rows in
  // registers (e, masked by element()); up to kRowRows of them at a time put
  // their rows and signs into the wave's rowbuf (kRowBuf doubles of LDS) at
  // their rank among the flagged lanes, and every four rows go into the
  // accumulator with one FP64 MFMA (16 x 16 x 4): no flagged variable waits
  // for a load from memory, and a wave whose lanes are all flagged (a free
  // set that changes wholesale) stages its rows eight at a time.
  __device__ __forceinline__ void corr_step(long long i, int flag, const In& e, double* rowbuf) {
    if (!(track && col > 0)) return;
    const int lane = threadIdx.x & 63;
    const unsigned long long fm = __ballot(flag != 0);
    if (!fm) return;
    if (lane == 0) ds->p4_flagged = 1;
    const int nf = __popcll(fm);
    const int rank = __popcll(fm & ((1ull << lane) - 1ull));
    for (int base = 0; base < nf; base += kRowRows) {
      if (flag != 0 && rank >= base && rank < base + kRowRows) {
        const int r = rank - base;
#pragma unroll
        for (int j = 0; j < kMaxM; j++) {
          rowbuf[r * 16 + j] = (double)e.yv[j];
          rowbuf[r * 16 + kMaxM + j] = (double)e.sv[j];
        }
        rowbuf[kRowRows * 16 + r] = (double)flag;
      }
      const int rows = nf - base < kRowRows ? nf - base : kRowRows;
      for (int q = 0; q < rows; q += 4)
        gram_feed(rowbuf + q * 16, rowbuf + kRowRows * 16 + q, rows - q);
    }
  }
  // Four rows at rows16 (16 doubles each; signs at sgn) into the accumulator,
  // the rows from `rows` on taken as zero: A[i][k] = w_k[i] and
  // B[k][j] = sign_k w_k[j] sit in lane i + 16 k and j + 16 k alike, so lane
  // l takes rows16[l] (exact in FP64: the products of two bulk values are).
  __device__ __forceinline__ void gram_feed(const double* rows16, const double* sgn, int rows) {
    const int lane = threadIdx.x & 63, k = lane >> 4;
    // the wave's LDS operations complete in order: the rows are in place
    // (the wave barriers keep the compiler from moving the reads up, or the
    // next rows' writes above them)
    __builtin_amdgcn_wave_barrier();
    // rows past `rows` are zero in both operands (their LDS is stale: the
    // sign slot may hold anything, a NaN included)
    const bool live = k < rows;
    const double a = live ? rows16[lane] : 0.0;
    const double b = live ? a * sgn[k] : 0.0;
    gacc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, gacc, 0, 0, 0);
    __builtin_amdgcn_wave_barrier();
  }
  // The workgroup's Gram into gram[16 * i + j] (256 doubles of LDS): the
  // waves' accumulators added in wave order, one wave after another.  All
  // threads of the workgroup call it (it holds nwaves barriers).
  __device__ __forceinline__ void corr_block_sum(double* gram, int nwaves) const {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int w = 0; w < nwaves; w++) {
      if (wid == w) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int q = ((lane >> 4) + 4 * r) * 16 + (lane & 15);
          gram[q] = w == 0 ? gacc[r] : gram[q] + gacc[r];
        }
      }
      __syncthreads();
    }
  }
};

// ---------------------------------------------------------------------------
// Subspace direction over the free variables and the step-to-box ratio
// (lbfgsbcpu.cpp:1472-1516).
template <typename T>
struct P5aOp {
  const T* __restrict__ x;
  const T* __restrict__ g;
  const T* __restrict__ l;
  const T* __restrict__ u;
  const uint8_t* __restrict__ iwhere;
  const T* __restrict__ W;
  T* __restrict__ dsub;
  int col;
  bool shortcut;
  double theta;
  CauchyOutcome<T> co;
  RingReads rr;
  double wy[kMaxM], wsv[kMaxM];
  double amin;
  long long ai;

  struct In {
    T x, g, l, u;
    uint8_t iwb;
    T yv[kMaxM], sv[kMaxM];
  };

  __device__ __forceinline__ void init(const DevScalars* ds, Pair<T> X, Pair<T> Gv, const T* l_,
                                       const T* u_, const uint8_t* iwhere_,
                                       const T* W_, T* dsub_) {
    const int cur = ds->phase & 1;
    x = X[cur];
    g = Gv[cur];
    l = l_;
    u = u_;
    iwhere = iwhere_;
    W = W_;
    dsub = dsub_;
    col = ds->col;
    shortcut = ds->shortcut;
    theta = ds->theta;
    co.init(ds);
#pragma unroll
    for (int j = 0; j < kMaxM; j++) {
      wy[j] = j < col ? ds->wv[j] / theta : 0.0;
      wsv[j] = j < col ? ds->wv[col + j] : 0.0;
    }
    rr = ring_reads(ds->head, ds->m, col, -1, ds->pslot_x);
    amin = 1.0;
    ai = 0x7fffffffffffffffLL;
  }
  __device__ __forceinline__ void load(In& e, long long i) const {
    e.iwb = iwhere[i];
    e.x = x[i];
    e.g = g[i];
    e.l = l[i];
    e.u = u[i];
    load_ring_raw(W, i, rr, e.yv, e.sv);
  }
  __device__ __forceinline__ void keep(const In& e) const {
    keep_loads(e.x, e.g, e.l, e.u, e.iwb, e.yv, e.sv);
  }
  __device__ __forceinline__ void element(long long i, In& e) {
    int iw = iw_get(e.iwb);
    const bool freenow = shortcut || iw <= 0;
    if (!freenow) return;
    mask_ring(rr, e.yv, e.sv);
    // x_cp and r as the Cauchy-point pass formed them (r rounded to T)
    const T zi = co.point(i, e.x, e.g, e.l, e.u, iw_nbd(e.iwb), iw);
    const T rt = (T)co.reduced(e.x, e.g, zi, e.yv, e.sv);
    double di = (double)rt;
#pragma unroll
    for (int j = 0; j < kMaxM; j++)
      if (j < col) di += (double)e.yv[j] * wy[j] + (double)e.sv[j] * wsv[j];
    const T dk = (T)(di / theta);
    st_nt(&dsub[i], dk);
    const int nb = iw_nbd(e.iwb);
    if (nb != 0) {
      double cand = 2.0;
      if (dk < 0 && nb <= 2) {
        const T t2 = e.l - zi;
        if (t2 >= 0)
          cand = 0.0;
        else if (dk < t2)
          cand = (double)(t2 / dk);
      } else if (dk > 0 && nb >= 2) {
        const T t2 = e.u - zi;
        if (t2 <= 0)
          cand = 0.0;
        else if (dk > t2)
          cand = (double)(t2 / dk);
      }
      if (cand < amin || (cand == amin && cand < 1.0 && i < ai)) {
        amin = cand;
        ai = i;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Backtrack (pinning the limiting variable), d = z - x, and the line-search
// set-up sums d'd, g'd and the nbd-aware max step (lbfgsbcpu.cpp:1517-1533,
// 343-345, 1263-1299).  z goes into the trial buffer X[cur ^ 1], where it is
// the first trial point of a unit step; the trial pass recomputes it with
// subspace_point() if a later unit-step trial finds it overwritten.
template <typename T>
struct P5bOp {
  const T* __restrict__ x;
  const T* __restrict__ g;
  const T* __restrict__ l;
  const T* __restrict__ u;
  const uint8_t* __restrict__ iwhere;
  T* __restrict__ z;
  const T* __restrict__ dsub;
  bool sub, shortcut, pin, want_stpmx;
  T alpha;
  long long ibd;
  CauchyOutcome<T> co;
  double acc[P5_COUNT];
  double smin;

  struct In {
    T x, g, l, u, dsub;
    uint8_t iwb;
  };

  // cur_buf: the buffer pair the subspace step was formed from; by default
  // the accepted one (phase & 1).  The trial pass names it (trial_src),
  // because a step that the line search accepts without evaluating it
  // (iback >= 20) flips the phase before that pass runs.
  __device__ __forceinline__ void init(const DevScalars* ds, Pair<T> X, Pair<T> Gv, const T* l_,
                              const T* u_, const uint8_t* iwhere_,
                              const T* dsub_, int cur_buf = -1) {
    const int cur = cur_buf >= 0 ? cur_buf : (ds->phase & 1);
    x = X[cur];
    g = Gv[cur];
    l = l_;
    u = u_;
    iwhere = iwhere_;
    z = X[cur ^ 1];
    dsub = dsub_;
    sub = !ds->skip_subsm;
    shortcut = ds->shortcut;
    alpha = (T)ds->alpha;
    pin = ds->alpha < 1.0;
    ibd = ds->ibd;
    co.init(ds);
    want_stpmx = ds->cnstnd && ds->iter > 0;
    acc[0] = acc[1] = 0.0;
    smin = 1.0e10;
  }
  __device__ __forceinline__ void load(In& e, long long i) const {
    e.x = x[i];
    e.g = g[i];
    e.l = l[i];
    e.u = u[i];
    e.iwb = iwhere[i];
    e.dsub = sub ? dsub[i] : (T)0;
  }
  __device__ __forceinline__ void keep(const In& e) const {
    keep_loads(e.x, e.g, e.l, e.u, e.iwb, e.dsub);
  }
  // The subspace point of variable i (x_cp where there is no subspace step).
  __device__ __forceinline__ T subspace_point(long long i, const In& e) const {
    int iw = iw_get(e.iwb);
    const bool freenow = shortcut || iw <= 0;
    T zi = co.point(i, e.x, e.g, e.l, e.u, iw_nbd(e.iwb), iw);  // x_cp
    if (sub && freenow) {
      T dk = e.dsub;
      if (pin && i == ibd) {
        if (dk > 0) {
          zi = e.u;
          dk = 0;
        } else if (dk < 0) {
          zi = e.l;
          dk = 0;
        }
      }
      zi = zi + alpha * dk;
    }
    return zi;
  }
  __device__ __forceinline__ void element(long long i, const In& e) {
    const T xi = e.x;
    const T zi = subspace_point(i, e);
    z[i] = zi;  // read back soon (objective, pass A on the trial point): normal stores
    const T di = zi - xi;  // d is not stored: z - x while z is the trial point
    acc[P5_DTD] += (double)di * (double)di;
    acc[P5_GD] += (double)e.g * (double)di;
    if (want_stpmx) {
      // the line search's largest feasible step is an FP64 scalar: each
      // ratio is formed in FP64 (the bound distance of two bulk values is
      // exact there), so the fp32 form takes the decisions the FP64 search
      // takes on the same data
      constexpr double big = 1.0e10;
      const int nb = iw_nbd(e.iwb);
      const double dd = (double)di;
      if (nb != 0) {
        if (di < 0 && nb <= 2) {
          const double a2 = (double)e.l - (double)xi;
          if (a2 >= 0)
            smin = 0.0;
          else if (dd * big < a2)
            smin = fmin(smin, a2 / dd);
        } else if (di > 0 && nb >= 2) {
          const double a2 = (double)e.u - (double)xi;
          if (a2 <= 0)
            smin = 0.0;
          else if (dd * big > a2)
            smin = fmin(smin, a2 / dd);
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// The line search on the trial buffer (lbfgsbcpu.cpp:1312-1335).  The
// subspace pass leaves z in X[cur ^ 1] and stores no d: while z is the trial
// point, d = z - x is the difference of the two x buffers; the first trial
// that moves off z stores d first (trial_d).

// d_i for g'd after an evaluation of the trial point.
template <typename T>
struct LsDirection {
  bool stored;
  const T* __restrict__ d;
  const T* __restrict__ x;
  const T* __restrict__ xt;
  __device__ __forceinline__ void init(const DevScalars* ds, Pair<T> X, const T* d_) {
    const int cur = ds->phase & 1;
    stored = ds->d_stored;
    d = d_;
    x = X[cur];
    xt = X[cur ^ 1];
  }
  __device__ __forceinline__ T at(long long i) const {
    return stored ? d[i] : (T)(xt[i] - x[i]);
  }
};

// Trial point i: x_t = stp d + x, or z for a unit step (recomputed when a
// trial since has overwritten it; then d is stored).
template <typename T>
__device__ __forceinline__ void trial_element(long long i, const DevScalars* ds,
                                              const P5bOp<T>& sub, const T* __restrict__ xa,
                                              T* __restrict__ xb, T* __restrict__ d, double stp,
                                              bool trial_d) {
  if (stp == 1.0) {
    typename P5bOp<T>::In e;
    sub.load(e, i);
    xb[i] = sub.subspace_point(i, e);
  } else if (trial_d) {
    const T di = xb[i] - xa[i];  // xb still holds z
    d[i] = di;
    xb[i] = (T)(stp * (double)di + (double)xa[i]);
  } else {
    xb[i] = (T)(stp * (double)d[i] + (double)xa[i]);
  }
  (void)ds;
}

}  // namespace lbfgsb
