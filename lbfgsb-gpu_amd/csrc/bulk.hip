// Bulk n-length passes of one L-BFGS-B iteration (HBM-streaming kernels).
//
// Pass map (one launch each, every launch reads each n-vector at most once):
//   k_active    projection of x0 and iwhere set-up     (lbfgsbcpu.cpp:467-515)
//   k_errclb    input validation counts                (lbfgsbcpu.cpp:871-905)
//   k_projgr    projected-gradient norm                (lbfgsbcpu.cpp:1389-1414)
//   k_pass_a    [iteration tail: projgr, ||dx||, y=g-g_old, y'y, ring write,
//                new Gram row]  + Cauchy classification, f1, p = W'd and the
//                breakpoint list                        (lbfgsbcpu.cpp:394-449,
//                                                        567-701, 1341-1387)
//   k_p4        Cauchy point, free/active counts, new formk column, signed
//               entering/leaving Gram, reduced gradient r and W'r
//                                                      (lbfgsbcpu.cpp:705-817,
//                                                        819-869, 907-1054,
//                                                        1135-1173, 1446-1458)
//   k_p5a       subspace direction and step-to-box ratio (lbfgsbcpu.cpp:1472-1516)
//   k_p5b       backtrack, d = z - x, d'd, g'd, stpmx     (lbfgsbcpu.cpp:1517-1533,
//                                                        343-345, 1263-1299)
//   k_trial     line-search trial point                  (lbfgsbcpu.cpp:1329-1335)
//   k_gd        g'd after an evaluation                   (lbfgsbcpu.cpp:1312)
//
// Every kernel is a grid-stride loop over 256-thread blocks; per-block
// partial sums go to part[v * G + block] and are finished by the
// single-workgroup FP64 kernels in small.hip.  Products of the bulk type are
// formed and accumulated in FP64 (an FP32 x FP32 product is exact in FP64),
// and the cross-block sums are compensated (two-sum), which is the
// compensated-dot scheme used for the FP32 bulk path.
#include "common.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

// Reclassification of one variable for the Cauchy search
// (lbfgsbcpu.cpp:634-700).  Idempotent: applying it to an already
// classified iwhere gives the same result, so later passes recompute the
// breakpoint t instead of storing it.
template <typename T>
__device__ inline void classify(T x, T g, T l, T u, int nb, int iw_in, int& iw,
                                T& neggi, bool& isfree, bool& isbrk, T& t) {
  neggi = -g;
  iw = iw_in;
  T tl = 0, tu = 0;
  if (iw != 3 && iw != -1) {
    if (nb <= 2) tl = x - l;
    if (nb >= 2) tu = u - x;
    const bool xlower = nb <= 2 && tl <= 0;
    const bool xupper = nb >= 2 && tu <= 0;
    iw = 0;
    if (xlower) {
      if (neggi <= 0) iw = 1;
    } else if (xupper) {
      if (neggi >= 0) iw = 2;
    } else {
      if (fabs(neggi) <= 0) iw = -3;
    }
  }
  isfree = (iw == 0 || iw == -1);
  isbrk = false;
  t = 0;
  if (isfree) {
    if (nb <= 2 && nb != 0 && neggi < 0) {
      isbrk = true;
      t = tl / (-neggi);
    } else if (nb >= 2 && neggi > 0) {
      isbrk = true;
      t = tu / neggi;
    }
  }
}

__device__ inline int grid_blocks() { return gridDim.x; }

// Everything the Cauchy-point pass reads for one element; yn / sn repeat
// the newest pair (position col - 1, a cache hit) so that no per-element
// select over positions is needed.
template <typename T>
struct BulkIn {
  T x, g, l, u;
  int nb;
  uint8_t iwb;
  T yv[kMaxM], sv[kMaxM];
  T yn, sn;
};
template <typename T>
__device__ inline void load_bulk_in(BulkIn<T>& e, long long i, const T* __restrict__ x,
                                    const T* __restrict__ g, const T* __restrict__ l,
                                    const T* __restrict__ u, const int* __restrict__ nbd,
                                    const uint8_t* __restrict__ iwhere,
                                    const T* __restrict__ W, const RingReads& rr) {
  e.x = x[i];
  e.g = g[i];
  e.l = l[i];
  e.u = u[i];
  e.nb = nbd[i];
  e.iwb = iwhere[i];
  load_ring(W, i, rr, e.yv, e.sv);
  const T* __restrict__ row = W + widx((size_t)i, 0);
  e.yn = row[rr.newest * kTile];
  e.sn = row[(8 + rr.newest) * kTile];
}
template <typename T>
__device__ inline void keep_bulk_in(const BulkIn<T>& e) {
  keep_loads(e.x, e.g, e.l, e.u, e.nb, e.iwb, e.yv, e.sv, e.yn, e.sn);
}

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
    // every variable starts in the free set of the (empty) previous partition
    iwhere[i] = iw_make(iw, true, true);
  }
  block_sum_store<3>(v, sh, part + blockIdx.x, gridDim.x);
}

// max_i |P(g)_i| (lbfgsbcpu.cpp:1389-1414); per-block maxima to pmax.
template <typename T>
__device__ inline T projected_gi(T xi, T gi, T li, T ui, int nb) {
  if (nb != 0) {
    if (gi < 0) {
      if (nb >= 2) {
        T a = xi - ui;
        gi = a > gi ? a : gi;
      }
    } else {
      if (nb <= 2) {
        T a = xi - li;
        gi = a > gi ? gi : a;
      }
    }
  }
  return gi;
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
template <typename T, bool FUSED>
__global__ void __launch_bounds__(kBlock)
k_pass_a(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
         const T* __restrict__ u, const int* __restrict__ nbd,
         uint8_t* __restrict__ iwhere, T* __restrict__ W, const T* __restrict__ d,
         DevScalars* ds, double* part, double* pmax, double* pmin, long long* pmini) {
  __shared__ double sh[kWavesPerBlock * PA_COUNT];
  if (ds->halt) return;
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  const T* __restrict__ xold = X[cur ^ 1];
  const T* __restrict__ gold = Gv[cur ^ 1];
  const int m = ds->m, col = ds->col, head = ds->head, itail = ds->itail;
  const bool upd = FUSED && ds->updatd;
  const T stp = (T)ds->stp;
  const bool stp1 = ds->stp == 1.0;
  const RingReads rr = ring_reads(head, m, col, upd ? itail : -1);

  double acc[PA_COUNT];
#pragma unroll
  for (int v = 0; v < PA_COUNT; v++) acc[v] = 0.0;
  double sbg = 0.0;
  double pny = 0.0, pns = 0.0;  // new pair's share of p (position col - 1)
  double bkmin = __builtin_inf();
  long long ibk = 0x7fffffffffffffffLL;

  // everything the pass reads for one element; the slot being replaced by
  // the new pair is not read
  struct In {
    T x, g, l, u, xo, go, dd;
    int nb;
    uint8_t iwb;
    T yv[kMaxM], sv[kMaxM];
  };
  auto load = [&](In& e, const long long i) __attribute__((always_inline)) {
    e.x = x[i];
    e.g = g[i];
    e.l = l[i];
    e.u = u[i];
    e.nb = nbd[i];
    e.iwb = iwhere[i];
    e.xo = FUSED ? xold[i] : (T)0;
    e.go = upd ? gold[i] : (T)0;
    e.dd = upd ? d[i] : (T)0;
    load_ring(W, i, rr, e.yv, e.sv);
  };
  auto keep = [&](const In& e) __attribute__((always_inline)) {
    keep_loads(e.x, e.g, e.l, e.u, e.nb, e.iwb, e.xo, e.go, e.dd, e.yv, e.sv);
  };
  auto element = [&](const long long i, In& e) __attribute__((always_inline)) {
    const T xi = e.x, gi = e.g, li = e.l, ui = e.u;
    const int nb = e.nb;
    const uint8_t iwb = e.iwb;
    T ys = 0, ss = 0;
    if (FUSED) {
      const T gp = projected_gi(xi, gi, li, ui, nb);
      sbg = fmax(sbg, (double)fabs(gp));
      const T xd = e.xo - xi;
      acc_prod(acc[PA_TF], xd, xd);
      if (upd) {
        ys = gi - e.go;
        ss = stp1 ? e.dd : e.dd * stp;
        acc_prod(acc[PA_RR], ys, ys);
        W[widx(i, itail)] = ys;
        W[widx(i, 8 + itail)] = ss;
      }
    }
    // Unused positions hold zeros (load_ring), and with an update the
    // newest position col - 1 (slot itail, not read) is zero too: the sums
    // run over all positions without guards, and the new pair's share of p
    // goes to pny / pns (added at position col - 1 after the loop).
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
        if (fabs(neggi) > 0) acc[PA_UNBND] = 1.0;
      }
    }
    iwhere[i] = iw_make(iw, iw_wasfree(iwb), iw_prevfree(iwb));
  };

  if constexpr (sizeof(T) == 4) {
    // FP32: two consecutive runs of 256 elements per block step, loads of
    // both issued before either is processed
    for (long long i0 = blockIdx.x * (long long)(2 * kBlock); i0 < n;
         i0 += (long long)gridDim.x * (2 * kBlock)) {
      const long long ia = i0 + threadIdx.x, ib = ia + kBlock;
      In ea, eb;
      load(ea, ia < n ? ia : n - 1);
      load(eb, ib < n ? ib : n - 1);
      keep(ea);
      keep(eb);
      if (ia < n) element(ia, ea);
      if (ib < n) element(ib, eb);
    }
  } else {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
      In e;
      load(e, i);
      keep(e);
      element(i, e);
    }
  }
  if (upd) {
#pragma unroll
    for (int j = 0; j < kMaxM; j++) {
      acc[PA_P + j] += j == col - 1 ? pny : 0.0;
      acc[PA_P + 8 + j] += j == col - 1 ? pns : 0.0;
    }
  }
  // PA_UNBND is an OR: keep it as a 0/1 value through the sum
  block_sum_store<PA_COUNT>(acc, sh, part + blockIdx.x, gridDim.x);
  if (FUSED) block_max_store(sbg, pmax + blockIdx.x);
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
// Signed entering/leaving Gram pairs: 36 YY (lower), 36 SS (lower), 64 SY.
__device__ inline void corr_pair(int t, int& cat, int& a, int& b) {
  if (t < 72) {
    cat = t < 36 ? 0 : 1;
    int q = t < 36 ? t : t - 36;
    // lower-triangular index q -> (a >= b)
    a = 0;
    while ((a + 1) * (a + 2) / 2 <= q) a++;
    b = q - a * (a + 1) / 2;
  } else {
    cat = 2;
    a = (t - 72) >> 3;  // S position
    b = (t - 72) & 7;   // Y position
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_p4(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
     const T* __restrict__ u, const int* __restrict__ nbd,
     uint8_t* __restrict__ iwhere, const T* __restrict__ W, T* __restrict__ z,
     T* __restrict__ r, DevScalars* ds, double* part) {
  __shared__ double sh[kWavesPerBlock * 52];
  __shared__ double s_corr[kWavesPerBlock][136];
  if (ds->halt) return;
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  const int m = ds->m, col = ds->col, head = ds->head;
  const bool shortcut = ds->shortcut;
  const int mode = ds->gcp_mode;
  const T tsum = (T)ds->gcp_tsum;
  const double t_last = ds->gcp_last_t;
  const long long i_last = ds->gcp_last_idx;
  const bool track = ds->iter > 0 && ds->cnstnd && !shortcut;
  const bool newcol = ds->updatd && col > 0;
  const double theta = ds->theta;
  const int upcl = ds->updatd ? col - 1 : col;
  const RingReads rr = ring_reads(head, m, col, -1);
  double mcy[kMaxM], mcs[kMaxM];
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    mcy[j] = j < col ? ds->mc[j] : 0.0;
    mcs[j] = j < col ? theta * ds->mc[col + j] : 0.0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

  // per-thread accumulators: 3 counts, 32 new-column sums, 16 W'r
  double acc[51];
#pragma unroll
  for (int v = 0; v < 51; v++) acc[v] = 0.0;
  // signed Gram pairs over entering (+1) / leaving (-1) variables: lane L of
  // every wave owns pairs L, L+64, L+128 (< 136); wa/wb are the W columns
  // (0..7 = Y slots, 8..15 = S slots) of the pair's two factors
  double corr[3] = {0.0, 0.0, 0.0};
  int wa[3], wb[3];
  bool pact[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int t = lane + 64 * q;
    int cat = 0, a = 0, b = 0;
    if (t < 136) corr_pair(t, cat, a, b);
    pact[q] = t < 136 && a < upcl && b < upcl;
    int sa = head + a, sb = head + b;
    if (sa >= m) sa -= m;
    if (sb >= m) sb -= m;
    wa[q] = cat == 0 ? sa : 8 + sa;
    wb[q] = cat == 1 ? 8 + sb : sb;
  }

  // one element: Cauchy point, free/entering/leaving flags, the new-column
  // Gram sums, r and W'r.  Returns +1 / -1 for an entering / leaving variable.
  auto element = [&](const long long i, const BulkIn<T>& e) __attribute__((always_inline)) -> int {
    int flag = 0;
    const T xi = e.x, gi = e.g, li = e.l, ui = e.u;
    const int nb = e.nb;
    const uint8_t iwb = e.iwb;
    const bool wasfree = iw_wasfree(iwb);
    int iw = iw_get(iwb);
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
      iwhere[i] = iw_make(iw, freenow, wasfree);
      z[i] = zi;
    }
    // Positions j >= col hold zeros (load_ring), so the sums below run over
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
      // r = -theta (z - x) - g + W (M c) over the free set
      // (lbfgsbcpu.cpp:847-867); 0 elsewhere, so W'r needs no branch
      double ri = -(double)gi;
      if (!shortcut) {
        ri = -theta * (double)(zi - xi) - (double)gi;
#pragma unroll
        for (int j = 0; j < kMaxM; j++) {
          ri = __builtin_fma((double)e.yv[j], mcy[j], ri);
          ri = __builtin_fma((double)e.sv[j], mcs[j], ri);
        }
      }
      const T rt = freenow ? (T)ri : (T)0;
      if (freenow) r[i] = rt;
#pragma unroll
      for (int j = 0; j < kMaxM; j++) {
        acc_prod(acc[P4_WV + j], e.yv[j], rt);
        acc_prod(acc[P4_WV + 8 + j], e.sv[j], rt);
      }
    }
    return flag;
  };

  // signed Gram over this wave's entering / leaving variables, one flagged
  // variable at a time in lane order (no block barrier)
  auto corr_step = [&](const long long i, const int flag) __attribute__((always_inline)) {
    if (!(track && col > 0)) return;
    unsigned long long fm = __ballot(flag != 0);
    if (fm && lane == 0) ds->p4_flagged = 1;
    while (fm) {
      const int src = __ffsll((long long)fm) - 1;
      fm &= fm - 1ull;
      const long long ei = __shfl(i, src, kWave);
      const double sg = (double)__shfl(flag, src, kWave);
      const double wrow = lane < 16 ? (double)W[widx(ei, lane)] : 0.0;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const double va = __shfl(wrow, wa[q], kWave);
        const double vb = __shfl(wrow, wb[q], kWave);
        if (pact[q]) corr[q] += sg * va * vb;
      }
    }
  };

  if constexpr (sizeof(T) == 4) {
    // FP32: each block takes two consecutive runs of 256 elements per step;
    // the loads of both are issued before either is processed, so every wave
    // has two elements' worth of requests in flight (the register footprint
    // leaves two waves per SIMD)
    for (long long i0 = blockIdx.x * (long long)(2 * kBlock); i0 < n;
         i0 += (long long)gridDim.x * (2 * kBlock)) {
      const long long ia = i0 + threadIdx.x, ib = ia + kBlock;
      BulkIn<T> ea, eb;
      load_bulk_in(ea, ia < n ? ia : n - 1, x, g, l, u, nbd, iwhere, W, rr);
      load_bulk_in(eb, ib < n ? ib : n - 1, x, g, l, u, nbd, iwhere, W, rr);
      keep_bulk_in(ea);
      keep_bulk_in(eb);
      corr_step(ia, ia < n ? element(ia, ea) : 0);
      corr_step(ib, ib < n ? element(ib, eb) : 0);
    }
  } else {
    // FP64: one element per step (two would not fit two waves per SIMD)
    for (long long i0 = blockIdx.x * (long long)kBlock; i0 < n;
         i0 += (long long)gridDim.x * kBlock) {
      const long long ia = i0 + threadIdx.x;
      BulkIn<T> ea;
      load_bulk_in(ea, ia < n ? ia : n - 1, x, g, l, u, nbd, iwhere, W, rr);
      keep_bulk_in(ea);
      corr_step(ia, ia < n ? element(ia, ea) : 0);
    }
  }
  block_sum_store<51>(acc, sh, part + blockIdx.x, gridDim.x);
  if (track && col > 0) {
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (lane + 64 * q < 136) s_corr[wid][lane + 64 * q] = corr[q];
    __syncthreads();
    for (int t = threadIdx.x; t < 136; t += blockDim.x) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kWavesPerBlock; w++) s += s_corr[w][t];
      part[(size_t)(P4_CORR + t) * gridDim.x + blockIdx.x] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// k_p5a: subspace Newton direction over the free variables and the
// step-to-box ratio (lbfgsbcpu.cpp:1472-1516).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_p5a(long long n, const T* __restrict__ l, const T* __restrict__ u,
      const int* __restrict__ nbd, const uint8_t* __restrict__ iwhere,
      const T* __restrict__ W, const T* __restrict__ z, const T* __restrict__ r,
      T* __restrict__ dsub, DevScalars* ds, double* pmin, long long* pmini) {
  if (ds->halt || ds->skip_subsm) return;
  const int m = ds->m, col = ds->col, head = ds->head;
  const bool shortcut = ds->shortcut;
  const double theta = ds->theta;
  double wy[kMaxM], wsv[kMaxM];
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    wy[j] = j < col ? ds->wv[j] / theta : 0.0;
    wsv[j] = j < col ? ds->wv[col + j] : 0.0;
  }
  const RingReads rr = ring_reads(head, m, col, -1);
  double amin = 1.0;
  long long ai = 0x7fffffffffffffffLL;
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    // every load of the element is issued before the first use, so a wave
    // keeps ~22 requests in flight instead of waiting on each W pair
    const uint8_t iwb = iwhere[i];
    const T ri = r[i], zi = z[i], li = l[i], ui = u[i];
    const int nb = nbd[i];
    T yv[kMaxM], sv[kMaxM];
    load_ring(W, i, rr, yv, sv);
    keep_loads(ri, zi, li, ui, nb, iwb, yv, sv);
    const bool freenow = shortcut || iw_get(iwb) <= 0;
    if (!freenow) continue;
    double di = (double)ri;
#pragma unroll
    for (int j = 0; j < kMaxM; j++)
      if (j < col) di += (double)yv[j] * wy[j] + (double)sv[j] * wsv[j];
    const T dk = (T)(di / theta);
    dsub[i] = dk;
    if (nb != 0) {
      double cand = 2.0;
      if (dk < 0 && nb <= 2) {
        const T t2 = li - zi;
        if (t2 >= 0)
          cand = 0.0;
        else if (dk < t2)
          cand = (double)(t2 / dk);
      } else if (dk > 0 && nb >= 2) {
        const T t2 = ui - zi;
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

// k_p5b: apply the backtrack (pinning the limiting variable), d = z - x,
// and the line-search set-up sums d'd, g'd and the nbd-aware max step.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_p5b(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
      const T* __restrict__ u, const int* __restrict__ nbd,
      const uint8_t* __restrict__ iwhere, T* __restrict__ z,
      const T* __restrict__ dsub, T* __restrict__ d, DevScalars* ds,
      double* part, double* pmin) {
  __shared__ double sh[kWavesPerBlock * P5_COUNT];
  if (ds->halt) return;
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  const bool sub = !ds->skip_subsm;
  const bool shortcut = ds->shortcut;
  const T alpha = (T)ds->alpha;
  const bool pin = ds->alpha < 1.0;
  const long long ibd = ds->ibd;
  const bool want_stpmx = ds->cnstnd && ds->iter > 0;
  const T big = (T)1.0e10;
  double acc[P5_COUNT] = {0.0, 0.0};
  double smin = 1.0e10;
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    // all loads first (see load_ring)
    const T xi = x[i], gi = g[i], li = l[i], ui = u[i];
    const T zr = shortcut ? (T)0 : z[i];
    const uint8_t iwb = iwhere[i];
    const T dl = sub ? dsub[i] : (T)0;
    const int nb = nbd[i];
    keep_loads(xi, gi, li, ui, zr, iwb, dl, nb);
    T zi = shortcut ? xi : zr;
    if (sub && (shortcut || iw_get(iwb) <= 0)) {
      T dk = dl;
      if (pin && i == ibd) {
        if (dk > 0) {
          zi = ui;
          dk = 0;
        } else if (dk < 0) {
          zi = li;
          dk = 0;
        }
      }
      zi = zi + alpha * dk;
      z[i] = zi;
    } else if (shortcut) {
      z[i] = zi;
    }
    const T di = zi - xi;
    d[i] = di;
    acc[0] += (double)di * (double)di;
    acc[1] += (double)gi * (double)di;
    if (want_stpmx) {
      if (nb != 0) {
        if (di < 0 && nb <= 2) {
          const T a2 = li - xi;
          if (a2 >= 0)
            smin = 0.0;
          else if (di * big < a2)
            smin = fmin(smin, (double)(a2 / di));
        } else if (di > 0 && nb >= 2) {
          const T a2 = ui - xi;
          if (a2 <= 0)
            smin = 0.0;
          else if (di * big > a2)
            smin = fmin(smin, (double)(a2 / di));
        }
      }
    }
  }
  block_sum_store<P5_COUNT>(acc, sh, part + blockIdx.x, gridDim.x);
  smin = -wave_max(-smin);
  __shared__ double shm[kWavesPerBlock];
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = smin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = shm[0];
    for (int w = 1; w < kWavesPerBlock; w++) v = fmin(v, shm[w]);
    pmin[blockIdx.x] = v;
  }
}

// Trial point x_b = stp*d + x_a, or the subspace point when stp == 1
// (lbfgsbcpu.cpp:1329-1335).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_trial(long long n, Pair<T> X, const T* __restrict__ z, const T* __restrict__ d,
        const DevScalars* ds) {
  if (!ds->write_trial) return;
  const T* __restrict__ xa = X[ds->trial_src];
  T* __restrict__ xb = X[ds->trial_dst];
  const double stp = ds->stp;
  const bool one = stp == 1.0;
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    if (one)
      xb[i] = z[i];
    else
      xb[i] = (T)(stp * (double)d[i] + (double)xa[i]);
  }
}

// g_b . d after an evaluation (lbfgsbcpu.cpp:1312).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gd(long long n, Pair<T> Gv, const T* __restrict__ d, const DevScalars* ds,
     double* part) {
  __shared__ double sh[kWavesPerBlock];
  const int cur = ds->phase & 1;
  const T* __restrict__ gb = Gv[cur ^ 1];
  double acc[1] = {0.0};
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    acc[0] += (double)gb[i] * (double)d[i];
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
                   Pair<T> Gv, const T* l, const T* u, const int* nbd,
                   uint8_t* iwhere, T* W, const T* d, DevScalars* ds, double* part,
                   double* pmax, double* pmin, long long* pmini) {
  if (fused)
    hipLaunchKernelGGL((k_pass_a<T, true>), dim3(G), dim3(kBlock), 0, s, n, X, Gv,
                       l, u, nbd, iwhere, W, d, ds, part, pmax, pmin, pmini);
  else
    hipLaunchKernelGGL((k_pass_a<T, false>), dim3(G), dim3(kBlock), 0, s, n, X,
                       Gv, l, u, nbd, iwhere, W, d, ds, part, pmax, pmin, pmini);
}
template <typename T>
void launch_p4(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
               const T* l, const T* u, const int* nbd, uint8_t* iwhere,
               const T* W, T* z, T* r, DevScalars* ds, double* part) {
  hipLaunchKernelGGL(k_p4<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, nbd,
                     iwhere, W, z, r, ds, part);
}
template <typename T>
void launch_p5a(hipStream_t s, int G, long long n, const T* l, const T* u,
                const int* nbd, const uint8_t* iwhere, const T* W, const T* z,
                const T* r, T* dsub, DevScalars* ds, double* pmin,
                long long* pmini) {
  hipLaunchKernelGGL(k_p5a<T>, dim3(G), dim3(kBlock), 0, s, n, l, u, nbd, iwhere,
                     W, z, r, dsub, ds, pmin, pmini);
}
template <typename T>
void launch_p5b(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
                const T* l, const T* u, const int* nbd, const uint8_t* iwhere,
                T* z, const T* dsub, T* d, DevScalars* ds, double* part,
                double* pmin) {
  hipLaunchKernelGGL(k_p5b<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, nbd,
                     iwhere, z, dsub, d, ds, part, pmin);
}
template <typename T>
void launch_trial(hipStream_t s, int G, long long n, Pair<T> X, const T* z,
                  const T* d, const DevScalars* ds) {
  hipLaunchKernelGGL(k_trial<T>, dim3(G), dim3(kBlock), 0, s, n, X, z, d, ds);
}
template <typename T>
void launch_gd(hipStream_t s, int G, long long n, Pair<T> Gv, const T* d,
               const DevScalars* ds, double* part) {
  hipLaunchKernelGGL(k_gd<T>, dim3(G), dim3(kBlock), 0, s, n, Gv, d, ds, part);
}

#define LBFGSB_INST_BULK(T)                                                         \
  template void launch_errclb<T>(hipStream_t, int, long long, const int*, const T*, \
                                 const T*, double*);                                \
  template void launch_active<T>(hipStream_t, int, long long, T*, const T*,         \
                                 const T*, const int*, uint8_t*, double*);          \
  template void launch_projgr<T>(hipStream_t, int, long long, const T*, const T*,   \
                                 const T*, const T*, const int*, double*);          \
  template void launch_pass_a<T>(hipStream_t, int, bool, long long, Pair<T>,      \
                                 Pair<T>, const T*, const T*, const int*,         \
                                 uint8_t*, T*, const T*, DevScalars*, double*,      \
                                 double*, double*, long long*);                     \
  template void launch_p4<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,     \
                             const T*, const T*, const int*, uint8_t*, const T*,    \
                             T*, T*, DevScalars*, double*);                         \
  template void launch_p5a<T>(hipStream_t, int, long long, const T*, const T*,      \
                              const int*, const uint8_t*, const T*, const T*,       \
                              const T*, T*, DevScalars*, double*, long long*);      \
  template void launch_p5b<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,    \
                              const T*, const T*, const int*, const uint8_t*, T*,   \
                              const T*, T*, DevScalars*, double*, double*);         \
  template void launch_trial<T>(hipStream_t, int, long long, Pair<T>, const T*,   \
                                const T*, const DevScalars*);                       \
  template void launch_gd<T>(hipStream_t, int, long long, Pair<T>, const T*,      \
                             const DevScalars*, double*);

LBFGSB_INST_BULK(float)
LBFGSB_INST_BULK(double)

}  // namespace lbfgsb
