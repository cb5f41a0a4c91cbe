// Device side of the breakpoint search of the generalized Cauchy point:
// candidate order, per-candidate contributions, the closed-form stop test and
// the conclusion of a round.  Used by the grid kernels of gcp.hip and by the
// per-problem batched kernel.  The method is described at the top of
// gcp.hip.
#pragma once

#include "common.hpp"
#include "control.hpp"
#include "dense.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

static constexpr int kSortTile = 2048;
static constexpr int kNV = 33;
static constexpr double kMachEpsG = 5e-16;

__device__ inline bool cand_less(const Cand& a, const Cand& b) {
  return a.t < b.t || (a.t == b.t && a.idx < b.idx);
}

// Whether component q of a state vector [P_y P_s Q_y Q_s D] can be nonzero
// with col pairs in the memory: the positions j < col of each block, and D.
// The others are exact zeros, so their sums and scans are skipped.
__device__ __forceinline__ bool nv_used(int q, int col) { return q == kNV - 1 || (q & 7) < col; }

// Contribution of candidate c: [d w (2col) ; t d w (2col) ; d^2].
template <typename T>
__device__ inline void contribution(const Cand& c, const T* __restrict__ g,
                                    const T* __restrict__ W, const RingReads& rr,
                                    double theta, double (&v)[kNV]) {
  const long long i = c.idx;
  // g_i and the used positions of the W row (a gather: every entry is its
  // own cache line) are loaded before any use; positions j >= col are zero
  // and not read
  T yv[kMaxM], sv[kMaxM];
  const T gi = g[i];
  const T* __restrict__ row = W + widx((size_t)i, 0);
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    yv[j] = (T)0;
    sv[j] = (T)0;
    if (rr.use[j]) {
      yv[j] = row[rr.rs[j] * kTile];
      sv[j] = row[(kWPhys + rr.rs[j]) * kTile];
    }
  }
  const double d = -(double)gi;
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    const double wy = (double)yv[j];
    const double ws = theta * (double)sv[j];
    v[j] = d * wy;
    v[8 + j] = d * ws;
    v[16 + j] = c.t * d * wy;
    v[24 + j] = c.t * d * ws;
  }
  v[32] = d * d;
}

// p at the start of the search and the middle matrix M in the padded
// layout of the breakpoint sums (positions 0..7 Y, 8..15 S, zero beyond
// col), staged in LDS by every thread of the block.
struct GcpConst {
  double pp[16];
  double Mp[256];
};
__device__ inline void stage_gcp_const(const DevScalars* ds, GcpConst& gc) {
  const int col = ds->col;
  for (int q = threadIdx.x; q < 256; q += blockDim.x) gc.Mp[q] = ds->Mx[q];
  if (threadIdx.x < 16) {
    const int q = threadIdx.x;
    const int j = q & 7;
    gc.pp[q] = j < col ? ds->p[(q < 8 ? 0 : col) + j] : 0.0;
  }
  __syncthreads();
}

// Layout of a state vector: [P_y(8) P_s(8) Q_y(8) Q_s(8) D].

// p'Mp and q'Mp at the state `st` (p = pp - P, q = Q) over the first C
// positions of each half: the positions from col on hold exact zeros, so
// leaving them out (C >= col) leaves every sum unchanged; the loops have
// compile-time indices, so the temporaries stay in registers.
template <int C>
__device__ __forceinline__ void stop_quad(const GcpConst& gc, const double (&st)[kNV], double& pmp,
                                          double& qmp) {
  double p[2 * C], mp[2 * C];
#pragma unroll
  for (int a = 0; a < 2 * C; a++) {
    const int q = a < C ? a : kMaxM + a - C;
    p[a] = gc.pp[q] - st[q];
  }
#pragma unroll
  for (int ra = 0; ra < 2 * C; ra++) {
    const int r = ra < C ? ra : kMaxM + ra - C;
    double s = 0.0;
#pragma unroll
    for (int ka = 0; ka < 2 * C; ka++) {
      const int k = ka < C ? ka : kMaxM + ka - C;
      s += gc.Mp[r * 16 + k] * p[ka];
    }
    mp[ra] = s;
  }
  pmp = 0.0;
  qmp = 0.0;
#pragma unroll
  for (int j = 0; j < C; j++) {
    pmp += p[j] * mp[j] + p[C + j] * mp[C + j];
    qmp += st[16 + j] * mp[j] + st[24 + j] * mp[C + j];
  }
}

// Stop test at the state `st` (passed breakpoints before t_next).
__device__ inline bool stop_test(const DevScalars* ds, const GcpConst& gc,
                                 const double (&st)[kNV], double t_prev, double t_next,
                                 bool floor_on, double (&out)[3]) {
  const double theta = ds->theta;
  const int col = ds->col;
  double pmp, qmp;
  if (col <= 2)
    stop_quad<2>(gc, st, pmp, qmp);
  else if (col <= 4)
    stop_quad<4>(gc, st, pmp, qmp);
  else
    stop_quad<kMaxM>(gc, st, pmp, qmp);
  const double D = ds->gcp_D0 - st[32];
  const double f2u = theta * D - pmp;
  const double f1 = -D + t_prev * f2u - qmp;
  double f2 = f2u;
  if (floor_on) {
    const double fl = kMachEpsG * ds->gcp_f2org;
    if (fl > f2) f2 = fl;
  }
  const double dtm = -f1 / f2;
  out[0] = dtm;
  out[1] = f1;
  out[2] = f2;
  return dtm < t_next - t_prev;
}

// The bound below which the next pass A lists breakpoints: gcp_near_grow
// (8 to start with) times the upper end of the t-range this search ended in,
// none when that range is unbounded.
static constexpr double kNearGrow0 = 8.0;
__device__ __forceinline__ void near_next(DevScalars* ds) {
  const double gr = ds->gcp_near_grow > 0.0 ? ds->gcp_near_grow : kNearGrow0;
  const double t = ds->gcp_tau_hi * gr;
  ds->gcp_near_tau = t < 1e300 ? t : 0.0;
}

__device__ __forceinline__ void gcp_done(DevScalars* ds) {
  near_next(ds);
  const int col = ds->col;
  if (col > 0 && d_bmv(ds->sy, ds->wt, col, ds->c, ds->mc) != 0) ds->internalinfo = -8;
  ds->nintol += ds->nint;
  ds->gcp_pending = 0;
  ds->status = ST_RUNNING;
  ds->halt = 0;
}

// Conclusion of a round (thread 0).  stop_pos = first candidate at which the
// search stops (its exclusive state in `st`), or -1 when every candidate
// passed (then `st` is the inclusive total).  last_state = exclusive state
// before the last candidate.
__device__ __forceinline__ void round_conclude(DevScalars* ds, const GcpConst& gc, const Cand* cand, long long K,
                               long long stop_pos, const double* stv, const double* last_state) {
  const int col = ds->col;
  // the state stays where the caller keeps it (LDS or global), read in place
  const double(&st)[kNV] = *reinterpret_cast<const double(*)[kNV]>(stv);
  double tk, dtm = 0.0;
  long long ik, kpass;
  if (stop_pos >= 0) {
    const long long pos = stop_pos;
    tk = pos > 0 ? cand[pos - 1].t : ds->gcp_t_prev;
    ik = pos > 0 ? cand[pos - 1].idx : ds->gcp_idx_prev;
    kpass = ds->gcp_kpassed + pos;
    double out[3];
    stop_test(ds, gc, st, tk, cand[pos].t, kpass > 0, out);
    dtm = out[0];
  } else {
    for (int q = 0; q < kNV; q++) ds->gcp_carry[q] = st[q];
    ds->gcp_kpassed += K;
    if (K > 0) {
      ds->gcp_t_prev = cand[K - 1].t;
      ds->gcp_idx_prev = cand[K - 1].idx;
    }
    if (ds->gcp_kpassed < ds->gcp_nbreak) {
      // next round: a 4x larger t-range; sized by the host
      ds->gcp_round++;
      ds->gcp_ranges++;
      ds->gcp_need_select = 1;
      ds->gcp_tau_lo = ds->gcp_tau_hi;
      const double nt = ds->gcp_tau_hi * 4.0;
      ds->gcp_tau_hi = (ds->gcp_round >= 24 || !(nt < 1e300)) ? __builtin_inf() : nt;
      ds->cand_count = 0;
      ds->status = ST_GCP_ROUNDS;
      ds->halt = 1;
      return;
    }
    kpass = ds->gcp_kpassed;
    tk = ds->gcp_t_prev;
    ik = ds->gcp_idx_prev;
    if (ds->gcp_nbreak == ds->n) {
      // every variable was a breakpoint and all were passed: the reference
      // returns before updating p with the last one (lbfgsbcpu.cpp:763-769);
      // c = t_K p_{K-1} + Q_{K-1}
      for (int j = 0; j < col; j++) {
        const double py = ds->p[j] - last_state[j];
        const double ps = ds->p[col + j] - last_state[8 + j];
        ds->c[j] = tk * py + last_state[16 + j];
        ds->c[col + j] = tk * ps + last_state[24 + j];
        ds->p[j] = py;
        ds->p[col + j] = ps;
      }
      ds->nint = (int)ds->gcp_nbreak;
      ds->gcp_mode = 3;
      ds->gcp_tsum = tk;
      ds->gcp_last_t = tk;
      ds->gcp_last_idx = ik;
      gcp_done(ds);
      return;
    }
    double out[3];
    stop_test(ds, gc, st, tk, tk, true, out);
    dtm = ds->gcp_unbnd ? out[0] : 0.0;
  }
  if (dtm <= 0) dtm = 0;
  for (int j = 0; j < col; j++) {
    const double py = ds->p[j] - st[j];
    const double ps = ds->p[col + j] - st[8 + j];
    ds->p[j] = py;
    ds->p[col + j] = ps;
    ds->c[j] = (tk + dtm) * py + st[16 + j];
    ds->c[col + j] = (tk + dtm) * ps + st[24 + j];
  }
  ds->gcp_dtm = dtm;
  ds->gcp_tsum = tk + dtm;
  ds->gcp_last_t = tk;
  ds->gcp_last_idx = ik;
  ds->gcp_mode = kpass > 0 ? 1 : 0;
  ds->nint = (int)(kpass + 1);
  ds->gcp_passed_total += kpass;
  gcp_done(ds);
}

// ---------------------------------------------------------------------------
// A whole round in one workgroup (any blockDim that is a multiple of 64, up
// to 1024): sort, then chunks of blockDim candidates scanned in order,
// stopping at the first chunk that holds the stop.
static constexpr int kSmallRound = 2048;  // candidates one workgroup sorts in LDS
static constexpr int kMaxRoundWaves = 16;

struct RoundLds {
  Cand sc[kSmallRound];
  double swave[kMaxRoundWaves * kNV];
  double scarry[kNV];
  double sstop_state[kNV];
  double slast[kNV];
  long long sstop;
  GcpConst gc;
};

// Bitonic sort of sc[0 .. P) in LDS, P a power of two.  Pair p of the stage
// with distance j = 2^lj is (i, i + j), i the p-th index whose bit lj is
// clear (shifts: no integer division).  The pairs a wave takes (p in
// 64-aligned runs) lie in 128-element blocks of its own for every stage with
// j <= 32, so such a stage waits for the workgroup only when the stage
// before it had j >= 64; between two of them the wave's own LDS operations,
// which complete in order, suffice.
__device__ inline void block_bitonic(Cand* sc, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    const int lj0 = __ffs(k) - 2;
    for (int lj = lj0; lj >= 0; lj--) {
      const int j = 1 << lj;
      if (j >= 64 || (j == 32 && lj != lj0))
        __syncthreads();
      else
        __builtin_amdgcn_wave_barrier();
      for (int p = threadIdx.x; p < P / 2; p += blockDim.x) {
        const int i = ((p >> lj) << (lj + 1)) | (p & (j - 1)), l = i + j;
        const bool up = (i & k) == 0;
        if (cand_less(sc[l], sc[i]) == up) {
          Cand tmp = sc[i];
          sc[i] = sc[l];
          sc[l] = tmp;
        }
      }
    }
  }
  __syncthreads();
}

// Sort K <= 64 candidates with one wave: lane-to-lane bitonic network in
// registers, then into LDS.  The other waves only take part in the barrier.
__device__ inline void wave_sort_small(const Cand* cand, long long K, Cand* sc) {
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    Cand c;
    if (lane < K) {
      c = cand[lane];
    } else {
      c.t = __builtin_inf();
      c.idx = 0x7fffffff;
      c.pad = 0;
    }
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        Cand o;
        o.t = __shfl_xor(c.t, j, kWave);
        o.idx = __shfl_xor(c.idx, j, kWave);
        o.pad = 0;
        // pair (i, i + j): ascending blocks keep the smaller at i
        const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
        const bool take = keep_min ? cand_less(o, c) : cand_less(c, o);
        if (take) c = o;
      }
    }
    sc[lane] = c;
  }
  __syncthreads();
}

// Load K <= kSmallRound candidates into LDS, padded to a power of two, and
// sort them.
__device__ inline void block_sort_small(const Cand* cand, long long K, Cand* sc) {
  if (K <= kWave) {
    wave_sort_small(cand, K, sc);
    return;
  }
  int P = 64;
  while (P < K) P <<= 1;
  for (int j = threadIdx.x; j < P; j += blockDim.x) {
    if (j < K) {
      sc[j] = cand[j];
    } else {
      sc[j].t = __builtin_inf();
      sc[j].idx = 0x7fffffff;
      sc[j].pad = 0;
    }
  }
  __syncthreads();
  block_bitonic(sc, P);
}

// Sort K candidates of a (global) with one workgroup: kSmallRound-long runs
// in LDS, then merge-path passes between a and b.  Returns the buffer that
// holds the sorted sequence.
__device__ LBFGSB_COLD Cand* block_sort_large(Cand* a, Cand* b, long long K, Cand* sc) {
  for (long long base = 0; base < K; base += kSmallRound) {
    const long long len = K - base < kSmallRound ? K - base : kSmallRound;
    block_sort_small(a + base, len, sc);
    for (int j = threadIdx.x; j < len; j += blockDim.x) a[base + j] = sc[j];
    __syncthreads();
  }
  Cand* src = a;
  Cand* dst = b;
  for (long long width = kSmallRound; width < K; width <<= 1) {
    const long long per = (K + blockDim.x - 1) / blockDim.x;
    const long long o0 = (long long)threadIdx.x * per;
    long long o = o0;
    while (o < K && o < o0 + per) {
      // output position o lies in the pair of runs starting at base
      const long long base = (o / (2 * width)) * 2 * width;
      const long long a0 = base, a1 = base + width < K ? base + width : K;
      const long long b0 = a1, b1 = base + 2 * width < K ? base + 2 * width : K;
      const long long la = a1 - a0, lb = b1 - b0;
      const long long diag = o - base;
      long long lo = diag - lb > 0 ? diag - lb : 0, hi = diag < la ? diag : la;
      while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (cand_less(src[a0 + mid], src[b0 + diag - 1 - mid]))
          lo = mid + 1;
        else
          hi = mid;
      }
      long long i = lo, j = diag - lo;
      const long long oend = (o0 + per < b1) ? o0 + per : b1;
      for (; o < oend; o++) {
        bool takeA;
        if (i >= la)
          takeA = false;
        else if (j >= lb)
          takeA = true;
        else
          takeA = !cand_less(src[b0 + j], src[a0 + i]);
        dst[o] = takeA ? src[a0 + i++] : src[b0 + j++];
      }
    }
    __syncthreads();
    Cand* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

// Scan sorted candidates cands[c_lo .. c_hi) of a round of K (LDS or global)
// in chunks of blockDim: prefix sums of the contributions from the state in
// L.scarry and, when `test`, the stop test at every position, ending at the
// first chunk that holds a stop.  On return L.sstop is the first stopping
// position (or the largest long long) with its exclusive state in
// L.sstop_state, L.scarry is the inclusive state after the last position
// scanned, and L.slast the exclusive state of position K - 1 if scanned.
// L.gc must be staged.
template <typename T>
__device__ __forceinline__ void block_scan_range(const DevScalars* ds, RoundLds& L, const Cand* cands,
                                        long long c_lo, long long c_hi, long long K, bool test,
                                        const T* g, const T* W) {
  const RingReads rr = ring_reads(ds->head, ds->m, ds->col, -1, ds->pslot_x);
  const int col = ds->col;
  const double theta = ds->theta;
  const long long kpassed = ds->gcp_kpassed;
  const double t_prev0 = ds->gcp_t_prev;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  if (threadIdx.x == 0) L.sstop = 0x7fffffffffffffffLL;
  __syncthreads();
  for (long long c0 = c_lo; c0 < c_hi; c0 += blockDim.x) {
    const long long pos = c0 + threadIdx.x;
    double v[kNV], ex[kNV];
#pragma unroll
    for (int q = 0; q < kNV; q++) v[q] = 0.0;
    Cand c;
    c.t = 0;
    c.idx = 0;
    c.pad = 0;
    if (pos < c_hi) {
      c = cands[pos];
      contribution(c, g, W, rr, theta, v);
    }
#pragma unroll
    for (int q = 0; q < kNV; q++) {
      double incl = v[q];
      if (nv_used(q, col)) incl = wave_scan_incl(incl);
      ex[q] = incl - v[q];
      if (lane == 63) L.swave[wid * kNV + q] = incl;
    }
    __syncthreads();
    // each component's wave offsets (the carried state plus the totals of the
    // waves before, added in wave order) in place of the wave totals, by one
    // thread per component, and the chunk's total into the carry: the other
    // threads then read their offsets independently
    if (threadIdx.x < kNV) {
      const int q = threadIdx.x;
      double run = L.scarry[q];
      for (int w = 0; w < nw; w++) {
        const double t = L.swave[w * kNV + q];
        L.swave[w * kNV + q] = run;
        run += t;
      }
      L.scarry[q] = run;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kNV; q++) ex[q] += L.swave[wid * kNV + q];
    bool stop = false;
    if (test && pos < c_hi && kpassed + pos > 0) {
      const double tp = pos > 0 ? cands[pos - 1].t : t_prev0;
      double out[3];
      stop = stop_test(ds, L.gc, ex, tp, c.t, true, out);
    }
    if (stop) atomicMin(&L.sstop, pos);
    if (pos == K - 1)
      for (int q = 0; q < kNV; q++) L.slast[q] = ex[q];
    __syncthreads();
    if (L.sstop != 0x7fffffffffffffffLL) {
      if (pos == L.sstop)
        for (int q = 0; q < kNV; q++) L.sstop_state[q] = ex[q];
      break;
    }
  }
  __syncthreads();
}

// A whole round's scan in one workgroup and its conclusion by thread 0.
template <typename T>
__device__ inline void block_round_scan(DevScalars* ds, RoundLds& L, const Cand* cands,
                                        long long K, const T* g, const T* W) {
  if (threadIdx.x < kNV) L.scarry[threadIdx.x] = ds->gcp_carry[threadIdx.x];
  block_scan_range(ds, L, cands, 0, K, K, true, g, W);
  if (threadIdx.x == 0) {
    const bool stopped = L.sstop != 0x7fffffffffffffffLL;
    round_conclude(ds, L.gc, cands, K, stopped ? L.sstop : -1,
                   stopped ? L.sstop_state : L.scarry, L.slast);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// A round of more than kSmallRound candidates in one workgroup, without a
// full sort: the candidates are counted into kBuckets buckets by their t's
// bit pattern (monotone in t >= 0, so equal t share a bucket and every
// bucket holds a t-interval), scattered into b in bucket order, and then
// taken as consecutive groups of whole buckets of at most kSmallRound
// candidates: each group is sorted in LDS and scanned from the carried
// state, and the first group that holds the stop concludes the round (the
// groups after it are never sorted).  Group by group this is the scan of the
// fully sorted round: the groups' t-intervals are disjoint and ascending, and
// between groups the carried state, the passed count and the last passed
// breakpoint advance as the scan's positions do.  Returns false, having
// changed nothing, when one bucket holds more than kSmallRound candidates
// (the caller sorts the round whole).
static constexpr int kBuckets = 8192;

struct BucketLds {
  unsigned cnt[kBuckets];  // counts, then (scanned) each bucket's end in b
  unsigned long long kmin, kmax;
  unsigned big;
  unsigned gmax;  // the current group's largest bucket
};

// A group whose buckets all hold at most this many candidates is sorted one
// bucket per thread (insertion sort in LDS) instead of by the bitonic
// network over the whole group.
static constexpr unsigned kBucketInsert = 32;

__device__ __forceinline__ unsigned long long cand_key(double t) {
  return (unsigned long long)__double_as_longlong(t);  // t >= 0: ordered as t
}

template <typename T>
__device__ LBFGSB_COLD bool block_round_bucketed(DevScalars* ds, RoundLds& L, BucketLds& B,
                                            const Cand* __restrict__ a, Cand* __restrict__ b,
                                            long long K, const T* g, const T* W) {
  constexpr int U = 4;  // candidates per thread per step, loads issued together
  const long long nt = blockDim.x;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) {
    B.kmin = ~0ull;
    B.kmax = 0ull;
    B.big = 0u;
  }
  for (int q = threadIdx.x; q < kBuckets; q += blockDim.x) B.cnt[q] = 0u;
  __syncthreads();
  // the key range
  unsigned long long kmn = ~0ull, kmx = 0ull;
  for (long long i0 = threadIdx.x; i0 < K; i0 += U * nt) {
    double tv[U];
#pragma unroll
    for (int u = 0; u < U; u++) tv[u] = i0 + u * nt < K ? a[i0 + u * nt].t : -1.0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (tv[u] >= 0.0) {
        const unsigned long long k = cand_key(tv[u]);
        kmn = k < kmn ? k : kmn;
        kmx = k > kmx ? k : kmx;
      }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long on = __shfl_xor(kmn, o, kWave), ox = __shfl_xor(kmx, o, kWave);
    kmn = on < kmn ? on : kmn;
    kmx = ox > kmx ? ox : kmx;
  }
  if (lane == 0) {
    atomicMin(&B.kmin, kmn);
    atomicMax(&B.kmax, kmx);
  }
  __syncthreads();
  const unsigned long long kmin = B.kmin, span = B.kmax - B.kmin;
  int shift = 0;
  while (shift < 64 && (span >> shift) >= (unsigned long long)kBuckets) shift++;
  // counts
  for (long long i0 = threadIdx.x; i0 < K; i0 += U * nt) {
    double tv[U];
#pragma unroll
    for (int u = 0; u < U; u++) tv[u] = i0 + u * nt < K ? a[i0 + u * nt].t : -1.0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (tv[u] >= 0.0) atomicAdd(&B.cnt[(cand_key(tv[u]) - kmin) >> shift], 1u);
  }
  __syncthreads();
  // a bucket too large for one group: the whole sort instead
  for (int q = threadIdx.x; q < kBuckets; q += blockDim.x)
    if (B.cnt[q] > (unsigned)kSmallRound) B.big = 1u;
  __syncthreads();
  if (B.big) return false;
  // exclusive scan of the counts in place (each thread a contiguous run)
  {
    const int per = (kBuckets + (int)blockDim.x - 1) / (int)blockDim.x;
    const int q0 = threadIdx.x * per;
    unsigned run = 0;
    for (int q = q0; q < q0 + per && q < kBuckets; q++) run += B.cnt[q];
    unsigned incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned up = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += up;
    }
    unsigned* wsum = (unsigned*)L.swave;  // one word per wave, free until the scans
    if (lane == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    unsigned base = incl - run;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) base += wsum[w];
    __syncthreads();
    for (int q = q0; q < q0 + per && q < kBuckets; q++) {
      const unsigned c = B.cnt[q];
      B.cnt[q] = base;
      base += c;
    }
  }
  __syncthreads();
  // scatter in bucket order; cnt[q] ends as bucket q's end
  for (long long i0 = threadIdx.x; i0 < K; i0 += U * nt) {
    Cand c[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      c[u].t = -1.0;
      if (i0 + u * nt < K) c[u] = a[i0 + u * nt];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (c[u].t >= 0.0) b[atomicAdd(&B.cnt[(cand_key(c[u].t) - kmin) >> shift], 1u)] = c[u];
  }
  __syncthreads();
  // groups of whole buckets, in order
  unsigned start = 0;
  int b0 = 0;
  while (b0 < kBuckets) {
    // the last bucket b >= b0 that still ends within kSmallRound of start
    // (bucket ends ascend; bucket b0 alone always fits)
    int lo = b0, hi = kBuckets - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (B.cnt[mid] - start <= (unsigned)kSmallRound)
        lo = mid;
      else
        hi = mid - 1;
    }
    const unsigned end = B.cnt[lo];
    const long long len = (long long)(end - start);
    const bool last = lo == kBuckets - 1;
    if (len > 0) {
      if (threadIdx.x == 0) B.gmax = 0u;
      for (long long j = threadIdx.x; j < len; j += nt) L.sc[j] = b[start + j];
      __syncthreads();
      // bucket q of the group spans [end of bucket q - 1, end of q) of b
      unsigned mx = 0u;
      for (int q = b0 + (int)threadIdx.x; q <= lo; q += (int)nt) {
        const unsigned sz = B.cnt[q] - (q > 0 ? B.cnt[q - 1] : 0u);
        mx = sz > mx ? sz : mx;
      }
      if (mx) atomicMax(&B.gmax, mx);
      __syncthreads();
      if (B.gmax <= kBucketInsert) {
        // the buckets' t-ranges are disjoint and ascending, so sorting each
        // bucket in place sorts the group
        for (int q = b0 + (int)threadIdx.x; q <= lo; q += (int)nt) {
          const int a = (int)((q > 0 ? B.cnt[q - 1] : 0u) - start);
          const int e = (int)(B.cnt[q] - start);
          for (int x = a + 1; x < e; x++) {
            const Cand key = L.sc[x];
            int y = x - 1;
            while (y >= a && cand_less(key, L.sc[y])) {
              L.sc[y + 1] = L.sc[y];
              y--;
            }
            L.sc[y + 1] = key;
          }
        }
        __syncthreads();
      } else {
        block_sort_small(L.sc, len, L.sc);
      }
      if (threadIdx.x < kNV) L.scarry[threadIdx.x] = ds->gcp_carry[threadIdx.x];
      block_scan_range(ds, L, L.sc, 0, len, len, true, g, W);
      const bool stopped = L.sstop != 0x7fffffffffffffffLL;
      if (stopped || last || end >= (unsigned)K) {
        if (threadIdx.x == 0)
          round_conclude(ds, L.gc, L.sc, len, stopped ? L.sstop : -1,
                         stopped ? L.sstop_state : L.scarry, L.slast);
        __syncthreads();
        return true;
      }
      // every candidate of the group passed: carry on from its end
      if (threadIdx.x == 0) {
        for (int q = 0; q < kNV; q++) ds->gcp_carry[q] = L.scarry[q];
        ds->gcp_kpassed += len;
        ds->gcp_t_prev = L.sc[len - 1].t;
        ds->gcp_idx_prev = L.sc[len - 1].idx;
      }
      __syncthreads();
    }
    start = end;
    b0 = lo + 1;
  }
  return true;
}

}  // namespace lbfgsb
