// Breakpoint search of the generalized Cauchy point when at least one
// breakpoint is passed (lbfgsbcpu.cpp:725-807).
//
// The reference walks the breakpoints in increasing t with a heap and updates
// f1, f2, c, p after each one.  After passing the k smallest breakpoints
// (t_1 <= ... <= t_k, direction components d_j = -g_j, rows
// w_j = [Y_j ; theta S_j]) the same quantities have a closed form:
//   p_k  = p_0 - P_k,          P_k = sum_{j<=k} d_j w_j
//   c_k  = t_k p_k + Q_k,      Q_k = sum_{j<=k} t_j d_j w_j
//   D_k  = D_0 - sum_{j<=k} d_j^2
//   f2_k = theta D_k - p_k' M p_k
//   f1_k = -D_k + t_k f2_k - Q_k' M p_k
// and the search stops before breakpoint k+1 when -f1_k/f2_k < t_{k+1}-t_k.
// So the walk becomes rounds of: select the breakpoints with t in a range
// [tau_lo, tau_hi), sort them by (t, index), prefix-sum the
// (4 col + 1)-vectors [d w, t d w, d^2], and test every k in parallel; the
// first k that stops ends the search.  The first range is [0, 2 dtm_0) with
// dtm_0 the minimiser of the model before any breakpoint; later ranges grow
// by 4x.  The candidates are recomputed from x, g, l, u (no breakpoint list
// is kept by the hot pass), and round 0 runs entirely on the device with its
// size read on the device when one workgroup can sort it (2048 candidates);
// larger or later rounds are
// sized by the host.
#include "common.hpp"
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

// Same classification as the bulk passes (lbfgsbcpu.cpp:634-700).
template <typename T>
__device__ inline bool breakpoint_of(T x, T g, T l, T u, int nb, int iw, T& t) {
  const T neggi = -g;
  if (iw != 0 && iw != -1) return false;
  if (nb <= 2 && nb != 0 && neggi < 0) {
    t = (x - l) / (-neggi);
    return true;
  }
  if (nb >= 2 && neggi > 0) {
    t = (u - x) / neggi;
    return true;
  }
  return false;
}

// Candidates t in [tau_lo, tau_hi) appended to cand (count in ds->cand_count;
// entries beyond `cap` are not written).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_select(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
             const T* __restrict__ u, const int* __restrict__ nbd,
             const uint8_t* __restrict__ iwhere, DevScalars* ds, Cand* cand, long long cap) {
  if (!ds->gcp_pending) return;
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  const double lo = ds->gcp_tau_lo, hi = ds->gcp_tau_hi;
  const int lane = threadIdx.x & 63;
  for (long long i0 = blockIdx.x * (long long)kBlock; i0 < n;
       i0 += (long long)gridDim.x * kBlock) {
    const long long i = i0 + threadIdx.x;
    bool take = false;
    T tb = 0;
    if (i < n) {
      if (breakpoint_of(x[i], g[i], l[i], u[i], nbd[i], iw_get(iwhere[i]), tb)) {
        const double t = (double)tb;
        take = t >= lo && t < hi;
      }
    }
    const unsigned long long mask = __ballot(take);
    if (mask) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(&ds->cand_count, (unsigned long long)__popcll(mask));
      base = __shfl(base, 0, kWave);
      const unsigned long long slot = base + __popcll(mask & ((1ull << lane) - 1ull));
      if (take && slot < (unsigned long long)cap) {
        Cand c;
        c.t = (double)tb;
        c.idx = (int)i;
        c.pad = 0;
        cand[slot] = c;
      }
    }
  }
}

__device__ inline long long round_size(const DevScalars* ds, long long cap) {
  const long long k = (long long)ds->cand_count;
  return k < cap ? k : cap;
}

// Sort each kSortTile-long tile of the candidates in LDS (bitonic).
__global__ void __launch_bounds__(kBlock)
k_sort_tile(Cand* a, const DevScalars* ds, long long cap) {
  __shared__ Cand sh[kSortTile];
  if (!ds->gcp_pending) return;
  const long long K = round_size(ds, cap);
  const long long base = (long long)blockIdx.x * kSortTile;
  if (base >= K) return;
  for (int j = threadIdx.x; j < kSortTile; j += blockDim.x) {
    if (base + j < K) {
      sh[j] = a[base + j];
    } else {
      sh[j].t = __builtin_inf();
      sh[j].idx = 0x7fffffff;
      sh[j].pad = 0;
    }
  }
  __syncthreads();
  for (int k = 2; k <= kSortTile; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int p = threadIdx.x; p < kSortTile / 2; p += blockDim.x) {
        const int i = (p / j) * 2 * j + (p % j), l = i + j;
        const bool up = (i & k) == 0;
        if (cand_less(sh[l], sh[i]) == up) {
          Cand tmp = sh[i];
          sh[i] = sh[l];
          sh[l] = tmp;
        }
      }
      __syncthreads();
    }
  }
  for (int j = threadIdx.x; j < kSortTile; j += blockDim.x)
    if (base + j < K) a[base + j] = sh[j];
}

// One merge-path pass: sorted runs of `width` -> runs of 2*width.
static constexpr int kMergeItems = 4;
__global__ void __launch_bounds__(kBlock)
k_merge(const Cand* __restrict__ src, Cand* __restrict__ dst, const DevScalars* ds,
        long long cap, long long width) {
  if (!ds->gcp_pending) return;
  const long long K = round_size(ds, cap);
  const long long o0 = ((long long)blockIdx.x * kBlock + threadIdx.x) * kMergeItems;
  if (o0 >= K) return;
  const long long base = (o0 / (2 * width)) * 2 * width;
  const long long a0 = base, a1 = min(base + width, K);
  const long long b0 = a1, b1 = min(base + 2 * width, K);
  const long long la = a1 - a0, lb = b1 - b0;
  const long long diag = o0 - base;
  long long lo = diag - lb > 0 ? diag - lb : 0, hi = diag < la ? diag : la;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (cand_less(src[a0 + mid], src[b0 + diag - 1 - mid]))
      lo = mid + 1;
    else
      hi = mid;
  }
  long long i = lo, j = diag - lo;
  for (int k = 0; k < kMergeItems && o0 + k < K; k++) {
    bool takeA;
    if (i >= la)
      takeA = false;
    else if (j >= lb)
      takeA = true;
    else
      takeA = !cand_less(src[b0 + j], src[a0 + i]);
    dst[o0 + k] = takeA ? src[a0 + i++] : src[b0 + j++];
  }
}

// Contribution of candidate c: [d w (2col) ; t d w (2col) ; d^2].
template <typename T>
__device__ inline void contribution(const Cand& c, const T* __restrict__ g,
                                    const T* __restrict__ W, const RingReads& rr,
                                    double theta, double (&v)[kNV]) {
  const long long i = c.idx;
  // the whole W row and g_i are loaded before any use (positions j >= col
  // read as zero)
  T yv[kMaxM], sv[kMaxM];
  const T gi = g[i];
  load_ring(W, i, rr, yv, sv);
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

// Layout of a state vector in this file: [P_y(8) P_s(8) Q_y(8) Q_s(8) D].
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_tile_reduce(const Cand* __restrict__ cand, Pair<T> Gv, const T* __restrict__ W,
                  const DevScalars* ds, long long cap, double* tsum, int tstride) {
  __shared__ double sh[kWavesPerBlock * kNV];
  if (!ds->gcp_pending) return;
  const long long K = round_size(ds, cap);
  if ((long long)blockIdx.x * kBlock >= K) return;
  const RingReads rr = ring_reads(ds->head, ds->m, ds->col, -1);
  const T* g = Gv[ds->phase & 1];
  double v[kNV];
#pragma unroll
  for (int q = 0; q < kNV; q++) v[q] = 0.0;
  const long long pos = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (pos < K) contribution(cand[pos], g, W, rr, ds->theta, v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kNV; q++) {
    double s = wave_sum(v[q]);
    if (lane == 0) sh[wid * kNV + q] = s;
  }
  __syncthreads();
  if (threadIdx.x < kNV) {
    double s = 0.0;
    for (int w = 0; w < kWavesPerBlock; w++) s += sh[w * kNV + threadIdx.x];
    tsum[(size_t)threadIdx.x * tstride + blockIdx.x] = s;
  }
}

// Exclusive scan over tiles (the carried state included) and the round total.
__global__ void s_gcp_tile_scan(DevScalars* ds, long long cap, const double* tsum,
                                double* tpre, int tstride, double* total) {
  __shared__ double sh[kWavesPerBlock][kWave];
  if (!ds->gcp_pending) return;
  const long long K = round_size(ds, cap);
  const int ntiles = (int)((K + kBlock - 1) / kBlock);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // one wave per state component, 64 tiles per step
  for (int q = w; q < kNV; q += kWavesPerBlock) {
    const double* row = tsum + (size_t)q * tstride;
    double* out = tpre + (size_t)q * tstride;
    double run = ds->gcp_carry[q];
    for (int t0 = 0; t0 < ntiles; t0 += kWave) {
      const int t = t0 + lane;
      const double v = t < ntiles ? row[t] : 0.0;
      double incl = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double up = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += up;
      }
      if (t < ntiles) out[t] = run + incl - v;
      run += __shfl(incl, 63, kWave);
    }
    if (lane == 0) total[q] = run;
  }
  (void)sh;
}

// Stop test at the state `st` (passed breakpoints before t_next).
// All loops run over the padded 16 positions with compile-time indices (the
// entries beyond col are exact zeros, which leaves every sum unchanged), so
// the temporaries stay in registers.
__device__ inline bool stop_test(const DevScalars* ds, const GcpConst& gc,
                                 const double (&st)[kNV], double t_prev, double t_next,
                                 bool floor_on, double (&out)[3]) {
  const double theta = ds->theta;
  double p[16], mp[16];
#pragma unroll
  for (int q = 0; q < 16; q++) p[q] = gc.pp[q] - st[q];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; k++) s += gc.Mp[r * 16 + k] * p[k];
    mp[r] = s;
  }
  double pmp = 0.0, qmp = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    pmp += p[j] * mp[j] + p[8 + j] * mp[8 + j];
    qmp += st[16 + j] * mp[j] + st[24 + j] * mp[8 + j];
  }
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

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_tile_eval(const Cand* __restrict__ cand, Pair<T> Gv, const T* __restrict__ W,
                const DevScalars* ds, long long cap, const double* tpre, int tstride,
                long long* tile_stop, double* tile_state, double* last_state) {
  __shared__ double sh[kWavesPerBlock * kNV];
  __shared__ long long sstop;
  __shared__ GcpConst gc;
  if (!ds->gcp_pending) return;
  const long long K = round_size(ds, cap);
  if ((long long)blockIdx.x * kBlock >= K) return;
  stage_gcp_const(ds, gc);
  const RingReads rr = ring_reads(ds->head, ds->m, ds->col, -1);
  const T* g = Gv[ds->phase & 1];
  const long long pos = (long long)blockIdx.x * kBlock + threadIdx.x;
  double v[kNV], ex[kNV];
#pragma unroll
  for (int q = 0; q < kNV; q++) v[q] = 0.0;
  Cand c;
  c.t = 0;
  c.idx = 0;
  c.pad = 0;
  if (pos < K) {
    c = cand[pos];
    contribution(c, g, W, rr, ds->theta, v);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) sstop = 0x7fffffffffffffffLL;
#pragma unroll
  for (int q = 0; q < kNV; q++) {
    double incl = v[q];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double up = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += up;
    }
    ex[q] = incl - v[q];
    if (lane == 63) sh[wid * kNV + q] = incl;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kNV; q++) {
    double off = tpre[(size_t)q * tstride + blockIdx.x];
    for (int w = 0; w < wid; w++) off += sh[w * kNV + q];
    ex[q] += off;
  }
  if (pos < K) {
    const long long k = ds->gcp_kpassed + pos;
    if (k > 0) {
      const double t_prev = pos > 0 ? cand[pos - 1].t : ds->gcp_t_prev;
      double out[3];
      if (stop_test(ds, gc, ex, t_prev, c.t, true, out)) atomicMin(&sstop, pos);
    }
    if (pos == K - 1)
      for (int q = 0; q < kNV; q++) last_state[q] = ex[q];
  }
  __syncthreads();
  if (sstop != 0x7fffffffffffffffLL && pos == sstop)
    for (int q = 0; q < kNV; q++) tile_state[(size_t)blockIdx.x * kNV + q] = ex[q];
  if (threadIdx.x == 0) tile_stop[blockIdx.x] = sstop;
}

__device__ void gcp_done(DevScalars* ds) {
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
__device__ void round_conclude(DevScalars* ds, const GcpConst& gc, const Cand* cand, long long K,
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

// End of a multi-block round: either the first stop is known (finish the
// Cauchy point) or every candidate passed (carry the sums; next round or
// terminal case).  `cap` is the candidate capacity the launches were sized
// for.
__global__ void s_gcp_round_end(DevScalars* ds, const Cand* __restrict__ cand, long long cap,
                                const long long* tile_stop, const double* tile_state,
                                const double* last_state, const double* total) {
  __shared__ long long sbest;
  if (!ds->gcp_pending) return;
  if ((long long)ds->cand_count > cap) {
    // too many candidates for this launch: the host sorts them
    if (threadIdx.x == 0) {
      ds->status = ST_GCP_ROUNDS;
      ds->halt = 1;
    }
    return;
  }
  __shared__ GcpConst gc;
  stage_gcp_const(ds, gc);
  const long long K = (long long)ds->cand_count;
  const int ntiles = (int)((K + kBlock - 1) / kBlock);
  if (threadIdx.x == 0) sbest = 0x7fffffffffffffffLL;
  __syncthreads();
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x)
    if (tile_stop[t] != 0x7fffffffffffffffLL) atomicMin(&sbest, tile_stop[t]);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const bool stopped = sbest != 0x7fffffffffffffffLL;
  const double* st = stopped ? tile_state + (size_t)(sbest / kBlock) * kNV : total;
  round_conclude(ds, gc, cand, K, stopped ? sbest : -1, st, last_state);
}

// A whole round in one workgroup when it has at most kSmallRound candidates
// (the common case once the active set settles): bitonic sort in LDS, then
// 256-candidate chunks scanned in order, stopping at the first chunk that
// holds the stop.  Larger rounds are left to the multi-block path.
static constexpr int kSmallRound = 2048;
template <typename T>
__global__ void __launch_bounds__(kBlock)
s_gcp_round_small(const Cand* __restrict__ cand, Pair<T> Gv, const T* __restrict__ W,
                  DevScalars* ds) {
  __shared__ Cand sc[kSmallRound];
  __shared__ double swave[kWavesPerBlock * kNV];
  __shared__ double scarry[kNV];
  __shared__ double sstop_state[kNV];
  __shared__ double slast[kNV];
  __shared__ long long sstop;
  __shared__ GcpConst gc;
  if (!ds->gcp_pending || ds->halt) return;
  const long long K = (long long)ds->cand_count;
  if (K > kSmallRound) {
    if (threadIdx.x == 0) {
      ds->status = ST_GCP_ROUNDS;
      ds->halt = 1;
    }
    return;
  }
  stage_gcp_const(ds, gc);
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
  if (threadIdx.x < kNV) scarry[threadIdx.x] = ds->gcp_carry[threadIdx.x];
  if (threadIdx.x == 0) sstop = 0x7fffffffffffffffLL;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int p = threadIdx.x; p < P / 2; p += blockDim.x) {
        const int i = (p / j) * 2 * j + (p % j), l = i + j;
        const bool up = (i & k) == 0;
        if (cand_less(sc[l], sc[i]) == up) {
          Cand tmp = sc[i];
          sc[i] = sc[l];
          sc[l] = tmp;
        }
      }
      __syncthreads();
    }
  }
  const RingReads rr = ring_reads(ds->head, ds->m, ds->col, -1);
  const double theta = ds->theta;
  const long long kpassed = ds->gcp_kpassed;
  const double t_prev0 = ds->gcp_t_prev;
  const T* g = Gv[ds->phase & 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (long long c0 = 0; c0 < K; c0 += kBlock) {
    const long long pos = c0 + threadIdx.x;
    double v[kNV], ex[kNV];
#pragma unroll
    for (int q = 0; q < kNV; q++) v[q] = 0.0;
    if (pos < K) contribution(sc[pos], g, W, rr, theta, v);
#pragma unroll
    for (int q = 0; q < kNV; q++) {
      double incl = v[q];
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double up = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += up;
      }
      ex[q] = incl - v[q];
      if (lane == 63) swave[wid * kNV + q] = incl;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kNV; q++) {
      double off = scarry[q];
      for (int w = 0; w < wid; w++) off += swave[w * kNV + q];
      ex[q] += off;
    }
    bool stop = false;
    if (pos < K && kpassed + pos > 0) {
      const double tp = pos > 0 ? sc[pos - 1].t : t_prev0;
      double out[3];
      stop = stop_test(ds, gc, ex, tp, sc[pos].t, true, out);
    }
    if (stop) atomicMin(&sstop, pos);
    if (pos == K - 1)
      for (int q = 0; q < kNV; q++) slast[q] = ex[q];
    __syncthreads();
    if (sstop != 0x7fffffffffffffffLL) {
      if (pos == sstop)
        for (int q = 0; q < kNV; q++) sstop_state[q] = ex[q];
      break;
    }
    // carry the chunk total
    if (threadIdx.x < kNV) {
      double t = scarry[threadIdx.x];
      for (int w = 0; w < kWavesPerBlock; w++) t += swave[w * kNV + threadIdx.x];
      scarry[threadIdx.x] = t;
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const bool stopped = sstop != 0x7fffffffffffffffLL;
  round_conclude(ds, gc, sc, K, stopped ? sstop : -1, stopped ? sstop_state : scarry, slast);
}

// ---------------------------------------------------------------------------
template <typename T>
void launch_gcp_select(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
                       const T* l, const T* u, const int* nbd, const uint8_t* iwhere,
                       DevScalars* ds, Cand* ca, long long buf_cap) {
  hipLaunchKernelGGL(k_gcp_select<T>, dim3(G), dim3(kBlock), 0, s, n, X, Gv, l, u, nbd, iwhere,
                     ds, ca, buf_cap);
}

// Sort + scan + round end for at most `cap` candidates (the device count
// decides how much of the launch does work).
template <typename T>
void launch_gcp_sortscan(hipStream_t s, Pair<T> Gv, const T* W, DevScalars* ds, Cand* ca,
                         Cand* cb, long long cap, const GcpScratch& w) {
  if (cap < 1) cap = 1;
  const long long tiles = (cap + kSortTile - 1) / kSortTile;
  hipLaunchKernelGGL(k_sort_tile, dim3((unsigned)tiles), dim3(kBlock), 0, s, ca, ds, cap);
  Cand* src = ca;
  Cand* dst = cb;
  for (long long wdt = kSortTile; wdt < cap; wdt *= 2) {
    const long long threads = (cap + kMergeItems - 1) / kMergeItems;
    const long long blocks = (threads + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_merge, dim3((unsigned)blocks), dim3(kBlock), 0, s, src, dst, ds, cap,
                       wdt);
    Cand* t = src;
    src = dst;
    dst = t;
  }
  const int ntiles = (int)((cap + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_gcp_tile_reduce<T>, dim3(ntiles), dim3(kBlock), 0, s, src, Gv, W, ds,
                     cap, w.tsum, w.tstride);
  hipLaunchKernelGGL(s_gcp_tile_scan, dim3(1), dim3(kBlock), 0, s, ds, cap, w.tsum, w.tpre,
                     w.tstride, w.total);
  hipLaunchKernelGGL(k_gcp_tile_eval<T>, dim3(ntiles), dim3(kBlock), 0, s, src, Gv, W, ds, cap,
                     w.tpre, w.tstride, w.tile_stop, w.tile_state, w.last_state);
  hipLaunchKernelGGL(s_gcp_round_end, dim3(1), dim3(kBlock), 0, s, ds, src, cap, w.tile_stop,
                     w.tile_state, w.last_state, w.total);
}

template <typename T>
void launch_gcp_round_small(hipStream_t s, const Cand* ca, Pair<T> Gv, const T* W,
                            DevScalars* ds) {
  hipLaunchKernelGGL(s_gcp_round_small<T>, dim3(1), dim3(kBlock), 0, s, ca, Gv, W, ds);
}

#define LBFGSB_INST_GCP(T)                                                                  \
  template void launch_gcp_round_small<T>(hipStream_t, const Cand*, Pair<T>, const T*,   \
                                          DevScalars*);                                    \
  template void launch_gcp_select<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,    \
                                     const T*, const T*, const int*, const uint8_t*,       \
                                     DevScalars*, Cand*, long long);                       \
  template void launch_gcp_sortscan<T>(hipStream_t, Pair<T>, const T*, DevScalars*,      \
                                       Cand*, Cand*, long long, const GcpScratch&);

LBFGSB_INST_GCP(float)
LBFGSB_INST_GCP(double)

}  // namespace lbfgsb
