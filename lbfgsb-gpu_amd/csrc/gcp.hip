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
// So the walk becomes: select the next batch of candidates by t, sort them
// by (t, index), prefix-sum the (4 col + 1)-vectors [d w, t d w, d^2], and
// test every k in parallel; the first k that stops ends the search.
// Batches grow geometrically, their t-ranges coming from a sorted sample of
// the breakpoint list, so the cost follows the number of breakpoints passed.
#include "common.hpp"
#include "dense.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

static constexpr int kSample = 2048;
static constexpr int kSortTile = 2048;
static constexpr long long kFirstRound = 4096;
static constexpr double kMachEpsG = 5e-16;

__device__ inline bool cand_less(const Cand& a, const Cand& b) {
  return a.t < b.t || (a.t == b.t && a.idx < b.idx);
}

// Bitonic sort of kSortTile doubles in LDS by 256 threads.
__device__ void bitonic_doubles(double* sh) {
  for (int k = 2; k <= kSortTile; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int p = threadIdx.x; p < kSortTile / 2; p += blockDim.x) {
        const int i = (p / j) * 2 * j + (p % j), l = i + j;
        const bool up = (i & k) == 0;
        const double a = sh[i], b = sh[l];
        if ((a > b) == up) {
          sh[i] = b;
          sh[l] = a;
        }
      }
      __syncthreads();
    }
  }
}

__device__ double next_tau(const DevScalars* ds, const double* samp, double lo_excl,
                           long long target) {
  const long long S = ds->gcp_nsample, nb = ds->gcp_nbreak;
  if (target >= nb) return __builtin_inf();
  long long q = (target * S + nb - 1) / nb;
  while (q < S && samp[q] <= lo_excl) q++;
  return q < S ? samp[q] : __builtin_inf();
}

// Set-up of the rounds: explicit M (columns M e_k), the sorted sample and the
// first t-range.
template <typename T>
__global__ void s_gcp_prepare(DevScalars* ds, const T* __restrict__ bp_t, double* samp,
                              unsigned long long* cnt) {
  __shared__ double sh[kSortTile];
  const int col = ds->col;
  if (threadIdx.x < 2 * col) {
    double e[16], out[16];
    for (int i = 0; i < 16; i++) e[i] = 0.0;
    e[threadIdx.x] = 1.0;
    d_bmv(ds->sy, ds->wt, col, e, out);
    for (int r = 0; r < 2 * col; r++) ds->Mx[r * 16 + threadIdx.x] = out[r];
  }
  const long long nb = ds->gcp_nbreak;
  const long long S = nb < kSample ? nb : kSample;
  for (int j = threadIdx.x; j < kSortTile; j += blockDim.x)
    sh[j] = j < S ? (double)bp_t[(long long)j * nb / S] : __builtin_inf();
  __syncthreads();
  bitonic_doubles(sh);
  for (int j = threadIdx.x; j < kSample; j += blockDim.x) samp[j] = sh[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    ds->gcp_nsample = S;
    for (int v = 0; v < 40; v++) ds->gcp_carry[v] = 0.0;
    ds->gcp_round = 0;
    ds->gcp_floor_hits = 0;
    ds->gcp_tau_lo = -__builtin_inf();
    ds->gcp_tau_hi = next_tau(ds, samp, ds->gcp_bkmin, kFirstRound);
    ds->gcp_last_round = ds->gcp_tau_hi == __builtin_inf();
    *cnt = 0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_select(const T* __restrict__ bp_t, const int* __restrict__ bp_i, long long nb,
             const DevScalars* ds, Cand* cand, unsigned long long* cnt) {
  const double lo = ds->gcp_tau_lo, hi = ds->gcp_tau_hi;
  const int lane = threadIdx.x & 63;
  for (long long i0 = blockIdx.x * (long long)kBlock; i0 < nb;
       i0 += (long long)gridDim.x * kBlock) {
    const long long i = i0 + threadIdx.x;
    bool take = false;
    double t = 0;
    if (i < nb) {
      t = (double)bp_t[i];
      take = t >= lo && t < hi;
    }
    const unsigned long long mask = __ballot(take);
    if (mask) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(cnt, (unsigned long long)__popcll(mask));
      base = __shfl(base, 0, kWave);
      const unsigned long long slot = base + __popcll(mask & ((1ull << lane) - 1ull));
      if (take && slot < (unsigned long long)nb) {
        Cand c;
        c.t = t;
        c.idx = bp_i[i];
        c.pad = 0;
        cand[slot] = c;
      }
    }
  }
}

// Sort each kSortTile-long tile of candidates in LDS.
__global__ void __launch_bounds__(kBlock)
k_sort_tile(Cand* a, long long K) {
  __shared__ Cand sh[kSortTile];
  const long long base = (long long)blockIdx.x * kSortTile;
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
k_merge(const Cand* __restrict__ src, Cand* __restrict__ dst, long long K, long long width) {
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
                                    const T* __restrict__ W, const DevScalars* ds, int col,
                                    int head, int m, double theta, double* v) {
  const long long i = c.idx;
  const double d = -(double)g[i];
  for (int j = 0; j < kMaxM; j++) {
    if (j < col) {
      int slot = head + j;
      if (slot >= m) slot -= m;
      const double wy = (double)W[widx(i, slot)];
      const double ws = theta * (double)W[widx(i, 8 + slot)];
      v[j] = d * wy;
      v[col + j] = d * ws;
      v[2 * col + j] = c.t * d * wy;
      v[3 * col + j] = c.t * d * ws;
    }
  }
  v[4 * col] = d * d;
}

static constexpr int kNV = 33;

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_tile_reduce(const Cand* __restrict__ cand, long long K, T* const* Gv,
                  const T* __restrict__ W, const DevScalars* ds, double* tsum, int ntiles) {
  __shared__ double sh[kWavesPerBlock * kNV];
  const int col = ds->col, nv = 4 * col + 1;
  const T* g = Gv[ds->phase & 1];
  double v[kNV];
  for (int q = 0; q < kNV; q++) v[q] = 0.0;
  const long long pos = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (pos < K) contribution(cand[pos], g, W, ds, col, ds->head, ds->m, ds->theta, v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < nv; q++) {
    double s = wave_sum(v[q]);
    if (lane == 0) sh[wid * kNV + q] = s;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double s = 0.0;
    for (int w = 0; w < kWavesPerBlock; w++) s += sh[w * kNV + threadIdx.x];
    tsum[(size_t)threadIdx.x * ntiles + blockIdx.x] = s;
  }
}

// Exclusive scan over tiles (carry included) and the round total.
__global__ void s_gcp_tile_scan(DevScalars* ds, const double* tsum, double* tpre, int ntiles,
                                double* total) {
  __shared__ double sh[kBlock];
  const int col = ds->col, nv = 4 * col + 1;
  const int per = (ntiles + kBlock - 1) / kBlock;
  for (int q = 0; q < nv; q++) {
    const double* row = tsum + (size_t)q * ntiles;
    double* out = tpre + (size_t)q * ntiles;
    const int b0 = threadIdx.x * per, b1 = min(b0 + per, ntiles);
    double s = 0.0;
    for (int b = b0; b < b1; b++) s += row[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double run = ds->gcp_carry[q];
      for (int t = 0; t < kBlock; t++) {
        const double x = sh[t];
        sh[t] = run;
        run += x;
      }
      total[q] = run;
    }
    __syncthreads();
    double run = sh[threadIdx.x];
    for (int b = b0; b < b1; b++) {
      out[b] = run;
      run += row[b];
    }
    __syncthreads();
  }
}

// State after passing the candidates before position pos: evaluates the
// stop test.  Returns dtm and the derived f1, f2 in out[0..2].
__device__ inline bool stop_test(const DevScalars* ds, const double* st, double t_prev,
                                 double t_next, bool floor_on, double* out) {
  const int col = ds->col, c2 = 2 * col;
  const double theta = ds->theta;
  double p[16], mp[16];
  for (int j = 0; j < c2; j++) p[j] = ds->p[j] - st[j];
  for (int r = 0; r < c2; r++) {
    double s = 0.0;
    for (int k = 0; k < c2; k++) s += ds->Mx[r * 16 + k] * p[k];
    mp[r] = s;
  }
  double pmp = 0.0, qmp = 0.0;
  for (int j = 0; j < c2; j++) {
    pmp += p[j] * mp[j];
    qmp += st[c2 + j] * mp[j];
  }
  const double D = ds->gcp_D0 - st[2 * c2];
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
k_gcp_tile_eval(const Cand* __restrict__ cand, long long K, T* const* Gv,
                const T* __restrict__ W, const DevScalars* ds, const double* tpre, int ntiles,
                long long* tile_stop, double* tile_state, double* last_state) {
  __shared__ double sh[kWavesPerBlock * kNV];
  __shared__ long long sstop;
  const int col = ds->col, nv = 4 * col + 1;
  const T* g = Gv[ds->phase & 1];
  const long long pos = (long long)blockIdx.x * kBlock + threadIdx.x;
  double v[kNV], ex[kNV];
  for (int q = 0; q < kNV; q++) v[q] = 0.0;
  Cand c;
  c.t = 0;
  c.idx = 0;
  if (pos < K) {
    c = cand[pos];
    contribution(c, g, W, ds, col, ds->head, ds->m, ds->theta, v);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) sstop = 0x7fffffffffffffffLL;
  for (int q = 0; q < nv; q++) {
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
  for (int q = 0; q < nv; q++) {
    double off = tpre[(size_t)q * ntiles + blockIdx.x];
    for (int w = 0; w < wid; w++) off += sh[w * kNV + q];
    ex[q] += off;
  }
  if (pos < K) {
    const long long k = ds->gcp_kpassed + pos;
    if (k > 0) {
      const double t_prev = pos > 0 ? cand[pos - 1].t : ds->gcp_t_prev;
      double out[3];
      if (stop_test(ds, ex, t_prev, c.t, true, out)) atomicMin(&sstop, pos);
    }
    if (pos == K - 1)
      for (int q = 0; q < nv; q++) last_state[q] = ex[q];
  }
  __syncthreads();
  if (sstop != 0x7fffffffffffffffLL && pos == sstop)
    for (int q = 0; q < nv; q++) tile_state[(size_t)blockIdx.x * kNV + q] = ex[q];
  if (threadIdx.x == 0) tile_stop[blockIdx.x] = sstop;
}

// End of a round: either the first stop is known (finish the Cauchy point)
// or every candidate passed (carry the sums, next round or terminal case).
__global__ void s_gcp_round_end(DevScalars* ds, const Cand* __restrict__ cand, long long K,
                                const long long* tile_stop, const double* tile_state,
                                const double* last_state, const double* total, int ntiles,
                                const double* samp, unsigned long long* cnt) {
  __shared__ long long sbest;
  __shared__ int stile;
  if (threadIdx.x == 0) {
    sbest = 0x7fffffffffffffffLL;
    stile = -1;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x) {
    if (tile_stop[t] != 0x7fffffffffffffffLL) atomicMin(&sbest, tile_stop[t]);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int col = ds->col, c2 = 2 * col, nv = 4 * col + 1;
  double st[kNV];
  double tk;
  long long ik;
  long long kpass;
  bool finished = false;
  double dtm = 0.0;
  int mode = 1;
  if (sbest != 0x7fffffffffffffffLL && K > 0) {
    const long long pos = sbest;
    const int tile = (int)(pos / kBlock);
    for (int q = 0; q < nv; q++) st[q] = tile_state[(size_t)tile * kNV + q];
    tk = pos > 0 ? cand[pos - 1].t : ds->gcp_t_prev;
    ik = pos > 0 ? cand[pos - 1].idx : ds->gcp_idx_prev;
    kpass = ds->gcp_kpassed + pos;
    double out[3];
    stop_test(ds, st, tk, cand[pos].t, kpass > 0, out);
    dtm = out[0];
    finished = true;
  } else {
    for (int q = 0; q < nv; q++) ds->gcp_carry[q] = total[q];
    ds->gcp_kpassed += K;
    if (K > 0) {
      ds->gcp_t_prev = cand[K - 1].t;
      ds->gcp_idx_prev = cand[K - 1].idx;
    }
    if (ds->gcp_kpassed >= ds->gcp_nbreak) {
      kpass = ds->gcp_kpassed;
      tk = ds->gcp_t_prev;
      ik = ds->gcp_idx_prev;
      if (ds->gcp_nbreak == ds->n) {
        // every variable was a breakpoint and all were passed: the
        // reference returns before updating p with the last one
        // (lbfgsbcpu.cpp:763-769); c = t_K p_{K-1} + Q_{K-1}
        for (int j = 0; j < c2; j++) {
          const double pj = ds->p[j] - last_state[j];
          ds->c[j] = tk * pj + last_state[c2 + j];
        }
        ds->nint = (int)ds->gcp_nbreak;
        ds->gcp_mode = 3;
        ds->gcp_tsum = tk;
        ds->gcp_last_t = tk;
        ds->gcp_last_idx = ik;
        for (int j = 0; j < c2; j++) ds->p[j] = ds->p[j] - last_state[j];
        if (col > 0 && d_bmv(ds->sy, ds->wt, col, ds->c, ds->mc) != 0) {
          ds->internalinfo = -8;
        }
        ds->nintol += ds->nint;
        ds->status = ST_RUNNING;
        ds->halt = 0;
        return;
      }
      for (int q = 0; q < nv; q++) st[q] = total[q];
      double out[3];
      stop_test(ds, st, tk, tk, true, out);
      if (!ds->gcp_unbnd)
        dtm = 0.0;
      else
        dtm = out[0];
      finished = true;
    } else {
      // next round
      ds->gcp_round++;
      ds->gcp_tau_lo = ds->gcp_tau_hi;
      long long target = kFirstRound;
      for (int r = 0; r < ds->gcp_round; r++) target *= 8;
      ds->gcp_tau_hi = next_tau(ds, samp, ds->gcp_tau_lo, ds->gcp_kpassed + target);
      *cnt = 0;
      ds->status = ST_GCP_ROUNDS;
      ds->halt = 1;
      return;
    }
  }
  if (finished) {
    if (dtm <= 0) dtm = 0;
    for (int j = 0; j < c2; j++) {
      const double pj = ds->p[j] - st[j];
      ds->p[j] = pj;
      ds->c[j] = (tk + dtm) * pj + st[c2 + j];
    }
    ds->gcp_dtm = dtm;
    ds->gcp_tsum = tk + dtm;
    ds->gcp_last_t = tk;
    ds->gcp_last_idx = ik;
    ds->gcp_mode = kpass > 0 ? 1 : 0;
    ds->nint = (int)(kpass + 1);
    if (col > 0 && d_bmv(ds->sy, ds->wt, col, ds->c, ds->mc) != 0) ds->internalinfo = -8;
    ds->nintol += ds->nint;
    ds->status = ST_RUNNING;
    ds->halt = 0;
  }
}

// ---------------------------------------------------------------------------
template <typename T>
void launch_gcp_prepare(hipStream_t s, DevScalars* ds, const T* bp_t, double* samp,
                        unsigned long long* cnt) {
  hipLaunchKernelGGL(s_gcp_prepare<T>, dim3(1), dim3(kBlock), 0, s, ds, bp_t, samp, cnt);
}
template <typename T>
void launch_gcp_select(hipStream_t s, int G, const T* bp_t, const int* bp_i, long long nb,
                       const DevScalars* ds, Cand* cand, unsigned long long* cnt) {
  hipLaunchKernelGGL(k_gcp_select<T>, dim3(G), dim3(kBlock), 0, s, bp_t, bp_i, nb, ds, cand,
                     cnt);
}
// Sorts K candidates; returns the buffer holding the result (a or b).
Cand* sort_candidates(hipStream_t s, Cand* a, Cand* b, long long K) {
  if (K <= 0) return a;
  const long long tiles = (K + kSortTile - 1) / kSortTile;
  hipLaunchKernelGGL(k_sort_tile, dim3((unsigned)tiles), dim3(kBlock), 0, s, a, K);
  Cand* src = a;
  Cand* dst = b;
  for (long long w = kSortTile; w < K; w *= 2) {
    const long long threads = (K + kMergeItems - 1) / kMergeItems;
    const long long blocks = (threads + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_merge, dim3((unsigned)blocks), dim3(kBlock), 0, s, src, dst, K, w);
    Cand* t = src;
    src = dst;
    dst = t;
  }
  return src;
}
template <typename T>
void launch_gcp_scan(hipStream_t s, const Cand* cand, long long K, T* const* Gv, const T* W,
                     DevScalars* ds, double* tsum, double* tpre, double* total,
                     long long* tile_stop, double* tile_state, double* last_state,
                     const double* samp, unsigned long long* cnt) {
  const int ntiles = (int)((K + kBlock - 1) / kBlock);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_gcp_tile_reduce<T>, dim3(ntiles), dim3(kBlock), 0, s, cand, K, Gv, W,
                       ds, tsum, ntiles);
  }
  hipLaunchKernelGGL(s_gcp_tile_scan, dim3(1), dim3(kBlock), 0, s, ds, tsum, tpre, ntiles,
                     total);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_gcp_tile_eval<T>, dim3(ntiles), dim3(kBlock), 0, s, cand, K, Gv, W,
                       ds, tpre, ntiles, tile_stop, tile_state, last_state);
  }
  hipLaunchKernelGGL(s_gcp_round_end, dim3(1), dim3(kBlock), 0, s, ds, cand, K, tile_stop,
                     tile_state, last_state, total, ntiles, samp, cnt);
}

#define LBFGSB_INST_GCP(T)                                                                   \
  template void launch_gcp_prepare<T>(hipStream_t, DevScalars*, const T*, double*,           \
                                      unsigned long long*);                                  \
  template void launch_gcp_select<T>(hipStream_t, int, const T*, const int*, long long,      \
                                     const DevScalars*, Cand*, unsigned long long*);         \
  template void launch_gcp_scan<T>(hipStream_t, const Cand*, long long, T* const*, const T*, \
                                   DevScalars*, double*, double*, double*, long long*,       \
                                   double*, double*, const double*, unsigned long long*);

LBFGSB_INST_GCP(float)
LBFGSB_INST_GCP(double)

}  // namespace lbfgsb
