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
// by 4x.  Pass A leaves every variable's breakpoint t (+inf for none) in an
// n-vector, so a selection reads one value per variable, and round 0 runs
// entirely on the device with its
// size read on the device when one workgroup can sort it (2048 candidates);
// larger or later rounds are
// sized by the host.
#include "gcp_device.hpp"
#include "pass_device.hpp"

namespace lbfgsb {

// Candidates t in [tau_lo, tau_hi) written to cand (count in ds->cand_count;
// entries beyond `cap` are not written).
//
// From pass A's lists (k_gcp_near_gather): when the t-range lies below the
// bound pass A listed breakpoints under (ds->gcp_near_tau) and no block's
// list overflowed, the round's candidates are read from the lists.  Each
// block of the gather takes a run of lists, filters them into LDS and
// reserves its range of cand with one atomic (the lists are short: a few
// candidates per thousand variables once the search settles).  The gather
// decides for the whole round (gcp_sel_near); when the lists do not cover
// it, the selection below runs in the chain (in_chain), or the chain halts
// with the range to be selected by the host.
//
// From the variables (k_gcp_count, k_gcp_select): every variable is
// classified again from x, g, l, u and its stored class (the same t pass A
// forms), in two launches with no global atomics (one counter bumped by
// every wave that finds a candidate serialises at about 20 ns per update,
// milliseconds for a round of a few million).  Block b visits the
// kSelTile-long chunks b, b + G, ... (kSelVec elements per thread, a block
// stride apart).  k_gcp_count leaves the block's count in sel[b] and in
// sel[G + b] a 64-bit mask of its chunk groups that hold any; k_gcp_select
// re-reads only those groups and writes their candidates at the offset the
// counts of the blocks before give.
constexpr int kSelVec = 4;
constexpr int kSelTile = kBlock * kSelVec;
constexpr int kGatherLists = 64;  // pass-A lists per block of the gather

// Whether pass A's lists hold every candidate of the current t-range (the
// largest list count in `nmax`).
__device__ __forceinline__ bool near_covers(const DevScalars* ds, unsigned nmax, int cap) {
  return ds->gcp_near_tau > 0.0 && ds->gcp_tau_hi <= ds->gcp_near_tau &&
         nmax <= (unsigned)cap;
}

// The largest list count and the sum of this thread's share of the
// G <= kMaxBlocks counts (loaded together).
__device__ __forceinline__ unsigned near_max_count(const unsigned* __restrict__ nearn, int G,
                                                   unsigned& sum) {
  unsigned mx = 0u;
  sum = 0u;
#pragma unroll
  for (int j = 0; j < kMaxBlocks / kBlock; j++) {
    const int b = threadIdx.x + j * kBlock;
    const unsigned c = b < G ? nearn[b] : 0u;
    mx = max(mx, c);
    sum += c;
  }
  return mx;
}

// The lists do not cover the round (one thread): a list that overflowed asks
// for a lower bound next time, a range beyond the bound for a higher one;
// unless the full selection follows in the chain (in_chain), the host
// selects the range (run_rounds).
__device__ __forceinline__ void near_miss(DevScalars* ds, unsigned nmax, int near_cap,
                                          int in_chain) {
  ds->gcp_sel_near = 0;
  if (ds->gcp_near_tau > 0.0) {
    ds->gcp_near_misses++;
    const double gr = ds->gcp_near_grow > 0.0 ? ds->gcp_near_grow : kNearGrow0;
    ds->gcp_near_grow = nmax > (unsigned)near_cap ? fmax(gr * 0.25, 1.0) : fmin(gr * 4.0, 1024.0);
  }
  if (!in_chain) {
    ds->gcp_need_select = 1;
    ds->status = ST_GCP_ROUNDS;
    ds->halt = 1;
  }
}

// Exclusive offsets of the G lists' entries in a flat order into off[0 .. G]
// (LDS, off[G] the total) and the largest count into *cmax (LDS, zero on
// entry): thread t scans the counts of lists 8t .. 8t + 7.
constexpr int kGatherU = 4;
__device__ __forceinline__ void near_offsets(const NearLists& nl, unsigned* off, unsigned* cmax) {
  constexpr int kPer = kMaxBlocks / kBlock;
  __shared__ unsigned s_wsum[kBlock / kWave];
  unsigned c[kPer];
  unsigned run = 0, mx = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int b = threadIdx.x * kPer + j;
    c[j] = b < nl.G ? nl.n[b] : 0u;
    run += c[j];
    mx = max(mx, c[j]);
  }
  __syncthreads();  // *cmax is zero
  atomicMax(cmax, mx);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned incl = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned up = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += up;
  }
  if (lane == 63) s_wsum[w] = incl;
  __syncthreads();
  unsigned base = incl - run;
  for (int v = 0; v < w; v++) base += s_wsum[v];
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int b = threadIdx.x * kPer + j;
    if (b <= nl.G) off[b] = base;
    base += c[j];
  }
  if (threadIdx.x == kBlock - 1) off[nl.G] = base;  // the total (also when G = kMaxBlocks)
  __syncthreads();
}

// Position in the lists of flat entry q < off[G]: the list b with
// off[b] <= q < off[b + 1] (binary search in LDS), entry q - off[b].
__device__ __forceinline__ size_t near_entry(const unsigned* off, const NearLists& nl,
                                             unsigned q) {
  int lo = 0, hi = nl.G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= q)
      lo = mid;
    else
      hi = mid - 1;
  }
  return (size_t)lo * nl.cap + (q - off[lo]);
}

__global__ void __launch_bounds__(kBlock)
k_gcp_near_gather(const Cand* __restrict__ nearc, const unsigned* __restrict__ nearn, int G,
                  int near_cap, DevScalars* ds, Cand* cand, long long cap, int in_chain) {
  __shared__ unsigned s_max, s_cnt, s_sum;
  __shared__ unsigned long long s_base;
  __shared__ Cand s_c[kGatherLists * 8];
  if (!ds->gcp_pending || ds->halt) return;
  if (threadIdx.x == 0) {
    s_max = 0u;
    s_cnt = 0u;
    s_sum = 0u;
  }
  __syncthreads();
  unsigned part;
  atomicMax(&s_max, near_max_count(nearn, G, part));
  atomicAdd(&s_sum, part);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) ds->gcp_near_listed = (long long)s_sum;
  const bool covers = near_covers(ds, s_max, near_cap);
  if (!covers) {
    if (blockIdx.x == 0 && threadIdx.x == 0) near_miss(ds, s_max, near_cap, in_chain);
    return;
  }
  const double lo = ds->gcp_tau_lo, hi = ds->gcp_tau_hi;
  // this block's lists: kGatherLists of them, kBlock / kGatherLists threads
  // per list taking every such entry (the lists' loads are independent)
  constexpr int kPerList = kBlock / kGatherLists;
  const int b = blockIdx.x * kGatherLists + (threadIdx.x % kGatherLists);
  const unsigned c = b < G ? nearn[b] : 0u;
  const Cand* lst = nearc + (size_t)(b < G ? b : 0) * near_cap;
  // candidates in range: LDS first (up to its size), the rest straight to
  // cand after the reservation
  for (unsigned k = threadIdx.x / kGatherLists; k < c; k += kPerList) {
    const Cand e = lst[k];
    if (e.t >= lo && e.t < hi) {
      const unsigned q = atomicAdd(&s_cnt, 1u);
      if (q < kGatherLists * 8) s_c[q] = e;
    }
  }
  __syncthreads();
  const unsigned cnt = s_cnt;
  if (threadIdx.x == 0) s_base = cnt ? atomicAdd(&ds->cand_count, (unsigned long long)cnt) : 0ull;
  __syncthreads();
  const unsigned long long base = s_base;
  if (cnt <= kGatherLists * 8) {
    for (unsigned q = threadIdx.x; q < cnt; q += kBlock)
      if (base + q < (unsigned long long)cap) cand[base + q] = s_c[q];
  } else {
    // more than the LDS holds: filter the lists again into the reservation
    // (the order inside it need not be the first pass's: rounds sort by
    // (t, index))
    if (threadIdx.x == 0) s_cnt = 0u;
    __syncthreads();
    for (unsigned k = threadIdx.x / kGatherLists; k < c; k += kPerList) {
      const Cand e = lst[k];
      if (e.t >= lo && e.t < hi) {
        const unsigned long long q = base + atomicAdd(&s_cnt, 1u);
        if (q < (unsigned long long)cap) cand[q] = e;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ds->gcp_sel_near = 1;
    // no block reads the flag; the host reads it after the launch
    ds->gcp_need_select = 0;
  }
}

// The breakpoints of kSelVec elements e, e + kBlock, ... of a chunk (t, or
// +inf where there is none), classified from x, g, l, u and the stored class
// as pass A classifies them.
template <typename T>
__device__ __forceinline__ void sel_load(const T* __restrict__ x, const T* __restrict__ g,
                                         const T* __restrict__ l, const T* __restrict__ u,
                                         const uint8_t* __restrict__ iwhere, long long n,
                                         long long e, double (&t)[kSelVec]) {
  T xv[kSelVec], gv[kSelVec], lv[kSelVec], uv[kSelVec];
  uint8_t bv[kSelVec];
#pragma unroll
  for (int q = 0; q < kSelVec; q++) {
    const long long i = e + (long long)q * kBlock < n ? e + (long long)q * kBlock : n - 1;
    xv[q] = x[i];
    gv[q] = g[i];
    lv[q] = l[i];
    uv[q] = u[i];
    bv[q] = iwhere[i];
  }
#pragma unroll
  for (int q = 0; q < kSelVec; q++) {
    int iw;
    T neggi, tb;
    bool isfree, isbrk;
    classify(xv[q], gv[q], lv[q], uv[q], iw_nbd(bv[q]), iw_get(bv[q]), iw, neggi, isfree, isbrk,
             tb);
    t[q] = (isbrk && e + (long long)q * kBlock < n) ? (double)tb : __builtin_inf();
  }
}

// chunks per mask bit, the same for every block
__device__ __forceinline__ long long sel_span(long long n) {
  const long long nchunk = (n + kSelTile - 1) / kSelTile;
  const long long per_block = (nchunk + gridDim.x - 1) / gridDim.x;
  return (per_block + 63) / 64;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_count(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
            const T* __restrict__ u, const uint8_t* __restrict__ iwhere, const DevScalars* ds,
            unsigned long long* __restrict__ sel) {
  __shared__ int scnt[kBlock / kWave];
  __shared__ unsigned long long smask[kBlock / kWave];
  if (!ds->gcp_pending || ds->halt || ds->gcp_sel_near) return;
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  const double lo = ds->gcp_tau_lo, hi = ds->gcp_tau_hi;
  const long long span = sel_span(n);
  int cnt = 0;
  unsigned long long mask = 0;
  long long k = 0;
  for (long long c = blockIdx.x; c * kSelTile < n; c += gridDim.x, k++) {
    double t[kSelVec];
    sel_load(x, g, l, u, iwhere, n, c * kSelTile + threadIdx.x, t);
    int here = 0;
#pragma unroll
    for (int q = 0; q < kSelVec; q++) here += (t[q] >= lo && t[q] < hi) ? 1 : 0;  // +inf never
    cnt += here;
    if (here) mask |= 1ull << (k / span);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cnt += __shfl_xor(cnt, o, kWave);
    mask |= __shfl_xor(mask, o, kWave);
  }
  if ((threadIdx.x & 63) == 0) {
    scnt[threadIdx.x >> 6] = cnt;
    smask[threadIdx.x >> 6] = mask;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0, msk = 0;
    for (int w = 0; w < kBlock / kWave; w++) {
      c += (unsigned long long)scnt[w];
      msk |= smask[w];
    }
    sel[blockIdx.x] = c;
    sel[gridDim.x + blockIdx.x] = msk;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_select(long long n, Pair<T> X, Pair<T> Gv, const T* __restrict__ l,
             const T* __restrict__ u, const uint8_t* __restrict__ iwhere, DevScalars* ds,
             Cand* cand, long long cap, const unsigned long long* __restrict__ sel) {
  __shared__ unsigned long long sred[kBlock / kWave];
  __shared__ int swave[kBlock / kWave];
  if (!ds->gcp_pending || ds->halt || ds->gcp_sel_near) return;
  const int cur = ds->phase & 1;
  const T* __restrict__ x = X[cur];
  const T* __restrict__ g = Gv[cur];
  const bool last = blockIdx.x == gridDim.x - 1;
  const unsigned long long mine = sel[blockIdx.x];
  if (!mine && !last) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // offset of this block: the candidates of blocks 0 .. blockIdx.x - 1
  unsigned long long off = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kBlock) off += sel[b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o, kWave);
  if (lane == 0) sred[wid] = off;
  __syncthreads();
  unsigned long long base = 0;
  for (int w = 0; w < kBlock / kWave; w++) base += sred[w];
  if (last && threadIdx.x == 0) {
    ds->cand_count = base + mine;
    // no block reads the flag; the host reads it after the launch
    ds->gcp_need_select = 0;
  }
  if (!mine) return;
  const unsigned long long mask = sel[gridDim.x + blockIdx.x];
  const double lo = ds->gcp_tau_lo, hi = ds->gcp_tau_hi;
  const long long span = sel_span(n);
  long long k = 0;
  for (long long c = blockIdx.x; c * kSelTile < n; c += gridDim.x, k++) {
    if (!((mask >> (k / span)) & 1ull)) continue;  // the same for the whole block
    const long long e = c * kSelTile + threadIdx.x;
    double t[kSelVec];
    sel_load(x, g, l, u, iwhere, n, e, t);
    int here = 0;
#pragma unroll
    for (int q = 0; q < kSelVec; q++) here += (t[q] >= lo && t[q] < hi) ? 1 : 0;
    // exclusive prefix of the per-thread counts over the block
    int incl = here;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int up = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += up;
    }
    __syncthreads();  // swave of the previous chunk is read
    if (lane == 63) swave[wid] = incl;
    __syncthreads();
    unsigned long long slot = base + (unsigned long long)(incl - here), chunk = 0;
    for (int w = 0; w < kBlock / kWave; w++) {
      if (w < wid) slot += (unsigned long long)swave[w];
      chunk += (unsigned long long)swave[w];
    }
#pragma unroll
    for (int q = 0; q < kSelVec; q++) {
      if (t[q] >= lo && t[q] < hi) {
        if (slot < (unsigned long long)cap) {
          Cand cd;
          cd.t = t[q];
          cd.idx = (int)(e + (long long)q * kBlock);
          cd.pad = 0;
          cand[slot] = cd;
        }
        slot++;
      }
    }
    base += chunk;
  }
}

// Whether a launch of the multi-block round does work.  mid = 0: the host
// sized it (after resuming the chain).  mid = 1: enqueued in the chain right
// after the one-workgroup round, it takes over a round that one workgroup
// could not sort (more than kSmallRound candidates, at most its capacity),
// with no host round trip; every other state leaves it idle (a concluded
// round, a range still to select, a larger round left to the host).
__device__ inline bool round_active(const DevScalars* ds, long long cap, int mid) {
  if (!ds->gcp_pending) return false;
  if (!mid) return true;
  const long long K = (long long)ds->cand_count;
  return ds->status == ST_GCP_ROUNDS && !ds->gcp_need_select && K > kSmallRound && K <= cap;
}

__device__ inline long long round_size(const DevScalars* ds, long long cap) {
  const long long k = (long long)ds->cand_count;
  return k < cap ? k : cap;
}

// Sort each kSortTile-long tile of the candidates in LDS (bitonic).
__global__ void __launch_bounds__(kBlock)
k_sort_tile(Cand* a, const DevScalars* ds, long long cap, int mid) {
  __shared__ Cand sh[kSortTile];
  if (!round_active(ds, cap, mid)) return;
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
    // pair p of the stage with distance j = 2^lj: (i, i + j), i the p-th
    // index whose bit lj is clear (shifts: no integer division)
    for (int lj = __ffs(k) - 2; lj >= 0; lj--) {
      const int j = 1 << lj;
      for (int p = threadIdx.x; p < kSortTile / 2; p += blockDim.x) {
        const int i = ((p >> lj) << (lj + 1)) | (p & (j - 1)), l = i + j;
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
        long long cap, long long width, int mid) {
  if (!round_active(ds, cap, mid)) return;
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

// Per-tile sums of the candidates' contributions.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_tile_reduce(const Cand* __restrict__ cand, Pair<T> Gv, const T* __restrict__ W,
                  const DevScalars* ds, long long cap, double* tsum, int tstride, int mid) {
  __shared__ double sh[kWavesPerBlock * kNV];
  if (!round_active(ds, cap, mid)) return;
  const long long K = round_size(ds, cap);
  if ((long long)blockIdx.x * kBlock >= K) return;
  const RingReads rr = ring_reads(ds->head, ds->m, ds->col, -1, ds->pslot_x);
  const T* g = Gv[ds->phase & 1];
  double v[kNV];
#pragma unroll
  for (int q = 0; q < kNV; q++) v[q] = 0.0;
  const long long pos = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (pos < K) contribution(cand[pos], g, W, rr, ds->theta, v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int col = ds->col;
#pragma unroll
  for (int q = 0; q < kNV; q++) {
    const double s = nv_used(q, col) ? wave_sum(v[q]) : 0.0;
    if (lane == 0) sh[wid * kNV + q] = s;
  }
  __syncthreads();
  if (threadIdx.x < kNV) {
    double s = 0.0;
    for (int w = 0; w < kWavesPerBlock; w++) s += sh[w * kNV + threadIdx.x];
    tsum[(size_t)threadIdx.x * tstride + blockIdx.x] = s;
  }
}

// Exclusive scan over tiles (the carried state included) and the round total:
// block q scans component q, each of its kScanThreads threads a contiguous
// run of tiles (local sum, block scan of the run sums, then the run's
// prefixes), so a round of millions of candidates takes one pass.
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads)
k_gcp_tile_scan(const DevScalars* ds, long long cap, const double* __restrict__ tsum,
                double* __restrict__ tpre, int tstride, double* __restrict__ total, int mid) {
  __shared__ double swave[kScanThreads / kWave];
  if (!round_active(ds, cap, mid)) return;
  const long long K = round_size(ds, cap);
  const long long ntiles = (K + kBlock - 1) / kBlock;
  const int q = blockIdx.x;
  const double* row = tsum + (size_t)q * tstride;
  double* out = tpre + (size_t)q * tstride;
  const long long per = (ntiles + kScanThreads - 1) / kScanThreads;
  const long long t0 = (long long)threadIdx.x * per;
  const long long t1 = t0 + per < ntiles ? t0 + per : ntiles;
  double run = 0.0;
  for (long long t = t0; t < t1; t++) run += row[t];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double incl = wave_scan_incl(run);
  if (lane == 63) swave[w] = incl;
  __syncthreads();
  double off = ds->gcp_carry[q];
  for (int v = 0; v < w; v++) off += swave[v];
  off += incl - run;
  for (long long t = t0; t < t1; t++) {
    const double v = row[t];
    out[t] = off;
    off += v;
  }
  if (threadIdx.x == kScanThreads - 1) total[q] = off;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gcp_tile_eval(const Cand* __restrict__ cand, Pair<T> Gv, const T* __restrict__ W,
                const DevScalars* ds, long long cap, const double* tpre, int tstride,
                long long* tile_stop, double* tile_state, double* last_state, int mid) {
  __shared__ double sh[kWavesPerBlock * kNV];
  __shared__ long long sstop;
  __shared__ GcpConst gc;
  if (!round_active(ds, cap, mid)) return;
  const long long K = round_size(ds, cap);
  if ((long long)blockIdx.x * kBlock >= K) return;
  stage_gcp_const(ds, gc);
  const RingReads rr = ring_reads(ds->head, ds->m, ds->col, -1, ds->pslot_x);
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
  const int col = ds->col;
  if (threadIdx.x == 0) sstop = 0x7fffffffffffffffLL;
#pragma unroll
  for (int q = 0; q < kNV; q++) {
    double incl = v[q];
    if (nv_used(q, col)) incl = wave_scan_incl(incl);
    ex[q] = incl - v[q];
    if (lane == 63) sh[wid * kNV + q] = incl;
  }
  __syncthreads();
  // each component's wave offsets (the tile's prefix plus the totals of the
  // waves before, in wave order) in place of the wave totals
  if (threadIdx.x < kNV) {
    const int q = threadIdx.x;
    double run = tpre[(size_t)q * tstride + blockIdx.x];
    for (int w = 0; w < kBlock / kWave; w++) {
      const double t = sh[w * kNV + q];
      sh[w * kNV + q] = run;
      run += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kNV; q++) ex[q] += sh[wid * kNV + q];
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

// End of a multi-block round: either the first stop is known (finish the
// Cauchy point) or every candidate passed (carry the sums; next round or
// terminal case).  `cap` is the candidate capacity the launches were sized
// for.
__global__ void s_gcp_round_end(DevScalars* ds, const Cand* __restrict__ cand, long long cap,
                                const long long* tile_stop, const double* tile_state,
                                const double* last_state, const double* total, int mid) {
  __shared__ long long sbest;
  if (!round_active(ds, cap, mid)) return;
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
  if (mid) {
    // a concluded search lets the chain go on to the Cauchy-point pass; a
    // range still to select goes to the host as before
    if (!ds->gcp_pending) {
      ds->status = ST_RUNNING;
      ds->halt = 0;
    } else {
      ds->status = ST_GCP_ROUNDS;
      ds->halt = 1;
    }
  }
}

// A whole round in one workgroup when it has at most kSmallRound candidates
// (the common case once the active set settles): bitonic sort in LDS, then
// chunks scanned in order (gcp_device.hpp).  Larger rounds are left to the
// multi-block path.
//
// from_near: the round's candidates come straight from pass A's lists (no
// selection launch before it): the workgroup checks that the lists cover the
// t-range, filters them into LDS, and sorts them there; a round of more than
// kSmallRound candidates is written to cand for the multi-block path, and a
// range the lists do not cover halts the chain for the host's selection.
template <typename T>
__global__ void __launch_bounds__(kBlock)
s_gcp_round_small(Cand* __restrict__ cand, Pair<T> Gv, const T* __restrict__ W,
                  DevScalars* ds, NearLists nl, int from_near) {
  __shared__ RoundLds L;
  __shared__ DevScalars sds;
  __shared__ unsigned s_max, s_cnt;
  __shared__ unsigned s_off[kMaxBlocks + 1];
  if (!ds->gcp_pending || ds->halt) return;
  long long K;
  if (from_near) {
    if (threadIdx.x == 0) {
      s_max = 0u;
      s_cnt = 0u;
    }
    // every listed entry in range, into LDS while it has room.  The entries
    // are taken in a flat order (list offsets from a scan of the counts,
    // which also gives the largest count), kGatherU per thread per step,
    // their loads issued together
    unsigned* off = s_off;
    near_offsets(nl, off, &s_max);
    if (!near_covers(ds, s_max, nl.cap)) {
      if (threadIdx.x == 0) near_miss(ds, s_max, nl.cap, 0);
      return;
    }
    const double lo = ds->gcp_tau_lo, hi = ds->gcp_tau_hi;
    const unsigned total = off[nl.G];
    for (unsigned q0 = threadIdx.x; q0 < total; q0 += kGatherU * kBlock) {
      Cand e[kGatherU];
#pragma unroll
      for (int u = 0; u < kGatherU; u++) {
        const unsigned q = q0 + u * kBlock;
        e[u].t = __builtin_inf();
        if (q < total) e[u] = nl.c[near_entry(off, nl, q)];
      }
#pragma unroll
      for (int u = 0; u < kGatherU; u++) {
        if (e[u].t >= lo && e[u].t < hi) {
          const unsigned k = atomicAdd(&s_cnt, 1u);
          if (k < (unsigned)kSmallRound) L.sc[k] = e[u];
        }
      }
    }
    __syncthreads();
    K = (long long)s_cnt;
    if (K > kSmallRound) {
      // too many for this workgroup: all of them to cand (in any order: the
      // rounds sort by (t, index)) for the multi-block path
      if (threadIdx.x == 0) s_cnt = 0u;
      __syncthreads();
      for (unsigned q = threadIdx.x; q < total; q += kBlock) {
        const Cand e = nl.c[near_entry(off, nl, q)];
        if (e.t >= lo && e.t < hi) cand[atomicAdd(&s_cnt, 1u)] = e;
      }
    }
  } else {
    K = (long long)ds->cand_count;
  }
  if (K > kSmallRound) {
    if (threadIdx.x == 0) {
      if (from_near) {
        ds->cand_count = (unsigned long long)K;
        ds->gcp_sel_near = 1;
        ds->gcp_need_select = 0;
      }
      ds->status = ST_GCP_ROUNDS;
      ds->halt = 1;
      ds->gcp_handoffs++;
    }
    return;
  }
  // the round's scalar logic (stop tests, conclusion) on an LDS copy
  ds_load(&sds, ds);
  DevScalars* s = &sds;
  if (from_near && threadIdx.x == 0) {
    s->cand_count = (unsigned long long)K;
    s->gcp_sel_near = 1;
    s->gcp_need_select = 0;
    s->gcp_near_listed = (long long)s_off[nl.G];
  }
  stage_gcp_const(s, L.gc);
  // from the lists the candidates are in LDS already (sorted in place)
  block_sort_small(from_near ? L.sc : cand, K, L.sc);
  block_round_scan(s, L, L.sc, K, (const T*)Gv[s->phase & 1], W);
  ds_store(ds, s);
}

// ---------------------------------------------------------------------------
template <typename T>
void launch_gcp_select(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                       const T* u, const uint8_t* iwhere, const NearLists& nl, DevScalars* ds,
                       Cand* ca, long long buf_cap, unsigned long long* sel, bool full) {
  const int gg = (nl.G + kGatherLists - 1) / kGatherLists;
  hipLaunchKernelGGL(k_gcp_near_gather, dim3(gg), dim3(kBlock), 0, s, nl.c, nl.n, nl.G, nl.cap,
                     ds, ca, buf_cap, full ? 1 : 0);
  if (!full) return;
  // no more blocks than chunks, so that every block has at least one
  const long long chunks = (n + kSelTile - 1) / kSelTile;
  const int g = (int)std::max<long long>(1, std::min<long long>(G, chunks));
  hipLaunchKernelGGL(k_gcp_count<T>, dim3(g), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere, ds,
                     sel);
  hipLaunchKernelGGL(k_gcp_select<T>, dim3(g), dim3(kBlock), 0, s, n, X, Gv, l, u, iwhere, ds,
                     ca, buf_cap, sel);
}

// Sort + scan + round end for at most `cap` candidates (the device count
// decides how much of the launch does work).
template <typename T>
void launch_gcp_sortscan(hipStream_t s, Pair<T> Gv, const T* W, DevScalars* ds, Cand* ca,
                         Cand* cb, long long cap, const GcpScratch& w, int mid) {
  if (cap < 1) cap = 1;
  const long long tiles = (cap + kSortTile - 1) / kSortTile;
  hipLaunchKernelGGL(k_sort_tile, dim3((unsigned)tiles), dim3(kBlock), 0, s, ca, ds, cap, mid);
  Cand* src = ca;
  Cand* dst = cb;
  for (long long wdt = kSortTile; wdt < cap; wdt *= 2) {
    const long long threads = (cap + kMergeItems - 1) / kMergeItems;
    const long long blocks = (threads + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_merge, dim3((unsigned)blocks), dim3(kBlock), 0, s, src, dst, ds, cap,
                       wdt, mid);
    Cand* t = src;
    src = dst;
    dst = t;
  }
  const int ntiles = (int)((cap + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_gcp_tile_reduce<T>, dim3(ntiles), dim3(kBlock), 0, s, src, Gv, W, ds,
                     cap, w.tsum, w.tstride, mid);
  hipLaunchKernelGGL(k_gcp_tile_scan, dim3(kNV), dim3(kScanThreads), 0, s, ds, cap, w.tsum, w.tpre,
                     w.tstride, w.total, mid);
  hipLaunchKernelGGL(k_gcp_tile_eval<T>, dim3(ntiles), dim3(kBlock), 0, s, src, Gv, W, ds, cap,
                     w.tpre, w.tstride, w.tile_stop, w.tile_state, w.last_state, mid);
  hipLaunchKernelGGL(s_gcp_round_end, dim3(1), dim3(kBlock), 0, s, ds, src, cap, w.tile_stop,
                     w.tile_state, w.last_state, w.total, mid);
}

template <typename T>
void launch_gcp_round_small(hipStream_t s, Cand* ca, Pair<T> Gv, const T* W, DevScalars* ds,
                            const NearLists& nl, bool from_near) {
  hipLaunchKernelGGL(s_gcp_round_small<T>, dim3(1), dim3(kBlock), 0, s, ca, Gv, W, ds, nl,
                     from_near ? 1 : 0);
}

#define LBFGSB_INST_GCP(T)                                                                  \
  template void launch_gcp_round_small<T>(hipStream_t, Cand*, Pair<T>, const T*,         \
                                          DevScalars*, const NearLists&, bool);            \
  template void launch_gcp_select<T>(hipStream_t, int, long long, Pair<T>, Pair<T>,      \
                                     const T*, const T*, const uint8_t*, const NearLists&, \
                                     DevScalars*, Cand*, long long, unsigned long long*,   \
                                     bool);                                                \
  template void launch_gcp_sortscan<T>(hipStream_t, Pair<T>, const T*, DevScalars*,      \
                                       Cand*, Cand*, long long, const GcpScratch&, int);

LBFGSB_INST_GCP(float)
LBFGSB_INST_GCP(double)

}  // namespace lbfgsb
