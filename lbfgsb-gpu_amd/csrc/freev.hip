// Free/active index sets in the reference's exact order (lbfgsbcpu.cpp:
// 1135-1173), built from the per-variable flags that the hot path keeps in
// iwhere.  The iteration itself works on those flags (masked streaming, no
// gathers); this order-preserving compaction materialises index / indx2 when
// a caller asks for them:
//   index[0 .. nfree)       free variables, ascending
//   index[nfree .. n)       active variables, descending
//   indx2[0 .. nenter)      entering variables, descending
//   indx2[n-nleave .. n)    leaving variables, descending
// Values are 1-based variable numbers, as in the reference.
// Three launches: per-block counts, one-workgroup scan, ordered scatter.
#include "common.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

static constexpr int kItems = 16;
static constexpr long long kChunk = (long long)kBlock * kItems;

__device__ inline void flags_of(uint8_t b, bool track, int f[4]) {
  const bool fr = iw_wasfree(b);     // free in the partition just built
  const bool pf = iw_prevfree(b);    // free in the partition before
  f[0] = fr;
  f[1] = !fr;
  f[2] = track && fr && !pf;
  f[3] = track && !fr && pf;
}

__global__ void __launch_bounds__(kBlock)
k_export_count(const uint8_t* __restrict__ iwhere, long long n, int track,
               unsigned* counts, int nblk) {
  __shared__ unsigned sh[4];
  if (threadIdx.x < 4) sh[threadIdx.x] = 0;
  __syncthreads();
  const long long b0 = blockIdx.x * kChunk, b1 = min(b0 + kChunk, n);
  unsigned c[4] = {0, 0, 0, 0};
  for (long long i = b0 + threadIdx.x; i < b1; i += kBlock) {
    int f[4];
    flags_of(iwhere[i], track, f);
    for (int q = 0; q < 4; q++) c[q] += f[q];
  }
  for (int q = 0; q < 4; q++) {
    unsigned v = (unsigned)wave_sum_int((int)c[q]);
    if ((threadIdx.x & 63) == 0) atomicAdd(&sh[q], v);
  }
  __syncthreads();
  if (threadIdx.x < 4) counts[threadIdx.x * nblk + blockIdx.x] = sh[threadIdx.x];
}

__global__ void s_export_scan(unsigned* counts, int nblk, unsigned long long* totals) {
  // one wave per category, sequential over chunks of 64 blocks
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (q >= 4) return;
  unsigned* row = counts + (size_t)q * nblk;
  unsigned long long run = 0;
  for (int b0 = 0; b0 < nblk; b0 += 64) {
    const int b = b0 + lane;
    unsigned v = b < nblk ? row[b] : 0;
    unsigned incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      unsigned up = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += up;
    }
    if (b < nblk) row[b] = (unsigned)(run + incl - v);
    run += __shfl(incl, 63, kWave);
  }
  if (lane == 0) totals[q] = run;
}

__global__ void __launch_bounds__(kBlock)
k_export_write(const uint8_t* __restrict__ iwhere, long long n, int track,
               const unsigned* offs, int nblk, const unsigned long long* totals,
               int* __restrict__ index, int* __restrict__ indx2) {
  __shared__ unsigned wc[4][kWavesPerBlock];
  __shared__ unsigned base[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x < 4) base[threadIdx.x] = offs[threadIdx.x * nblk + blockIdx.x];
  __syncthreads();
  const long long nenter = (long long)totals[2];
  const long long b0 = blockIdx.x * kChunk, b1 = min(b0 + kChunk, n);
  for (long long s = b0; s < b1; s += kBlock) {
    const long long i = s + threadIdx.x;
    int f[4] = {0, 0, 0, 0};
    if (i < b1) flags_of(iwhere[i], track, f);
    unsigned long long m[4];
    for (int q = 0; q < 4; q++) {
      m[q] = __ballot(f[q] != 0);
      if (lane == 0) wc[q][wid] = (unsigned)__popcll(m[q]);
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      if (!f[q]) continue;
      unsigned r = base[q];
      for (int w = 0; w < wid; w++) r += wc[q][w];
      r += (unsigned)__popcll(m[q] & ((1ull << lane) - 1ull));
      const int v = (int)(i + 1);
      if (q == 0) index[r] = v;
      if (q == 1) index[n - 1 - r] = v;
      if (q == 2) indx2[nenter - 1 - r] = v;
      if (q == 3) indx2[n - 1 - r] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      unsigned t = 0;
      for (int w = 0; w < kWavesPerBlock; w++) t += wc[threadIdx.x][w];
      base[threadIdx.x] += t;
    }
    __syncthreads();
  }
}

void launch_export_index(hipStream_t s, const uint8_t* iwhere, long long n, int track,
                         unsigned* counts, unsigned long long* totals, int* index, int* indx2) {
  const int nblk = (int)((n + kChunk - 1) / kChunk);
  hipLaunchKernelGGL(k_export_count, dim3(nblk), dim3(kBlock), 0, s, iwhere, n, track, counts,
                     nblk);
  hipLaunchKernelGGL(s_export_scan, dim3(1), dim3(kBlock), 0, s, counts, nblk, totals);
  hipLaunchKernelGGL(k_export_write, dim3(nblk), dim3(kBlock), 0, s, iwhere, n, track, counts,
                     nblk, totals, index, indx2);
}

long long export_count_blocks(long long n) { return (n + kChunk - 1) / kChunk; }

}  // namespace lbfgsb
