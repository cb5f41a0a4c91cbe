// Device-side helpers shared by the L-BFGS-B kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbfgsb {

constexpr int kWave = 64;
constexpr int kBlock = 256;           // threads per block for bulk passes
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxM = 8;              // L-BFGS memory (the reference fixes m <= 8)
constexpr int kTile = 64;             // W is stored as [n/64][kWCols][64] tiles
// W holds kWPhys = kMaxM + 1 physical pair slots: the ring's kMaxM pairs
// and a spare that a new pair is written into before its step is accepted.
constexpr int kWPhys = kMaxM + 1;
constexpr int kWCols = 2 * kWPhys;
constexpr int kMaxBlocks = 2048;      // grid cap for the bulk passes

// Control routines that one thread runs once per step, or a workgroup once
// per large breakpoint round.  The batched kernel is one persistent
// function, so its register allocation is the largest any of its steps
// needs; it defines LBFGSB_OUTLINE_CONTROL before its includes, which keeps
// these out of line there, so that their registers and spills stay out of
// the bulk passes' loops.  Everywhere else they are inlined.
#if defined(LBFGSB_OUTLINE_CONTROL) && defined(__HIP_DEVICE_COMPILE__)
#define LBFGSB_COLD __noinline__
#else
#define LBFGSB_COLD inline
#endif

// ---------------------------------------------------------------------------
// W = [Y | S] in HBM: element i, column j (j < kWPhys: Y of physical slot
// j; j >= kWPhys: S of physical slot j - kWPhys) lives at
//   ((i / 64) * kWCols + j) * 64 + (i % 64)
// so one wave reading one column of 64 consecutive elements touches one
// contiguous 256 B (fp32) / 512 B (fp64) segment, and a new pair is written
// with fully coalesced stores.  Ring slot s of the memory is physical slot
// s ^ pslot_x[s] (DevScalars; all zero = the identity).
__host__ __device__ inline size_t widx(size_t i, int j) {
  return ((i >> 6) * kWCols + (size_t)j) * 64 + (i & 63);
}
__host__ __device__ inline size_t wsize(size_t n) { return ((n + 63) / 64) * kWCols * 64; }

// The two buffers of a double-buffered n-vector (accepted point and trial);
// kernels pick one with the device-side phase bit.  Passed by value, so the
// compiler sees kernel-argument pointers (global memory, global_load) rather
// than pointers read from memory (flat_load).
template <typename T>
struct Pair {
  T* b0;
  T* b1;
  // a select of the two values as integers: `s ? b1 : b0` may be folded into
  // a load through a selected address, which keeps the enclosing struct in
  // scratch memory
  __host__ __device__ T* operator[](int s) const {
    const uintptr_t a = (uintptr_t)b0, b = (uintptr_t)b1;
    const uintptr_t mask = (uintptr_t)0 - (uintptr_t)(s != 0);
    return (T*)(a ^ ((a ^ b) & mask));
  }
};

// Breakpoint candidate of the Cauchy search: sorted by (t, idx).
// 16-byte aligned wherever it is placed (LDS arrays of it included), so that
// an entry moves with one 128-bit LDS access: the arrays' offsets inside the
// workgroup's LDS structures do not depend on the sizes of what precedes them.
struct alignas(16) Cand {
  double t;
  int idx;
  int pad;
};

// ---------------------------------------------------------------------------
// iwhere is kept as one byte per variable:
//   bits 0..3: iwhere + 4   (values -3, -1, 0, 1, 2, 3 -> 1, 3, 4, 5, 6, 7)
//   bit 4    : free in the last free/active partition
//   bit 5    : free in the partition before it (entering/leaving sets)
//   bits 6..7: the variable's bound type nbd (0..3), copied in at set-up so
//              that the passes read one byte per variable instead of the
//              int32 nbd array as well
__device__ __host__ inline int iw_get(uint8_t b) { return (int)(b & 15) - 4; }
__device__ __host__ inline bool iw_wasfree(uint8_t b) { return (b >> 4) & 1; }
__device__ __host__ inline bool iw_prevfree(uint8_t b) { return (b >> 5) & 1; }
__device__ __host__ inline int iw_nbd(uint8_t b) { return b >> 6; }
__device__ __host__ inline uint8_t iw_make(int iw, bool freelast, bool freeprev, int nb) {
  return (uint8_t)((iw + 4) | (freelast ? 16 : 0) | (freeprev ? 32 : 0) | (nb << 6));
}

// ---------------------------------------------------------------------------
// Wave / block reductions over 64 lanes.  __shfl_xor compiles to
// ds_bpermute (the LDS crossbar); the prefix sums below use DPP moves instead
// (VALU only), where a reduction runs once per candidate or pass step.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ inline int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline int wave_or_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}
// One DPP move of a double (CTRL a DPP control, ROWS the rows of 16 lanes
// that take part): a lane whose source lane lies outside its row, or whose
// row is not in ROWS, receives 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// Inclusive prefix sum of v over the 64 lanes of a wave, by DPP moves (VALU
// only, no LDS crossbar): within each row of 16 lanes by shifts of 1, 2, 4
// and 8 (row_shr), then lane 15's running sum into the next row for rows 1
// and 3 (row_bcast:15), then lane 31's into rows 2 and 3 (row_bcast:31).
// Every lane of the wave takes part.
__device__ __forceinline__ double wave_scan_incl(double v) {
  v += dpp_f64<0x111, 0xf>(v);
  v += dpp_f64<0x112, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return v;
}
// (value, index) minimum with ties to the smallest index.
__device__ inline void wave_argmin(double& v, long long& idx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o, kWave);
    long long oi = __shfl_xor(idx, o, kWave);
    if (ov < v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// acc += a * b with the product formed in FP64.  For FP32 operands the
// product is exact in FP64, so a fused multiply-add rounds exactly like the
// separate multiply and add; FP64 operands keep the separate operations.
__device__ inline void acc_prod(double& acc, float a, float b) {
  acc = __builtin_fma((double)a, (double)b, acc);
}
__device__ inline void acc_prod(double& acc, double a, double b) { acc += a * b; }

// Error-free transformation a + b = s + e (Knuth two-sum).
__device__ __host__ inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

// Compensated sum of n doubles laid out with stride `stride`.
__device__ inline double neumaier_sum(const double* p, int n, int stride) {
  double s = 0.0, c = 0.0;
  for (int i = 0; i < n; i++) {
    double e;
    two_sum(s, p[(size_t)i * stride], s, e);
    c += e;
  }
  return s + c;
}

// Reduce NV per-thread values over a block of kBlock threads; writes the
// block totals to out[v * out_stride].  `sh` must hold kWavesPerBlock*NV
// doubles.  Deterministic (fixed tree).
template <int NV>
__device__ inline void block_sum_store(double (&vals)[NV], double* sh, double* out,
                                       size_t out_stride) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    double s = wave_sum(vals[v]);
    if (lane == 0) sh[wid * NV + v] = s;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < NV; v += blockDim.x) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; w++) s += sh[w * NV + v];
    out[(size_t)v * out_stride] = s;
  }
  __syncthreads();
}

// Marks loaded values as needed at this point.  Placed after the last load
// of an element, it keeps the compiler from sinking any of those loads into
// later conditional code, where each would be issued only after the waits
// before it: the element's loads go out together and the wave waits once.
// Emits no instruction.
__device__ inline void keep_one(float v) { asm volatile("" ::"v"(v)); }
__device__ inline void keep_one(double v) { asm volatile("" ::"v"(v)); }
__device__ inline void keep_one(int v) { asm volatile("" ::"v"(v)); }
__device__ inline void keep_one(uint8_t v) { asm volatile("" ::"v"((int)v)); }
template <typename T, int N>
__device__ inline void keep_one(const T (&v)[N]) {
#pragma unroll
  for (int k = 0; k < N; k++) keep_one(v[k]);
}
template <typename... A>
__device__ inline void keep_loads(const A&... a) {
  (keep_one(a), ...);
}

// ---------------------------------------------------------------------------
// Reading the S/Y memory of one element.
// Ring slot of each memory position j = 0..kMaxM-1 (j = 0 oldest).
__device__ inline void ring_slots(int head, int m, int (&slot)[kMaxM]) {
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    slot[j] = head + j;
    if (slot[j] >= m) slot[j] -= m;
  }
}

// Which W slots the kernel reads for each memory position j (uniform over
// the grid).  Positions that are not used (j >= col, or the slot that is
// being overwritten, skip_slot) read the slot of the first used position
// instead: an address the same wave is reading anyway, so the load hits in
// cache and the loop needs no branch around it.
// rs[j] and newest are physical slots (pslot_x maps ring slots, above).
// Physical W slot of ring slot s under the exchange table pslot_x.  A
// translation unit whose kernels never move the spare slot in (the batched
// kernel) defines LBFGSB_RING_IDENTITY before its includes: its device code
// then uses the identity, which the zero table encodes anyway, without the
// table lookups (they cost the batched kernel registers).  Host code is the
// same in every translation unit.
__host__ __device__ inline int ring_map(int s, const int* pslot_x) {
#if defined(LBFGSB_RING_IDENTITY) && defined(__HIP_DEVICE_COMPILE__)
  (void)pslot_x;
  return s;
#else
  return s ^ pslot_x[s];
#endif
}

struct RingReads {
  int rs[kMaxM];
  bool use[kMaxM];
  int newest;  // physical slot of position col - 1 (0 when the memory is empty)
};
__device__ inline RingReads ring_reads(int head, int m, int col, int skip_slot,
                                       const int* pslot_x) {
  RingReads r;
  int slot[kMaxM];
  ring_slots(head, m, slot);
  const int nl = col > 0 ? (head + col - 1) % m : 0;
  r.newest = ring_map(nl, pslot_x);
  int fb = -1;
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    r.use[j] = j < col && slot[j] != skip_slot;
    if (r.use[j] && fb < 0) fb = slot[j];
  }
  if (fb < 0) fb = 0;
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    const int s = r.use[j] ? slot[j] : fb;
    r.rs[j] = ring_map(s, pslot_x);
  }
  return r;
}

// The retained Y and S entries of element i in ring order, as loaded: an
// unused position holds the entry of the first used one.  The 16 loads are
// unconditional and issued ahead of any use, so a wave has the whole W row
// in flight at once (one 256 B segment per slot in FP32) instead of waiting
// on each pair.
template <typename T>
__device__ inline void load_ring_raw(const T* __restrict__ W, long long i, const RingReads& rr,
                                     T (&yv)[kMaxM], T (&sv)[kMaxM]) {
  const T* __restrict__ row = W + widx((size_t)i, 0);
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    yv[j] = row[rr.rs[j] * kTile];
    sv[j] = row[(kWPhys + rr.rs[j]) * kTile];
  }
}

// Zeros the unused positions of a row from load_ring_raw.  Kept apart from
// the loads so that a loop can issue the next element's loads and process
// the current one before waiting for them.
template <typename T>
__device__ inline void mask_ring(const RingReads& rr, T (&yv)[kMaxM], T (&sv)[kMaxM]) {
#pragma unroll
  for (int j = 0; j < kMaxM; j++) {
    yv[j] = rr.use[j] ? yv[j] : (T)0;
    sv[j] = rr.use[j] ? sv[j] : (T)0;
  }
}

// The retained Y and S entries of element i in ring order (0 for unused
// positions).
template <typename T>
__device__ inline void load_ring(const T* __restrict__ W, long long i, const RingReads& rr,
                                 T (&yv)[kMaxM], T (&sv)[kMaxM]) {
  load_ring_raw(W, i, rr, yv, sv);
  mask_ring(rr, yv, sv);
}

// Start of chunk c (of nchunks, C elements each) of a grid-stride pass, the
// chunks taken in ascending or (rev) descending order; the elements inside a
// chunk stay ascending, so the accesses stay coalesced.  Consecutive passes
// alternate directions: each starts where the previous one ended, on data
// that is still in the device's last-level cache.
__device__ __forceinline__ long long chunk_start(long long c, long long nchunks, int C,
                                                 bool rev) {
  return (rev ? nchunks - 1 - c : c) * (long long)C;
}

template <typename T>
__device__ inline T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Streaming store of a bulk pass's output (nontemporal hint).
template <typename T>
__device__ inline void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

}  // namespace lbfgsb
