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
//
// A large problem is solved by a gang of S workgroups, each taking one
// 64-aligned segment of its elements.  Every member runs the same control
// logic on the same reduced values (so every member takes the same
// decisions): a pass's block-level results are exchanged through HBM at a
// gang barrier and summed in member order, which makes the outcome a
// function of the problem and S only.  The breakpoint rounds of the Cauchy
// search are sorted and scanned by the gang together (gang_round).  Only
// member 0 writes the problem's results and control state back.
// the compact bmv: the kernel's registers are shared with its bulk passes
#define LBFGSB_COMPACT_BMV 1
// the ring is never exchanged with the spare slot here (pass A runs after the
// acceptance): physical W slots are the ring slots
#define LBFGSB_RING_IDENTITY 1
// formk + subsm's solves, the line-search step and the large-round sorts out
// of line (LBFGSB_COLD): 16.8k -> 19.8k solves/s (fp32) on the 4096-QP
// workload, with the same results
#define LBFGSB_OUTLINE_CONTROL 1
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
  int gflag;                 // gang error flag of the last barrier
  unsigned near_n;           // pass A's count of listed breakpoints
  BucketLds bk;              // bucket counts of a large breakpoint round
  union {
    RoundLds round;
    struct {
      double sh[kMaxRoundWaves * P4_CORR];
      double gram[256];  // the Cauchy-point pass's signed Gram
      long long shi[kMaxRoundWaves];
      double wrow[kMaxRoundWaves][kRowBuf];  // its W rows (corr_step)
    } r;
  } u;
};

template <typename T>
struct Prob {
  long long n;
  long long lo, hi;  // this workgroup's segment [lo, hi) of the elements
  Pair<T> X, G;
  const T* l;
  const T* u;
  const int* nbd;
  uint8_t* iwhere;
  T* W;
  T* d;
  T* dsub;
  Cand* ca;
  Cand* cb;
  ObjDesc obj;
};

__device__ __forceinline__ int nwaves() { return blockDim.x >> 6; }

// Entries of a problem's breakpoint list (pass A, one workgroup per
// problem): the list lives in the problem's second candidate buffer.
constexpr long long kListCap = 8192;
template <typename T>
__device__ __forceinline__ unsigned list_cap(const Prob<T>& P) {
  return (unsigned)(P.n < kListCap ? P.n : kListCap);
}

// Sums of the first NR of NV per-thread values over the workgroup into
// out[0 .. NR): each wave's sum is its DPP prefix sum's last lane (VALU
// moves, no LDS crossbar), then the waves' sums in wave order.
template <int NR, int NV>
__device__ __forceinline__ void bsum_first(const double (&v)[NV], double* sh, double* out) {
  static_assert(NR <= NV, "more sums than values");
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = nwaves();
#pragma unroll
  for (int q = 0; q < NR; q++) {
    const double s = wave_scan_incl(v[q]);
    if (lane == 63) sh[wid * NR + q] = s;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < NR; q += blockDim.x) {
    double t = 0.0;
    for (int w = 0; w < nw; w++) t += sh[w * NR + q];
    out[q] = t;
  }
  __syncthreads();
}

// Sums of NV per-thread values over the workgroup into out[0 .. NV).
template <int NV>
__device__ __forceinline__ void bsum(const double (&v)[NV], double* sh, double* out) {
  bsum_first<NV>(v, sh, out);
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
// gangs
// ON = false is the one-workgroup form: every gang step folds away at
// compile time.
template <bool ON>
struct Gang {
  int S;            // members
  int j;            // this workgroup's member index
  long long item0;  // queue entry of member 0
  GangCtl* ctl;
  double* part;     // exchange areas, 2 x kGangPart doubles per queue entry
  unsigned epoch;   // barriers passed in this launch
  __device__ __forceinline__ bool on() const { return ON && S > 1; }
};

// Barrier of the gang's workgroups (release of this workgroup's global
// writes, acquire of the others').  Returns false when the gang is in error:
// a member waited longer than kGangWaitTicks, here or before.
//
// The L2 of an XCD is not coherent with the others': every wave releases
// its own stores (write-back of the L2 after its stores completed), thread 0
// arrives and polls the counter with relaxed agent-scope loads (which read
// the coherent copy without invalidating the L2 under every workgroup of
// the XCD at each poll), and one wave invalidates the CU's and the XCD's
// caches once the gang has arrived.
__device__ __forceinline__ bool gang_sync(Gang<true>& G, int* flag) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&G.ctl->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (G.epoch + 1u) * (unsigned)G.S;
    const long long t0 = (long long)wall_clock64();
    int err = 0;
    while (__hip_atomic_load(&G.ctl->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           target) {
      err = __hip_atomic_load(&G.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (err) break;
      if ((long long)wall_clock64() - t0 > kGangWaitTicks) {
        __hip_atomic_store(&G.ctl->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        err = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
    if (!err) err = __hip_atomic_load(&G.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *flag = err;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  G.epoch++;
  const bool ok = *flag == 0;
  __syncthreads();
  return ok;
}

// Combine a pass's block-level results over the gang: red[0 .. nsum) are
// summed in member order, rv[k] is combined by mode[k] (0 none, 1 max,
// 2 min, 3 argmin with index ri[0], 4 argmin with index ri[1]), and the
// Part-4 flag of entering/leaving variables is or-ed.  A gang error ends
// the problem (info -2).
template <bool ON>
__device__ __forceinline__ void gang_combine(Gang<ON>& G, BatchLds& L, int nsum, int m0, int m1,
                                             int m2, int m3, int* flag) {
  if constexpr (!ON) {
    return;
  } else {
  if (!G.on()) return;
  const unsigned par = G.epoch & 1u;
  double* mine = G.part + ((G.item0 + G.j) * 2 + par) * kGangPart;
  for (int q = threadIdx.x; q < nsum; q += blockDim.x) mine[q] = L.red[q];
  if (threadIdx.x < 4) mine[kGangPart - 8 + threadIdx.x] = L.rv[threadIdx.x];
  if (threadIdx.x < 2) mine[kGangPart - 4 + threadIdx.x] = __longlong_as_double(L.ri[threadIdx.x]);
  if (threadIdx.x == 0) mine[kGangPart - 2] = (double)L.s.p4_flagged;
  if (!gang_sync(G, flag)) {
    if (threadIdx.x == 0) set_done(&L.s, -2);
    __syncthreads();
    return;
  }
  const double* base = G.part + (G.item0 * 2 + par) * kGangPart;
  const long long stride = 2LL * kGangPart;
  for (int q = threadIdx.x; q < nsum; q += blockDim.x) {
    double t = 0.0;
    for (int k = 0; k < G.S; k++) t += base[k * stride + q];
    L.red[q] = t;
  }
  if (threadIdx.x == 0) {
    const int mode[4] = {m0, m1, m2, m3};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (mode[c] == 0) continue;
      double v = base[kGangPart - 8 + c];
      const int ix = mode[c] - 3;
      long long id = mode[c] >= 3 ? __double_as_longlong(base[kGangPart - 4 + ix]) : 0;
      for (int k = 1; k < G.S; k++) {
        const double w = base[k * stride + kGangPart - 8 + c];
        if (mode[c] == 1) {
          v = fmax(v, w);
        } else if (mode[c] == 2) {
          v = fmin(v, w);
        } else {
          const long long wi = __double_as_longlong(base[k * stride + kGangPart - 4 + ix]);
          if (w < v || (w == v && wi < id)) {
            v = w;
            id = wi;
          }
        }
      }
      L.rv[c] = v;
      if (mode[c] >= 3) L.ri[ix] = id;
    }
    int fl = 0;
    for (int k = 0; k < G.S; k++) fl |= base[k * stride + kGangPart - 2] != 0.0;
    L.s.p4_flagged = fl;
  }
  __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// A workgroup is the only one on its CU (the kernel's register allocation
// leaves two waves per SIMD), so a pass that reads few streams per element
// and takes one element per thread at a time waits a full memory latency
// for every few bytes.  The light passes (the subspace point, g'd, the
// breakpoint selection, the objective) therefore take kLightU elements per
// thread per step, i0 + threadIdx.x + k * blockDim.x for k < kLightU
// (coalesced for every k, and in the order a one-element loop visits them,
// so every per-thread sum is unchanged), and issue the loads of all of them
// before processing any.  The heavy passes (pass A, the Cauchy-point pass,
// the subspace direction) keep one element per step: their accumulators
// fill the registers.
constexpr int kLightU = 4;

template <int U, typename Op, typename F>
__device__ __forceinline__ void batch_pass(Op& op, long long lo, long long hi, F&& body) {
  const long long nt = blockDim.x;
  for (long long i0 = lo + threadIdx.x; i0 < hi; i0 += U * nt) {
    typename Op::In e[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      // lanes past the end load element hi - 1 (a valid address) and skip it
      const long long i = i0 + k * nt;
      op.load(e[k], i < hi ? i : hi - 1);
    }
#pragma unroll
    for (int k = 0; k < U; k++) op.keep(e[k]);
#pragma unroll
    for (int k = 0; k < U; k++)
      if (i0 + k * nt < hi) body(i0 + k * nt, e[k]);
  }
}

// ---------------------------------------------------------------------------
// set-up passes (lbfgsbcpu.cpp:258-278)
template <typename T, bool GO>
__device__ __forceinline__ void b_errclb(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  double v[2] = {0.0, 0.0};
  for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x) {
    const int nb = P.nbd[i];
    if (nb < 0 || nb > 3) v[0] += 1.0;
    if (nb == 2 && P.l[i] > P.u[i]) v[1] += 1.0;
  }
  bsum<2>(v, L.u.r.sh, L.red);
  gang_combine(G, L, 2, 0, 0, 0, 0, &L.gflag);
  if (threadIdx.x == 0) logic_setup_errclb(&L.s, L.red[0], L.red[1]);
  __syncthreads();
}

template <typename T, bool GO>
__device__ __forceinline__ void b_active(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  T* x = P.X[0];
  double v[3] = {0.0, 0.0, 0.0};
  for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x) {
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
    P.iwhere[i] = iw_make(iw, true, true, nb & 3);
  }
  bsum<3>(v, L.u.r.sh, L.red);
  gang_combine(G, L, 3, 0, 0, 0, 0, &L.gflag);
  if (threadIdx.x == 0) {
    const double r[3] = {L.red[0], L.red[1], L.red[2]};
    logic_setup_active(&L.s, r);
  }
  __syncthreads();
}

// This thread's share of f over terms [k0, k1) of an objective of kind
// KIND, with the gradient stored: kLightU terms per thread per step, every
// one evaluated before the first gradient store (the terms' loads go out
// together), added to the sum in the order a one-term loop adds them.  Out
// of line, like the control routines: inlined, the five kinds' loops shared
// the kernel's register allocation (20.1k against 18.6k solves/s, fp32).
template <int KIND, typename T>
__device__ __noinline__ double obj_terms_sum(const ObjDesc& o, long long n, const T* x, T* g,
                                                long long k0, long long k1) {
  const long long nt = blockDim.x;
  double acc = 0.0;
  for (long long kb = k0 + threadIdx.x; kb < k1; kb += kLightU * nt) {
    ObjOut<T> r[kLightU];
    double fk[kLightU];
#pragma unroll
    for (int q = 0; q < kLightU; q++) {
      const long long k = kb + q * nt;
      fk[q] = obj_eval_kind<KIND>(o, n, x, k < k1 ? k : k1 - 1, r[q]);
    }
#pragma unroll
    for (int q = 0; q < kLightU; q++) {
      const long long k = kb + q * nt;
      if (k < k1) {
        obj_store_kind<KIND>(g, k, r[q]);
        acc += fk[q];
      }
    }
  }
  return acc;
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  typedef float4 type;
};
template <>
struct Vec4<double> {
  typedef double4 type;
};

// The pentadiagonal terms [k0, k1) (k0 a multiple of four), four consecutive
// terms per thread: x, a1 and a2 at a term's neighbours come from the same
// 4-vector loads as its own entries (one load of each vector instead of
// eleven scalar loads per term; the arrays of a problem start 64-element
// aligned and run to a multiple of 64).  Each term's gradient and f share
// are formed as obj_eval_kind forms them, in the same order.
template <typename T>
__device__ __noinline__ double obj_penta_sum4(const ObjDesc& o, long long n,
                                              const T* __restrict__ x, T* __restrict__ g,
                                              long long k0, long long k1) {
  using V = typename Vec4<T>::type;
  const V* __restrict__ xv = (const V*)x;
  const V* __restrict__ a0v = (const V*)o.a0;
  const V* __restrict__ a1v = (const V*)o.a1;
  const V* __restrict__ a2v = (const V*)o.a2;
  const V* __restrict__ bv = (const V*)o.b;
  const V zero = {(T)0, (T)0, (T)0, (T)0};
  double acc = 0.0;
  const long long j1 = (k1 + 3) / 4;
  for (long long j = k0 / 4 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long k = 4 * j;
    // the neighbour vectors only where they lie inside the problem
    const bool prev = k >= 4, next = k + 4 < n;
    const V xc = xv[j], xm = prev ? xv[j - 1] : zero, xp = next ? xv[j + 1] : zero;
    const V A0 = a0v[j], A1 = a1v[j], A2 = a2v[j], B = bv[j];
    const V A1m = prev ? a1v[j - 1] : zero, A2m = prev ? a2v[j - 1] : zero;
    // x[k-2 .. k+5], a1[k-1 .. k+3], a2[k-2 .. k+3]
    const T xs[8] = {xm.z, xm.w, xc.x, xc.y, xc.z, xc.w, xp.x, xp.y};
    const T a1s[5] = {A1m.w, A1.x, A1.y, A1.z, A1.w};
    const T a2s[6] = {A2m.z, A2m.w, A2.x, A2.y, A2.z, A2.w};
    const T a0s[4] = {A0.x, A0.y, A0.z, A0.w};
    const T bs[4] = {B.x, B.y, B.z, B.w};
    T go[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const long long ke = k + e;
      const double xi = (double)xs[2 + e];
      double ax = (double)a0s[e] * xi;
      if (ke + 1 < n) ax += (double)a1s[1 + e] * (double)xs[3 + e];
      if (ke + 2 < n) ax += (double)a2s[2 + e] * (double)xs[4 + e];
      if (ke >= 1) ax += (double)a1s[e] * (double)xs[1 + e];
      if (ke >= 2) ax += (double)a2s[e] * (double)xs[e];
      go[e] = (T)(ax - (double)bs[e]);
      if (ke < k1) acc += 0.5 * ax * xi - (double)bs[e] * xi;
    }
    if (k + 4 <= k1) {
      V gv;
      gv.x = go[0];
      gv.y = go[1];
      gv.z = go[2];
      gv.w = go[3];
      reinterpret_cast<V*>(g)[j] = gv;
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (k + e < k1) g[k + e] = go[e];
    }
  }
  return acc;
}

// The grid terms [k0, k1) (k0 a multiple of four) of a GRID_QUAD or GRID_COS
// objective, four consecutive points of a row-major grid per thread: the
// points' own x, b and c as 4-vectors, the left and right neighbours from
// the same vector plus one load at each end, the rows below and above as
// 4-vectors where rows start 4-aligned (nx a multiple of four), else one
// load per point.  Each term's gradient and f share are formed as
// obj_eval_kind forms them, in the same order.
template <int KIND, typename T>
__device__ __noinline__ double obj_grid_sum4(const ObjDesc& o, long long n,
                                             const T* __restrict__ x, T* __restrict__ g,
                                             long long k0, long long k1) {
  using V = typename Vec4<T>::type;
  const V* __restrict__ xv = (const V*)x;
  const V* __restrict__ bv = (const V*)o.b;
  const V* __restrict__ cv = (const V*)o.c;
  const long long nx = o.nx, ny = o.ny;
  const double h2 = o.h * o.h;
  const bool rows4 = (nx & 3) == 0;
  const V zero = {(T)0, (T)0, (T)0, (T)0};
  double acc = 0.0;
  const long long j1 = (k1 + 3) / 4;
  for (long long j = k0 / 4 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long k = 4 * j;
    const V xc = xv[j], B = bv[j];
    V C = zero;
    if constexpr (KIND == OBJK_GRID_COS) C = cv[j];
    // x[k-1 .. k+4]
    const T xs[6] = {k >= 1 ? x[k - 1] : (T)0, xc.x, xc.y, xc.z, xc.w, k + 4 < n ? x[k + 4] : (T)0};
    T dn[4], up[4];
    if (rows4) {
      const V d = k >= nx ? xv[(k - nx) / 4] : zero;
      const V u = k + nx < n ? xv[(k + nx) / 4] : zero;
      dn[0] = d.x, dn[1] = d.y, dn[2] = d.z, dn[3] = d.w;
      up[0] = u.x, up[1] = u.y, up[2] = u.z, up[3] = u.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        dn[e] = k + e >= nx && k + e < n ? x[k + e - nx] : (T)0;
        up[e] = k + e + nx < n ? x[k + e + nx] : (T)0;
      }
    }
    const T bs[4] = {B.x, B.y, B.z, B.w};
    const T cs[4] = {C.x, C.y, C.z, C.w};
    long long ix, iy;
    grid_pos(k, n, nx, ix, iy);
    T go[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (e > 0) ix++;
      while (ix >= nx) {
        ix -= nx;
        iy++;
      }
      const double v = (double)xs[1 + e];
      const double vl = ix > 0 ? (double)xs[e] : 0.0;
      const double vr = ix < nx - 1 ? (double)xs[2 + e] : 0.0;
      const double vd = iy > 0 ? (double)dn[e] : 0.0;
      const double vu = iy < ny - 1 ? (double)up[e] : 0.0;
      const double bk = (double)bs[e];
      double gk = (4.0 * v - vl - vr - vd - vu) - h2 * bk;
      const double er = vr - v, eu = vu - v;
      double ee = er * er + eu * eu;
      if (ix == 0) ee += v * v;
      if (iy == 0) ee += v * v;
      double fk = 0.5 * ee - h2 * bk * v;
      if constexpr (KIND == OBJK_GRID_COS) {
        const double ck = (double)cs[e];
        fk += h2 * ck * cos(v);
        gk -= h2 * ck * sin(v);
      }
      go[e] = (T)gk;
      if (k + e < k1) acc += fk;
    }
    if (k + 4 <= k1) {
      V gv;
      gv.x = go[0];
      gv.y = go[1];
      gv.z = go[2];
      gv.w = go[3];
      reinterpret_cast<V*>(g)[j] = gv;
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (k + e < k1) g[k + e] = go[e];
    }
  }
  return acc;
}

// Objective at X[which] into G[which]; L.f = f rounded to the bulk type, as
// a callback returning `real& f` would hand it over.
// In a gang, member j evaluates the terms of its segment (ROSEN: the pairs
// inside it, segments being even-aligned) after a barrier, since the
// other kinds read neighbours that other members have just written.
template <typename T, bool GO>
__device__ __forceinline__ void b_objective(const Prob<T>& P, BatchLds& L, int which,
                                            Gang<GO>& G) {
  const T* x = P.X[which];
  T* g = P.G[which];
  if constexpr (GO) {
    if (G.on() && P.obj.kind != OBJK_ROSEN && !gang_sync(G, &L.gflag)) {
      if (threadIdx.x == 0) set_done(&L.s, -2);
      __syncthreads();
    }
  }
  double acc[1] = {0.0};
  const bool pairs = P.obj.kind == OBJK_ROSEN;
  const long long k0 = pairs ? P.lo / 2 : P.lo;
  const long long k1 = pairs ? P.hi / 2 : P.hi;
  // one loop per kind: its body is straight-line code, so the terms' loads
  // can go out together
  switch (P.obj.kind) {
    case OBJK_GRID_QUAD:
      acc[0] = (k0 & 3) == 0 ? obj_grid_sum4<OBJK_GRID_QUAD, T>(P.obj, P.n, x, g, k0, k1)
                             : obj_terms_sum<OBJK_GRID_QUAD>(P.obj, P.n, x, g, k0, k1);
      break;
    case OBJK_GRID_COS:
      acc[0] = (k0 & 3) == 0 ? obj_grid_sum4<OBJK_GRID_COS, T>(P.obj, P.n, x, g, k0, k1)
                             : obj_terms_sum<OBJK_GRID_COS>(P.obj, P.n, x, g, k0, k1);
      break;
    case OBJK_DSSCFG:
      acc[0] = obj_terms_sum<OBJK_DSSCFG>(P.obj, P.n, x, g, k0, k1);
      break;
    case OBJK_ROSEN:
      acc[0] = obj_terms_sum<OBJK_ROSEN>(P.obj, P.n, x, g, k0, k1);
      break;
    default:
      acc[0] = (k0 & 3) == 0 ? obj_penta_sum4<T>(P.obj, P.n, x, g, k0, k1)
                             : obj_terms_sum<OBJK_PENTA>(P.obj, P.n, x, g, k0, k1);
      break;
  }
  bsum<1>(acc, L.u.r.sh, L.red);
  gang_combine(G, L, 1, 0, 0, 0, 0, &L.gflag);
  if (threadIdx.x == 0) L.f = (double)(T)L.red[0];
  __syncthreads();
}

template <typename T, bool GO>
__device__ __forceinline__ void b_projgr(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  const T* x = P.X[0];
  const T* g = P.G[0];
  double m = 0.0;
  for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x)
    m = fmax(m, (double)fabs(projected_gi(x[i], g[i], P.l[i], P.u[i], P.nbd[i])));
  m = bmax(m, L.u.r.sh);
  if (threadIdx.x == 0) L.rv[0] = m;
  __syncthreads();
  gang_combine(G, L, 0, 1, 0, 0, 0, &L.gflag);
  if (threadIdx.x == 0) logic_setup_projgr(&L.s, L.rv[0], L.f);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// iteration passes
template <typename T, bool FUSED, bool GO>
__device__ __forceinline__ void b_pass_a(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  if (L.s.halt) return;
  using Op = PassAOp<T, FUSED>;
  Op op;
  // one workgroup per problem: the breakpoints below the predicted bound go
  // to the problem's list in cb (the class is then not stored: the search
  // and the Cauchy-point pass reclassify from the stored one); a gang
  // stores the class and selects from the variables
  Cand* lst = nullptr;
  const unsigned cap = list_cap(P);
  if constexpr (!GO) {
    if (L.s.gcp_near_tau > 0.0) lst = P.cb;
    if (threadIdx.x == 0) L.near_n = 0u;
    __syncthreads();
  }
  op.init(&L.s, P.X, P.G, P.l, P.u, P.iwhere, P.W, P.d, lst, &L.near_n, cap);
  for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x) {
    typename Op::In e;
    op.load(e, i);
    op.keep(e);
    op.element(i, e);
  }
  op.finish();
  // rows 0 .. PA_GD - 1: the g'd row of the plain engine's speculative pass
  // is not formed here (the line search has its own g'd pass)
  bsum_first<PA_GD>(op.acc, L.u.r.sh, L.red);
  const double sbg = bmax(op.sbg, L.u.r.sh);
  double bk = op.bkmin;
  long long ibk = op.ibk;
  bargmin(bk, ibk, L.u.r.sh, L.u.r.shi);
  if (threadIdx.x == 0) {
    L.rv[0] = sbg;
    L.rv[1] = bk;
    L.ri[0] = ibk;
    L.red[PA_GD] = 0.0;
    L.s.gcp_list_n = lst ? (long long)L.near_n : -1;
  }
  __syncthreads();
  gang_combine(G, L, PA_GD, 1, 3, 0, 0, &L.gflag);
}

template <typename T, bool GO>
__device__ __forceinline__ void b_p4(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  if (L.s.halt) return;
  P4Op<T> op;
  op.init(&L.s, P.X, P.G, P.l, P.u, P.iwhere, P.W);
  const long long hi = P.hi;
  double* rowbuf = L.u.r.wrow[threadIdx.x >> 6];
  for (long long i0 = P.lo; i0 < hi; i0 += blockDim.x) {
    const long long i = i0 + threadIdx.x;
    typename P4Op<T>::In e;
    op.load(e, i < hi ? i : hi - 1);
    op.keep(e);
    op.corr_step(i, i < hi ? op.element(i, e) : 0, e, rowbuf);
  }
  bsum<P4_CORR>(op.acc, L.u.r.sh, L.red);
  const bool tracked = op.track && op.col > 0;
  if (tracked) op.corr_block_sum(L.u.r.gram, nwaves());
  __syncthreads();
  const int upcl = L.s.updatd ? op.col - 1 : op.col;
  for (int t = threadIdx.x; t < 136; t += blockDim.x)
    L.red[P4_CORR + t] = tracked ? corr_pair_value(t, L.u.r.gram, upcl) : 0.0;
  __syncthreads();
  gang_combine(G, L, P4_COUNT, 0, 0, 0, 0, &L.gflag);
}

template <typename T, bool GO>
__device__ __forceinline__ void b_p5a(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  if (L.s.halt || L.s.skip_subsm) return;
  P5aOp<T> op;
  op.init(&L.s, P.X, P.G, P.l, P.u, P.iwhere, P.W, P.dsub);
  for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x) {
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
  gang_combine(G, L, 0, 0, 0, 4, 0, &L.gflag);
}

template <typename T, bool GO>
__device__ __forceinline__ void b_p5b(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  if (L.s.halt) return;
  P5bOp<T> op;
  op.init(&L.s, P.X, P.G, P.l, P.u, P.iwhere, P.dsub);
  batch_pass<kLightU>(op, P.lo, P.hi,
                      [&](long long i, typename P5bOp<T>::In& e) { op.element(i, e); });
  bsum<P5_COUNT>(op.acc, L.u.r.sh, L.red);
  const double smin = bmin(op.smin, L.u.r.sh);
  if (threadIdx.x == 0) L.rv[3] = smin;
  __syncthreads();
  gang_combine(G, L, P5_COUNT, 0, 0, 0, 2, &L.gflag);
}

// Trial point x_b = stp d + x_a, or the subspace point z when stp == 1
// (lbfgsbcpu.cpp:1329-1335).  The subspace pass leaves z in the trial
// buffer, so a unit step only gets here (write_trial) when a trial since has
// overwritten it; z is then formed again from the state it came from.  The
// first trial off z stores d = z - x (trial_element).
template <typename T>
__device__ __forceinline__ void b_trial(const Prob<T>& P, BatchLds& L) {
  if (!L.s.write_trial) return;
  const T* xa = P.X[L.s.trial_src];
  T* xb = P.X[L.s.trial_dst];
  const double stp = L.s.stp;
  const bool trial_d = L.s.trial_d;
  P5bOp<T> sub;
  sub.init(&L.s, P.X, P.G, P.l, P.u, P.iwhere, P.dsub, L.s.trial_src);
  for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x)
    trial_element(i, &L.s, sub, xa, xb, P.d, stp, trial_d);
  __syncthreads();
}

// g_b . d after an evaluation (lbfgsbcpu.cpp:1312).
template <typename T, bool GO>
__device__ __forceinline__ void b_gd(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  const T* gb = P.G[(L.s.phase & 1) ^ 1];
  LsDirection<T> dir;
  dir.init(&L.s, P.X, P.d);
  double acc[1] = {0.0};
  const long long nt = blockDim.x, hi = P.hi;
  for (long long i0 = P.lo + threadIdx.x; i0 < hi; i0 += kLightU * nt) {
    T gv[kLightU], dv[kLightU];
#pragma unroll
    for (int k = 0; k < kLightU; k++) {
      const long long i = i0 + k * nt < hi ? i0 + k * nt : hi - 1;
      gv[k] = gb[i];
      dv[k] = dir.at(i);
    }
#pragma unroll
    for (int k = 0; k < kLightU; k++)
      if (i0 + k * nt < hi) acc[0] += (double)gv[k] * (double)dv[k];
  }
  bsum<1>(acc, L.u.r.sh, L.red);
  gang_combine(G, L, 1, 0, 0, 0, 0, &L.gflag);
}

// The round's candidates from pass A's list (cb, one workgroup per problem)
// when its t-range lies below the bound the list was made under and the
// list did not overflow: the entries in [tau_lo, tau_hi) into P.ca, as the
// selection from the variables would give them (the same t, in another
// order: the round sorts them).  Otherwise (false) the bound adapts as in
// the plain form (x4 for a range beyond it, /4 after an overflow) and the
// caller selects from the variables.
template <typename T>
__device__ LBFGSB_COLD bool b_select_listed(const Prob<T>& P, BatchLds& L) {
  DevScalars* s = &L.s;
  const long long ln = s->gcp_list_n;
  const double ntau = s->gcp_near_tau;
  const bool covers = ln >= 0 && ntau > 0.0 && s->gcp_tau_hi <= ntau && ln <= (long long)list_cap(P);
  __syncthreads();
  if (!covers) {
    if (threadIdx.x == 0 && ln >= 0 && ntau > 0.0) {
      s->gcp_near_misses++;
      const double gr = s->gcp_near_grow > 0.0 ? s->gcp_near_grow : kNearGrow0;
      s->gcp_near_grow = ln > (long long)list_cap(P) ? fmax(gr * 0.25, 1.0) : fmin(gr * 4.0, 1024.0);
    }
    if (threadIdx.x == 0) s->gcp_sel_near = 0;
    __syncthreads();
    return false;
  }
  const double lo = s->gcp_tau_lo, hi = s->gcp_tau_hi;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) {
    s->gcp_need_select = 0;
    s->gcp_sel_near = 1;
    s->gcp_near_listed = ln;
  }
  for (long long j0 = 0; j0 < ln; j0 += blockDim.x) {
    const long long j = j0 + threadIdx.x;
    Cand c;
    c.t = -1.0;
    if (j < ln) c = P.cb[j];
    const bool take = j < ln && c.t >= lo && c.t < hi;
    const unsigned long long mask = __ballot(take);
    if (mask) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(&s->cand_count, (unsigned long long)__popcll(mask));
      base = __shfl(base, 0, kWave);
      const unsigned long long slot = base + __popcll(mask & ((1ull << lane) - 1ull));
      if (take && slot < (unsigned long long)P.n) P.ca[slot] = c;
    }
  }
  __syncthreads();
  return true;
}

template <typename T, bool GO>
__device__ __forceinline__ void b_select(const Prob<T>& P, BatchLds& L, Gang<GO>& G);

// b_select out of line: the fp32 one-workgroup kernel's register allocation
// came out better with it (4096-QP workload 31.3k -> 33.0k solves/s, the
// 512-problem grid workload 564 -> 611), the fp64 one's worse (24.3k ->
// 23.7k), so fp64 keeps it inline.
template <typename T, bool GO>
__device__ LBFGSB_COLD void b_select_cold(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  b_select(P, L, G);
}

// Breakpoint candidates with t in [tau_lo, tau_hi) of this workgroup's
// segment into P.ca (slots from the LDS count, or in a gang from the gang's
// HBM count).
template <typename T, bool GO>
__device__ __forceinline__ void b_select(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  DevScalars* s = &L.s;
  const int cur = s->phase & 1;
  const T* __restrict__ x = P.X[cur];
  const T* __restrict__ g = P.G[cur];
  const double lo = s->gcp_tau_lo, hi = s->gcp_tau_hi;
  const int lane = threadIdx.x & 63;
  __syncthreads();
  if (threadIdx.x == 0) s->gcp_need_select = 0;
  const long long nt = blockDim.x, ph = P.hi;
  for (long long c0 = P.lo; c0 < ph; c0 += kLightU * nt) {
    // the loads of kLightU elements per thread first
    T xv[kLightU], gv[kLightU], lv[kLightU], uv[kLightU];
    uint8_t bv[kLightU];
#pragma unroll
    for (int q = 0; q < kLightU; q++) {
      const long long i = c0 + threadIdx.x + q * nt < ph ? c0 + threadIdx.x + q * nt : ph - 1;
      xv[q] = x[i];
      gv[q] = g[i];
      lv[q] = P.l[i];
      uv[q] = P.u[i];
      bv[q] = P.iwhere[i];
    }
#pragma unroll
    for (int q = 0; q < kLightU; q++) {
      if (c0 + q * nt >= ph) break;  // uniform: the whole run is past the end
      const long long i = c0 + threadIdx.x + q * nt;
      bool take = false;
      T tb = 0;
      const uint8_t iwb = bv[q];
      if (i < ph) {
        // the stored class is pass A's or the one before it (classify is
        // idempotent): the t pass A forms
        int iw;
        T neggi;
        bool isfree, isbrk;
        classify(xv[q], gv[q], lv[q], uv[q], iw_nbd(iwb), iw_get(iwb), iw, neggi, isfree, isbrk, tb);
        const double t = (double)tb;
        take = isbrk && t >= lo && t < hi;
      }
      const unsigned long long mask = __ballot(take);
      if (mask) {
        unsigned long long base = 0;
        if (lane == 0) {
          const unsigned long long cnt = (unsigned long long)__popcll(mask);
          if constexpr (GO)
            base = atomicAdd(&G.ctl->cand, cnt);
          else
            base = atomicAdd(&s->cand_count, cnt);
        }
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
  }
  __syncthreads();
}

// What the breakpoint rounds read of a problem.
template <typename T>
struct RoundArgs {
  long long n;
  const T* g;
  const T* W;
  Cand* ca;
  Cand* cb;
};

template <typename T>
__device__ __forceinline__ RoundArgs<T> round_args(const Prob<T>& P, const DevScalars* s) {
  return RoundArgs<T>{P.n, P.G[s->phase & 1], P.W, P.ca, P.cb};
}

// One breakpoint round over the candidates in ca (count in the control
// state): sort in LDS or HBM, then the closed-form scan.
template <typename T>
__device__ __forceinline__ void b_round(RoundArgs<T> P, BatchLds& L) {
  DevScalars* s = &L.s;
  const long long cnt = (long long)s->cand_count;
  const long long K = cnt < P.n ? cnt : P.n;
  stage_gcp_const(s, L.u.round.gc);
  const T* g = P.g;
  if (K <= kSmallRound) {
    block_sort_small(P.ca, K, L.u.round.sc);
    block_round_scan(s, L.u.round, L.u.round.sc, K, g, P.W);
  } else if (!block_round_bucketed(s, L.u.round, L.bk, P.ca, P.cb, K, g, P.W)) {
    const Cand* sorted = block_sort_large(P.ca, P.cb, K, L.u.round.sc);
    block_round_scan(s, L.u.round, sorted, K, g, P.W);
  }
  // the larger rounds work through cb, which held pass A's list
  if (K > kSmallRound && threadIdx.x == 0) s->gcp_list_n = -1;
  if (s->gcp_pending) {
    // every candidate of the t-range passed: the next range
    if (threadIdx.x == 0) logic_resume(s, 0);
    __syncthreads();
  }
}

// Copy of the control state between LDS and HBM.
__device__ __forceinline__ void ds_copy(unsigned long long* dst, const unsigned long long* src) {
  for (int i = threadIdx.x; i < kBDsWords; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// Boundary of chunk c (0 <= c <= S) when K positions are split into S
// chunks as evenly as possible.
__device__ __forceinline__ long long chunk_at(long long K, int S, int c) {
  return K * (long long)c / S;
}

// The chunk that holds position o (0 <= o < K).
__device__ __forceinline__ int chunk_of(long long K, int S, long long o) {
  int c = (int)((o * S) / K);
  while (c + 1 < S && chunk_at(K, S, c + 1) <= o) c++;
  while (c > 0 && chunk_at(K, S, c) > o) c--;
  return c;
}

// Sort a[0 .. K) over the gang: member j sorts chunk j in place (LDS, or
// LDS runs merged through b), then ceil(log2 S) merge-path passes between
// a and b, each merging pairs of runs of 2^p chunks with the outputs split
// evenly among the members.  Returns the buffer that holds the result.
__device__ __forceinline__ Cand* gang_sort(Cand* a, Cand* b, long long K, Cand* sc,
                                           Gang<true>& G, int* flag, bool& ok) {
  const int S = G.S;
  const long long lo = chunk_at(K, S, G.j), hi = chunk_at(K, S, G.j + 1), len = hi - lo;
  if (len > 0) {
    if (len <= kSmallRound) {
      block_sort_small(a + lo, len, sc);
      for (long long i = threadIdx.x; i < len; i += blockDim.x) a[lo + i] = sc[i];
    } else {
      Cand* r = block_sort_large(a + lo, b + lo, len, sc);
      if (r != a + lo)
        for (long long i = threadIdx.x; i < len; i += blockDim.x) a[lo + i] = r[i];
    }
  }
  ok = gang_sync(G, flag);
  Cand* src = a;
  Cand* dst = b;
  for (int w = 1; ok && w < S; w <<= 1) {
    const long long per = (len + blockDim.x - 1) / blockDim.x;
    const long long o0 = lo + (long long)threadIdx.x * per;
    const long long oe = o0 + per < hi ? o0 + per : hi;
    long long o = o0;
    while (o < oe) {
      // output position o lies in the pair of runs starting at chunk cb
      const int cb = (chunk_of(K, S, o) / (2 * w)) * 2 * w;
      const long long a0 = chunk_at(K, S, cb);
      const long long a1 = chunk_at(K, S, cb + w < S ? cb + w : S);
      const long long b0 = a1;
      const long long b1 = chunk_at(K, S, cb + 2 * w < S ? cb + 2 * w : S);
      const long long la = a1 - a0, lb = b1 - b0;
      const long long diag = o - a0;
      long long l = diag - lb > 0 ? diag - lb : 0, h = diag < la ? diag : la;
      while (l < h) {
        const long long mid = (l + h) >> 1;
        if (cand_less(src[a0 + mid], src[b0 + diag - 1 - mid]))
          l = mid + 1;
        else
          h = mid;
      }
      long long i = l, j = diag - l;
      const long long oend = oe < b1 ? oe : b1;
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
    ok = gang_sync(G, flag);
    Cand* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

// One breakpoint round over the gang.  The candidates of the t-range (count
// in the gang's HBM counter) are sorted over the gang; member j scans the
// j-th chunk of the sorted sequence from the state before it (the round's
// carry plus the chunk totals of members 0 .. j-1, summed in member order);
// the first stop of the gang, with its state, concludes the round in every
// member alike.  Returns the gang's barrier count after the round.
template <typename T>
__device__ __forceinline__ unsigned gang_round(RoundArgs<T> P, BatchLds& L, Gang<true> G) {
  DevScalars* s = &L.s;
  RoundLds& R = L.u.round;
  const int S = G.S;
  const long long cnt =
      (long long)__hip_atomic_load(&G.ctl->cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const long long K = cnt < P.n ? cnt : P.n;
  stage_gcp_const(s, R.gc);
  const T* g = P.g;
  bool ok = true;
  const Cand* sorted = gang_sort(P.ca, P.cb, K, R.sc, G, &L.gflag, ok);
  // every member has read the count: it starts from zero in the next round
  // (no member adds to it before the barriers below)
  if (ok && G.j == 0 && threadIdx.x == 0)
    __hip_atomic_store(&G.ctl->cand, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const long long lo = chunk_at(K, S, G.j), hi = chunk_at(K, S, G.j + 1);
  const long long stride = 2LL * kGangPart;
  // chunk totals
  if (ok) {
    if (threadIdx.x < kNV) R.scarry[threadIdx.x] = 0.0;
    block_scan_range(s, R, sorted, lo, hi, K, false, g, P.W);
    const unsigned par = G.epoch & 1u;
    double* mine = G.part + ((G.item0 + G.j) * 2 + par) * kGangPart;
    if (threadIdx.x < kNV) mine[threadIdx.x] = R.scarry[threadIdx.x];
    ok = gang_sync(G, &L.gflag);
    const double* base = G.part + (G.item0 * 2 + par) * kGangPart;
    if (ok && threadIdx.x < kNV) {
      double c = s->gcp_carry[threadIdx.x];
      for (int k = 0; k < G.j; k++) c += base[k * stride + threadIdx.x];
      R.scarry[threadIdx.x] = c;
      // the inclusive total of the round, summed alike in every member
      double t = s->gcp_carry[threadIdx.x];
      for (int k = 0; k < S; k++) t += base[k * stride + threadIdx.x];
      R.swave[threadIdx.x] = t;
    }
    __syncthreads();
  }
  // the scan of this chunk and the exchange of the stops
  if (ok) {
    const int q = threadIdx.x;
    double total = q < kNV ? R.swave[q] : 0.0;
    block_scan_range(s, R, sorted, lo, hi, K, true, g, P.W);
    const unsigned par = G.epoch & 1u;
    double* mine = G.part + ((G.item0 + G.j) * 2 + par) * kGangPart;
    if (q == 0) mine[0] = __longlong_as_double(R.sstop);
    if (q < kNV) {
      mine[1 + q] = R.sstop_state[q];
      mine[1 + kNV + q] = R.slast[q];
    }
    ok = gang_sync(G, &L.gflag);
    if (ok) {
      const double* base = G.part + (G.item0 * 2 + par) * kGangPart;
      long long stop = 0x7fffffffffffffffLL;
      int owner = -1;
      for (int k = 0; k < S; k++) {
        const long long sk = __double_as_longlong(base[k * stride]);
        if (sk < stop) {
          stop = sk;
          owner = k;
        }
      }
      const int last_owner = K > 0 ? chunk_of(K, S, K - 1) : 0;
      if (q < kNV) {
        R.sstop_state[q] = owner >= 0 ? base[owner * stride + 1 + q] : total;
        R.slast[q] = K > 0 ? base[last_owner * stride + 1 + kNV + q] : s->gcp_carry[q];
      }
      if (q == 0) R.sstop = stop;
      __syncthreads();
      if (q == 0) {
        const bool stopped = R.sstop != 0x7fffffffffffffffLL;
        round_conclude(s, R.gc, sorted, K, stopped ? R.sstop : -1, R.sstop_state, R.slast);
      }
      __syncthreads();
    }
  }
  if (!ok) {
    if (threadIdx.x == 0) set_done(s, -2);
    __syncthreads();
    return G.epoch;
  }
  if (s->gcp_pending) {
    // every candidate of the t-range passed: the next range
    if (threadIdx.x == 0) logic_resume(s, 0);
    __syncthreads();
  }
  return G.epoch;
}

// The breakpoint rounds of the Cauchy search.  A gang selects over its
// segments, then sorts and scans the round together (gang_round).
template <typename T, bool GO>
__device__ __forceinline__ void b_gcp(const Prob<T>& P, BatchLds& L, Gang<GO>& G) {
  DevScalars* s = &L.s;
  if (!s->gcp_pending || s->halt) return;
  for (int guard = 0; guard < 64 && s->gcp_pending; guard++) {
    if constexpr (!GO) {
      if (!b_select_listed(P, L)) {
        if constexpr (sizeof(T) == 4)
          b_select_cold(P, L, G);
        else
          b_select(P, L, G);
      }
      b_round(round_args(P, s), L);
    } else {
      b_select(P, L, G);
      if (!gang_sync(G, &L.gflag)) {
        if (threadIdx.x == 0) set_done(s, -2);
        __syncthreads();
        return;
      }
      G.epoch = gang_round(round_args(P, s), L, G);
    }
  }
}

// One step of the chain (the block form of the plain engine's kernels,
// engine.cpp: launch_step).
template <typename T, bool GO>
__device__ __forceinline__ void b_step(int step, double f, const Prob<T>& P, BatchLds& L,
                                       Gang<GO>& G) {
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
      b_pass_a<T, false, GO>(P, L, G);
      break;
    case K_PASSA_FUSED:
      b_pass_a<T, true, GO>(P, L, G);
      break;
    case K_UPD_REDUCE:
      if (s->halt) {
        if (t0) s->write_trial = 0;
        break;
      }
      if (t0) logic_update_gcp(s, 1, L.red, L.rv[1], L.ri[0]);
      __syncthreads();
      post_update_block(s);
      build_middle(s);
      break;
    case K_STOP:
      if (s->halt) {
        if (t0) s->write_trial = 0;
        break;
      }
      if (t0) logic_stop_tests(s, L.red, L.rv[0], L.rv[1], L.ri[0]);
      __syncthreads();
      post_update_block(s);
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
      post_update_block(s);
      build_middle(s);
      break;
    case K_GCP:
      b_gcp(P, L, G);
      break;
    case K_P4:
      b_p4(P, L, G);
      break;
    case K_FORMK:
      if (s->halt) break;
      logic_formk_subsm(s, L.red, s->p4_flagged != 0);
      break;
    case K_P5A:
      b_p5a(P, L, G);
      break;
    case K_ALPHA:
      if (s->halt || s->skip_subsm) break;
      if (t0) logic_alpha(s, L.rv[2], L.ri[1]);
      break;
    case K_P5B:
      b_p5b(P, L, G);
      break;
    case K_LSINIT:
      if (s->halt) break;
      if (t0) logic_ls_init(s, L.red[P5_DTD], L.red[P5_GD], L.rv[3]);
      break;
    case K_TRIAL:
      b_trial(P, L);
      break;
    case K_GD:
      b_gd(P, L, G);
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

template <typename T, bool GO>
__device__ __forceinline__ void b_program(int prog, double f, const Prob<T>& P, BatchLds& L,
                                          Gang<GO>& G) {
  if (prog < PROG_ITER_START || prog > PROG_AFTER_ACCEPT) return;
  for (int k = 0; k < 16; k++) {
    const int step = kProgSteps[prog][k];
    if (step < 0) break;
    b_step(step, f, P, L, G);
  }
}

// One problem (or one gang member's share of it): set-up (first launch
// only) and the chain loop of the host driver (capi.cpp: drive), with the
// objective evaluated in the workgroup.  Two variants: the one-workgroup
// form (GO = false) has no gang step in it.  Both are inlined into the
// kernel, where the arrays are known to be global memory.
template <typename T, bool GO>
__device__ __forceinline__ void solve_one(const BatchProb& bp, const BatchBufs<T>& B,
                                          DevScalars* dsg, BatchResult* res, int iter_budget,
                                          BatchLds& L, Gang<GO>& G) {
  Prob<T> P;
  P.n = bp.n;
  if constexpr (GO) {
    // member j's segment: chunks of 64 elements split as evenly as possible
    const long long chunks = (bp.n + 63) / 64;
    P.lo = ((long long)G.j * chunks / G.S) * 64;
    const long long hi = ((long long)(G.j + 1) * chunks / G.S) * 64;
    P.hi = hi < bp.n ? hi : bp.n;
  } else {
    P.lo = 0;
    P.hi = bp.n;
  }
  P.X = Pair<T>{B.x + bp.off, B.xb + bp.off};
  P.G = Pair<T>{B.g + bp.off, B.gb + bp.off};
  P.l = B.l + bp.off;
  P.u = B.u + bp.off;
  P.nbd = B.nbd + bp.off;
  P.iwhere = B.iwhere + bp.off;
  P.W = B.W + bp.woff;
  P.d = B.d + bp.off;
  P.dsub = B.dsub + bp.off;
  P.ca = B.ca + bp.off;
  P.cb = B.cb + bp.off;
  P.obj = bp.obj;

  const bool lead = !GO || G.j == 0;
  ds_copy((unsigned long long*)&L.s, (const unsigned long long*)dsg);
  DevScalars* s = &L.s;
  if (!s->b_started) {
    if (threadIdx.x == 0 && lead) res->clk_start = (long long)wall_clock64();
    b_errclb(P, L, G);
    if (s->status != ST_DONE) {
      b_active(P, L, G);
      b_objective(P, L, 0, G);
      b_projgr(P, L, G);
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
    b_program(prog, fv, P, L, G);
    const int st = s->status;
    if (st == ST_NEED_EVAL) {
      b_objective(P, L, s->trial_dst, G);
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
    for (long long i = P.lo + threadIdx.x; i < P.hi; i += blockDim.x) P.X[0][i] = P.X[1][i];
  }
  if (!lead) {
    __syncthreads();
    return;
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
  ds_copy((unsigned long long*)dsg, (const unsigned long long*)&L.s);
}

// The persistent kernel over one queue: GO = true takes the gang entries
// items[0 .. count[0]) through next[0], GO = false the one-workgroup
// entries items[count[0] .. count[0] + count[1]) through next[1].  The two
// are separate kernels so that each has the register allocation of its own
// form.
template <typename T, int NT, bool GO>
__global__ void __launch_bounds__(NT)
k_batch(const BatchProb* __restrict__ probs, const BatchItem* __restrict__ items,
        const int* __restrict__ count_p, unsigned* next, GangCtl* gang, double* gpart,
        BatchBufs<T> B, DevScalars* dsa, BatchResult* res, int iter_budget) {
  __shared__ BatchLds L;
  __shared__ int s_k;
  const int count = count_p[GO ? 0 : 1];
  const int first = GO ? 0 : count_p[0];
  for (;;) {
    if (threadIdx.x == 0) s_k = (int)atomicAdd(next + (GO ? 0 : 1), 1u);
    __syncthreads();
    const int k = s_k;
    __syncthreads();
    if (k >= count) return;
    const BatchItem it = items[first + k];
    const BatchProb bp = probs[it.prob];
    Gang<GO> G{GO ? bp.gang : 1, it.member, (long long)k - it.member, gang + it.prob, gpart, 0u};
    solve_one<T, GO>(bp, B, dsa + it.prob, res + it.prob, iter_budget, L, G);
  }
}

__global__ void k_batch_clear(unsigned* next, GangCtl* gang, int nprob) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) next[i] = 0u;
  if (i < nprob) {
    gang[i].arrive = 0u;
    gang[i].err = 0;
    gang[i].cand = 0ull;
  }
}

}  // namespace

template <typename T, int NT>
void launch_pair(hipStream_t s, int blocks_gang, int blocks_one, const BatchProb* probs,
                 const BatchItem* items, const int* count, unsigned* next, GangCtl* gang,
                 double* gpart, BatchBufs<T> B, DevScalars* dsa, BatchResult* res,
                 int iter_budget) {
  if (blocks_gang > 0)
    hipLaunchKernelGGL((k_batch<T, NT, true>), dim3(blocks_gang), dim3(NT), 0, s, probs, items,
                       count, next, gang, gpart, B, dsa, res, iter_budget);
  if (blocks_one > 0)
    hipLaunchKernelGGL((k_batch<T, NT, false>), dim3(blocks_one), dim3(NT), 0, s, probs, items,
                       count, next, gang, gpart, B, dsa, res, iter_budget);
}

template <typename T>
int launch_batch(hipStream_t s, int threads, int blocks_gang, int blocks_one,
                 const BatchProb* probs, const BatchItem* items, const int* count, unsigned* next,
                 GangCtl* gang, double* gpart, BatchBufs<T> B, DevScalars* dsa,
                 BatchResult* res, int iter_budget) {
  switch (threads) {
    case 256:
      launch_pair<T, 256>(s, blocks_gang, blocks_one, probs, items, count, next, gang, gpart, B,
                          dsa, res, iter_budget);
      return 0;
    case 512:
      launch_pair<T, 512>(s, blocks_gang, blocks_one, probs, items, count, next, gang, gpart, B,
                          dsa, res, iter_budget);
      return 0;
    default:
      return -1;
  }
}

void launch_batch_clear(hipStream_t s, unsigned* next, GangCtl* gang, int nprob) {
  const int n = nprob > 1 ? nprob : 1;
  hipLaunchKernelGGL(k_batch_clear, dim3((n + 255) / 256), dim3(256), 0, s, next, gang, nprob);
}

template <typename T>
int batch_blocks_per_cu(int threads) {
  int a = 0, b = 0;
  hipError_t e1, e2;
  if (threads == 512) {
    e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_batch<T, 512, true>, 512, 0);
    e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_batch<T, 512, false>, 512, 0);
  } else {
    e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_batch<T, 256, true>, 256, 0);
    e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_batch<T, 256, false>, 256, 0);
  }
  if (e1 != hipSuccess || e2 != hipSuccess) return 0;
  return a < b ? a : b;
}

template int launch_batch<float>(hipStream_t, int, int, int, const BatchProb*,
                                 const BatchItem*, const int*, unsigned*, GangCtl*, double*,
                                 BatchBufs<float>, DevScalars*, BatchResult*, int);
template int launch_batch<double>(hipStream_t, int, int, int, const BatchProb*,
                                  const BatchItem*, const int*, unsigned*, GangCtl*, double*,
                                  BatchBufs<double>, DevScalars*, BatchResult*, int);
template int batch_blocks_per_cu<float>(int);
template int batch_blocks_per_cu<double>(int);

}  // namespace lbfgsb
