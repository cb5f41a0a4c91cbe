// Per-element terms of the built-in objectives (device side), shared by the
// grid-wide objective kernel and the per-problem batched kernel.
//   GRID_QUAD  f = 0.5 sum_edges (v_a - v_b)^2 - h^2 sum b v  (zero ring)
//   GRID_COS   f = 0.5 sum_edges (dv)^2 + h^2 sum (c cos v - b v)
//   ROSEN      f = sum_i 100 (x_{2i} - x_{2i-1}^2)^2 + (1 - x_{2i-1})^2
//   PENTA      f = 0.5 x'Ax - b'x, A symmetric pentadiagonal
//   DSSCFG     MINPACK-2 steady-state combustion on nx x ny points, two
//              linear triangles per point, zero outside the grid
//              (examples/dsscfg/dsscfg_cpu.cpp:17-150), lambda in o.h
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "objectives.hpp"

namespace lbfgsb {

// Number of terms k of an objective over n variables (ROSEN: pairs).
__device__ inline long long obj_terms(const ObjDesc& o, long long n) {
  return o.kind == OBJK_ROSEN ? n / 2 : n;
}

// The gradient entries term k owns (ROSEN: the pair 2k, 2k + 1; the other
// kinds: entry k in g0).
template <typename T>
struct ObjOut {
  T g0, g1;
};

// Column and row of grid point k (row-major, nx per row): a 32-bit division
// where the grid allows one (the 64-bit one is a long instruction sequence).
__device__ __forceinline__ void grid_pos(long long k, long long n, long long nx, long long& ix,
                                         long long& iy) {
  if (n <= 0x7fffffffLL) {
    const unsigned q = (unsigned)k / (unsigned)nx;
    iy = q;
    ix = k - (long long)q * nx;
  } else {
    iy = k / nx;
    ix = k - iy * nx;
  }
}

// Term k of an objective of kind KIND: its share of f, with its gradient
// entries in out (not stored, so that a caller can evaluate several terms
// before the first store).
template <int KIND, typename T>
__device__ inline double obj_eval_kind(const ObjDesc& o, long long n, const T* __restrict__ x,
                                       long long k, ObjOut<T>& out) {
  if constexpr (KIND == OBJK_GRID_QUAD || KIND == OBJK_GRID_COS) {
    const T* b = (const T*)o.b;
    const double h2 = o.h * o.h;
    const long long nx = o.nx, ny = o.ny;
    long long ix, iy;
    grid_pos(k, n, nx, ix, iy);
    const double v = (double)x[k];
    const double vl = ix > 0 ? (double)x[k - 1] : 0.0;
    const double vr = ix < nx - 1 ? (double)x[k + 1] : 0.0;
    const double vd = iy > 0 ? (double)x[k - nx] : 0.0;
    const double vu = iy < ny - 1 ? (double)x[k + nx] : 0.0;
    const double bk = (double)b[k];
    double gk = (4.0 * v - vl - vr - vd - vu) - h2 * bk;
    const double er = vr - v, eu = vu - v;
    double e = er * er + eu * eu;
    if (ix == 0) e += v * v;
    if (iy == 0) e += v * v;
    double fk = 0.5 * e - h2 * bk * v;
    if constexpr (KIND == OBJK_GRID_COS) {
      const double ck = (double)((const T*)o.c)[k];
      fk += h2 * ck * cos(v);
      gk -= h2 * ck * sin(v);
    }
    out.g0 = (T)gk;
    return fk;
  } else if constexpr (KIND == OBJK_DSSCFG) {
    // Point k owns its lower triangle (v, right, top) and its upper one
    // (v, left, bottom).  Its gradient gathers the five contributions the
    // reference example scatters (its own, then from the bottom, left,
    // right and top neighbours' triangles), in that order.
    const long long nx = o.nx, ny = o.ny;
    long long ix, iy;
    grid_pos(k, n, nx, ix, iy);
    const double lam = o.h;
    const double hx = 1.0 / (double)(nx + 1), hy = 1.0 / (double)(ny + 1);
    const double area = 0.5 * hx * hy;
    const double v = (double)x[k];
    const double vr = ix != nx - 1 ? (double)x[k + 1] : 0.0;
    const double vt = iy != ny - 1 ? (double)x[k + nx] : 0.0;
    const double vl = ix != 0 ? (double)x[k - 1] : 0.0;
    const double vb = iy != 0 ? (double)x[k - nx] : 0.0;
    const double dxl = (vr - v) / hx, dyl = (vt - v) / hy;
    const double dxu = (v - vl) / hx, dyu = (v - vb) / hy;
    const double ev = exp(v), ler = lam * ev / 3.0;
    double own = 0.0 + (-dxl / hx - dyl / hy - ler);
    own = own + (dxu / hx + dyu / hy - ler);
    double gk = own;
    gk = gk + (iy != ny - 1 ? -dyl / hy - ler : 0.0);
    gk = gk + (ix != nx - 1 ? -dxl / hx - ler : 0.0);
    gk = gk + (ix != 0 ? dxu / hx - ler : 0.0);
    gk = gk + (iy != 0 ? dyu / hy - ler : 0.0);
    out.g0 = (T)(gk * area);
    const double fq = dxl * dxl + dyl * dyl + (dxu * dxu + dyu * dyu);
    const double fe = lam * (ev + exp(vr) + exp(vt)) / 3.0 + lam * (exp(vb) + exp(vl) + ev) / 3.0;
    return area * (0.5 * fq - fe);
  } else if constexpr (KIND == OBJK_ROSEN) {
    // the pair (x_{2k}, x_{2k+1}) as one 2-vector load: a wave reads one
    // contiguous run instead of two strided ones
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    const V2 xv = reinterpret_cast<const V2*>(x)[k];
    const double a = (double)xv.x, bb = (double)xv.y;
    const double t1 = bb - a * a, t2 = 1.0 - a;
    out.g0 = (T)(-400.0 * t1 * a - 2.0 * t2);
    out.g1 = (T)(200.0 * t1);
    return 100.0 * t1 * t1 + t2 * t2;
  } else {
    const T* a0 = (const T*)o.a0;
    const T* a1 = (const T*)o.a1;
    const T* a2 = (const T*)o.a2;
    const T* b = (const T*)o.b;
    const double xi = (double)x[k];
    double ax = (double)a0[k] * xi;
    if (k + 1 < n) ax += (double)a1[k] * (double)x[k + 1];
    if (k + 2 < n) ax += (double)a2[k] * (double)x[k + 2];
    if (k >= 1) ax += (double)a1[k - 1] * (double)x[k - 1];
    if (k >= 2) ax += (double)a2[k - 2] * (double)x[k - 2];
    out.g0 = (T)(ax - (double)b[k]);
    return 0.5 * ax * xi - (double)b[k] * xi;
  }
}

// Stores the gradient entries of term k (kind KIND).
template <int KIND, typename T>
__device__ inline void obj_store_kind(T* __restrict__ g, long long k, const ObjOut<T>& out) {
  if constexpr (KIND == OBJK_ROSEN) {
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    V2 gv;
    gv.x = out.g0;
    gv.y = out.g1;
    reinterpret_cast<V2*>(g)[k] = gv;
  } else {
    g[k] = out.g0;
  }
}

}  // namespace lbfgsb
