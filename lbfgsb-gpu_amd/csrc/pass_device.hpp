// Per-element helpers of the bulk passes, shared by the grid-wide kernels
// (bulk.hip) and the per-problem batched kernel.
#pragma once

#include "common.hpp"

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

// Projected gradient component (lbfgsbcpu.cpp:1389-1414).
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

}  // namespace lbfgsb
