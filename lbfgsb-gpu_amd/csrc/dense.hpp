// FP64 small dense algebra and the Moré-Thuente line-search scalars, as
// device functions run by one thread of a single-workgroup kernel.  The
// algorithms are those of the reference CPU path; each function cites the
// lines it follows.  Matrices are row-major with an explicit leading
// dimension; a(i,j) below is 1-based as in the reference.
#pragma once

#include <hip/hip_runtime.h>

namespace lbfgsb {

#define A1(a, ld, i, j) (a)[((i)-1) * (ld) + ((j)-1)]

// ALGLIB blocked dot product order (culbfgsb/apvt.h:86-131).
__device__ inline double ddot(const double* a, int sa, const double* b, int sb, int n) {
  double r = 0.0;
  int nb = n / 4;
  for (int i = 0; i < nb; i++) {
    r += a[0] * b[0] + a[sa] * b[sb] + a[2 * sa] * b[2 * sb] + a[3 * sa] * b[3 * sb];
    a += 4 * sa;
    b += 4 * sb;
  }
  for (int i = 0; i < n % 4; i++) {
    r += a[0] * b[0];
    a += sa;
    b += sb;
  }
  return r;
}

// ---------------------------------------------------------------------------
// LINPACK dpofa / dtrsl and bmv for one thread, with the vector (or matrix)
// being solved in registers: every loop is unrolled to a fixed maximum size
// and masked by the actual size, so each entry sees the reference's
// operations in the reference's order without a dependent LDS round trip
// per operation.

// ALGLIB blocked dot of column j of t (rows 1 .. len) with b[0 .. len).
template <int NMAX>
__device__ __forceinline__ double ddot_col_r(const double* __restrict__ t, int ldt, int j,
                                             const double (&b)[NMAX], int len) {
  double r = 0.0;
#pragma unroll
  for (int blk = 0; blk < NMAX / 4; blk++) {
    if (4 * blk + 4 <= len) {
      const int i = 4 * blk + 1;
      r += A1(t, ldt, i, j) * b[i - 1] + A1(t, ldt, i + 1, j) * b[i] +
           A1(t, ldt, i + 2, j) * b[i + 1] + A1(t, ldt, i + 3, j) * b[i + 2];
    }
  }
#pragma unroll
  for (int i = 1; i <= NMAX; i++) {
    if (i > (len / 4) * 4 && i <= len) r += A1(t, ldt, i, j) * b[i - 1];
  }
  return r;
}

// dtrsl (lbfgsbcpu.cpp:1910-1986) on a register vector b[0 .. n), n <= NMAX.
template <int NMAX>
__device__ __forceinline__ int d_dtrsl_r(const double* __restrict__ t, int ldt, int n,
                                         double (&b)[NMAX], int job) {
  for (int j = 1; j <= n; j++)
    if (A1(t, ldt, j, j) == 0.0) return j;
  if (job == 1) {
#pragma unroll
    for (int k = 0; k < NMAX; k++)
      if (k == n - 1) b[k] = b[k] / A1(t, ldt, n, n);
#pragma unroll
    for (int j = NMAX - 1; j >= 1; j--) {
      if (j <= n - 1) {
        const double temp = -b[j];
#pragma unroll
        for (int i = 1; i <= j; i++) b[i - 1] += temp * A1(t, ldt, i, j + 1);
        b[j - 1] = b[j - 1] / A1(t, ldt, j, j);
      }
    }
  } else {  // job 11
    b[0] = b[0] / A1(t, ldt, 1, 1);
#pragma unroll
    for (int j = 2; j <= NMAX; j++) {
      if (j <= n) {
        const double v = ddot_col_r<NMAX>(t, ldt, j, b, j - 1);
        b[j - 1] = b[j - 1] - v;
        b[j - 1] = b[j - 1] / A1(t, ldt, j, j);
      }
    }
  }
  return 0;
}

// dpofa (lbfgsbcpu.cpp:1876-1908) on the upper triangle of a register
// matrix a[r][c] (0-based), n <= NMAX.
template <int NMAX>
__device__ __forceinline__ bool d_dpofa_r(double (&a)[NMAX][NMAX], int n) {
#pragma unroll
  for (int j = 1; j <= NMAX; j++) {
    if (j > n) continue;
    double s = 0.0;
#pragma unroll
    for (int k = 1; k < NMAX; k++) {
      if (k > j - 1) continue;
      // ddot(a(1:k-1, k), a(1:k-1, j)) in the blocked order
      double v = 0.0;
#pragma unroll
      for (int blk = 0; blk < NMAX / 4; blk++) {
        if (4 * blk + 4 <= k - 1) {
          const int i = 4 * blk;
          v += a[i][k - 1] * a[i][j - 1] + a[i + 1][k - 1] * a[i + 1][j - 1] +
               a[i + 2][k - 1] * a[i + 2][j - 1] + a[i + 3][k - 1] * a[i + 3][j - 1];
        }
      }
#pragma unroll
      for (int i = 0; i < NMAX; i++)
        if (i >= ((k - 1) / 4) * 4 && i < k - 1) v += a[i][k - 1] * a[i][j - 1];
      double t = a[k - 1][j - 1] - v;
      t = t / a[k - 1][k - 1];
      a[k - 1][j - 1] = t;
      s = s + t * t;
    }
    s = a[j - 1][j - 1] - s;
    if (s <= 0.0) return false;
    a[j - 1][j - 1] = sqrt(s);
  }
  return true;
}

// p = M v with the 2col x 2col middle matrix through sy and the Cholesky
// factor of T in wt (lbfgsbcpu.cpp:517-565), compact form: sy, wt (ld 8)
// read where used, branches on col.  The two halves of v and p (Y part,
// S part) are kept in registers.
__device__ inline int d_bmv_compact(const double* __restrict__ sy, const double* __restrict__ wt,
                            int col, const double* v, double* p) {
  if (col == 0) return 0;
  double vy[8], vs[8], py[8], ps[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    vy[q] = q < col ? v[q] : 0.0;
    vs[q] = q < col ? v[col + q] : 0.0;
    py[q] = 0.0;
    ps[q] = 0.0;
  }
  ps[0] = vs[0];
#pragma unroll
  for (int i = 2; i <= 8; i++) {
    if (i > col) continue;
    double s = 0.0;
#pragma unroll
    for (int k = 1; k < 8; k++)
      if (k <= i - 1) s = s + A1(sy, 8, i, k) * vy[k - 1] / A1(sy, 8, k, k);
    ps[i - 1] = vs[i - 1] + s;
  }
  int info = d_dtrsl_r<8>(wt, 8, col, ps, 11);
  if (info != 0) return info;
#pragma unroll
  for (int i = 1; i <= 8; i++)
    if (i <= col) py[i - 1] = vy[i - 1] / sqrt(A1(sy, 8, i, i));
  info = d_dtrsl_r<8>(wt, 8, col, ps, 1);
  if (info != 0) return info;
#pragma unroll
  for (int i = 1; i <= 8; i++)
    if (i <= col) py[i - 1] = -py[i - 1] / sqrt(A1(sy, 8, i, i));
#pragma unroll
  for (int i = 1; i <= 8; i++) {
    if (i > col) continue;
    double s = 0.0;
#pragma unroll
    for (int k = 2; k <= 8; k++)
      if (k >= i + 1 && k <= col) s = s + A1(sy, 8, k, i) * ps[k - 1] / A1(sy, 8, i, i);
    py[i - 1] = py[i - 1] + s;
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (q < col) {
      p[q] = py[q];
      p[col + q] = ps[q];
    }
  }
  return 0;
}

// The same product, hoisted form: every entry
// of sy's lower and wt's upper triangle is loaded before the first
// operation, and the loops run over the fixed size 8 with the entries beyond
// col masked out by selects, not branches: the loads go out together and the
// independent divisions overlap.  Each used entry sees the reference's
// operations in the reference's order; entries beyond col are never stored.
__device__ inline int d_bmv_hoisted(const double* __restrict__ sy, const double* __restrict__ wt,
                                    int col, const double* v, double* p) {
  if (col == 0) return 0;
  double S[8][8], Tu[8][8];  // S[i][k], k <= i; Tu[k][j], k <= j (0-based)
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (k <= i) S[i][k] = sy[i * 8 + k];
      if (i <= k) Tu[i][k] = wt[i * 8 + k];
    }
  double vy[8], vs[8], py[8], ps[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const double a = v[q], b = v[col + q < 16 ? col + q : 15];
    vy[q] = q < col ? a : 0.0;
    vs[q] = q < col ? b : 0.0;
  }
  // dtrsl's zero-pivot test comes before any arithmetic: the first zero
  // diagonal entry of T within col
  int info = 0;
#pragma unroll
  for (int j = 7; j >= 0; j--)
    if (j < col && Tu[j][j] == 0.0) info = j + 1;
  if (info != 0) return info;
  // ps_i = vs_i + sum_{k < i} sy(i,k) vy_k / sy(k,k)
  ps[0] = vs[0];
#pragma unroll
  for (int i = 1; i < 8; i++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < i; k++) s = s + S[i][k] * vy[k] / S[k][k];
    ps[i] = vs[i] + s;
  }
  // solve T' x = ps (dtrsl job 11): forward, ALGLIB blocked dots
  ps[0] = ps[0] / Tu[0][0];
#pragma unroll
  for (int j = 1; j < 8; j++) {
    if (j < col) {
      double r = 0.0;
#pragma unroll
      for (int blk = 0; blk < 2; blk++) {
        if (4 * blk + 4 <= j) {
          const int i = 4 * blk;
          r += Tu[i][j] * ps[i] + Tu[i + 1][j] * ps[i + 1] + Tu[i + 2][j] * ps[i + 2] +
               Tu[i + 3][j] * ps[i + 3];
        }
      }
#pragma unroll
      for (int i = 0; i < 8; i++)
        if (i >= (j / 4) * 4 && i < j) r += Tu[i][j] * ps[i];
      ps[j] = ps[j] - r;
      ps[j] = ps[j] / Tu[j][j];
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) py[i] = vy[i] / sqrt(S[i][i]);
  // solve T x = ps (dtrsl job 01): backward
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (k == col - 1) ps[k] = ps[k] / Tu[k][k];
#pragma unroll
  for (int j = 7; j >= 1; j--) {
    if (j <= col - 1) {
      const double temp = -ps[j];
#pragma unroll
      for (int i = 0; i < j; i++) ps[i] += temp * Tu[i][j];
      ps[j - 1] = ps[j - 1] / Tu[j - 1][j - 1];
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) py[i] = -py[i] / sqrt(S[i][i]);
  // py_i += sum_{i < k < col} sy(k,i) ps_k / sy(i,i)
#pragma unroll
  for (int i = 0; i < 8; i++) {
    double s = 0.0;
#pragma unroll
    for (int k = i + 1; k < 8; k++) {
      const double term = S[k][i] * ps[k] / S[i][i];
      s = k < col ? s + term : s;
    }
    py[i] = py[i] + s;
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (q < col) {
      p[q] = py[q];
      p[col + q] = ps[q];
    }
  }
  return 0;
}

// The single-workgroup control kernels take the hoisted form (its loads
// overlap; it needs ~200 registers).  The batched kernel, whose register
// allocation its bulk passes share, is built with LBFGSB_COMPACT_BMV set and
// takes the compact one.  Both give the same results bit for bit.
#ifndef LBFGSB_COMPACT_BMV
#define LBFGSB_COMPACT_BMV 0
#endif
__device__ inline int d_bmv(const double* __restrict__ sy, const double* __restrict__ wt,
                            int col, const double* v, double* p) {
  if constexpr (LBFGSB_COMPACT_BMV)
    return d_bmv_compact(sy, wt, col, v, p);
  else
    return d_bmv_hoisted(sy, wt, col, v, p);
}

__device__ inline double dmax_(double a, double b) { return a > b ? a : b; }
__device__ inline double dmin_(double a, double b) { return a > b ? b : a; }

// MINPACK-2 dcstep (lbfgsbcpu.cpp:1737-1874).
__device__ inline void d_dcstep(double& stx, double& fx, double& dx, double& sty, double& fy,
                                double& dy, double& stp, double fp, double dp, int& brackt,
                                double stpmin, double stpmax) {
  double gamma, p, q, r, s, stpc, stpf, stpq, theta;
  double sgnd = dp * (dx / fabs(dx));
  if (fp > fx) {
    theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
    s = dmax_(fabs(theta), dmax_(fabs(dx), fabs(dp)));
    gamma = s * sqrt((theta / s) * (theta / s) - dx / s * (dp / s));
    if (stp < stx) gamma = -gamma;
    p = gamma - dx + theta;
    q = gamma - dx + gamma + dp;
    r = p / q;
    stpc = stx + r * (stp - stx);
    stpq = stx + dx / ((fx - fp) / (stp - stx) + dx) / 2 * (stp - stx);
    if (fabs(stpc - stx) < fabs(stpq - stx))
      stpf = stpc;
    else
      stpf = stpc + (stpq - stpc) / 2;
    brackt = 1;
  } else if (sgnd < 0) {
    theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
    s = dmax_(fabs(theta), dmax_(fabs(dx), fabs(dp)));
    gamma = s * sqrt((theta / s) * (theta / s) - dx / s * (dp / s));
    if (stp > stx) gamma = -gamma;
    p = gamma - dp + theta;
    q = gamma - dp + gamma + dx;
    r = p / q;
    stpc = stp + r * (stx - stp);
    stpq = stp + dp / (dp - dx) * (stx - stp);
    if (fabs(stpc - stp) > fabs(stpq - stp))
      stpf = stpc;
    else
      stpf = stpq;
    brackt = 1;
  } else if (fabs(dp) < fabs(dx)) {
    theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
    s = dmax_(fabs(theta), dmax_(fabs(dx), fabs(dp)));
    gamma = s * sqrt(dmax_(0.0, (theta / s) * (theta / s) - dx / s * (dp / s)));
    if (stp > stx) gamma = -gamma;
    p = gamma - dp + theta;
    q = gamma + (dx - dp) + gamma;
    r = p / q;
    if (r < 0 && gamma != 0)
      stpc = stp + r * (stx - stp);
    else if (stp > stx)
      stpc = stpmax;
    else
      stpc = stpmin;
    stpq = stp + dp / (dp - dx) * (stx - stp);
    if (brackt) {
      if (fabs(stpc - stp) < fabs(stpq - stp))
        stpf = stpc;
      else
        stpf = stpq;
      if (stp > stx)
        stpf = dmin_(stp + 0.66 * (sty - stp), stpf);
      else
        stpf = dmax_(stp + 0.66 * (sty - stp), stpf);
    } else {
      if (fabs(stpc - stp) > fabs(stpq - stp))
        stpf = stpc;
      else
        stpf = stpq;
      stpf = dmin_(stpmax, stpf);
      stpf = dmax_(stpmin, stpf);
    }
  } else {
    if (brackt) {
      theta = 3 * (fp - fy) / (sty - stp) + dy + dp;
      s = dmax_(fabs(theta), dmax_(fabs(dy), fabs(dp)));
      gamma = s * sqrt((theta / s) * (theta / s) - dy / s * (dp / s));
      if (stp > sty) gamma = -gamma;
      p = gamma - dp + theta;
      q = gamma - dp + gamma + dy;
      r = p / q;
      stpc = stp + r * (sty - stp);
      stpf = stpc;
    } else if (stp > stx) {
      stpf = stpmax;
    } else {
      stpf = stpmin;
    }
  }
  if (fp > fx) {
    sty = stp;
    fy = fp;
    dy = dp;
  } else {
    if (sgnd < 0) {
      sty = stx;
      fy = fx;
      dy = dx;
    }
    stx = stp;
    fx = fp;
    dx = dp;
  }
  stp = stpf;
}

// MINPACK-2 dcsrch (lbfgsbcpu.cpp:1541-1735).  task: 0 START, 1 FG,
// 2 ERROR, 3 WARNING, 4 CONVERGENCE.
__device__ inline void d_dcsrch(double f, double g, double& stp, double ftol, double gtol,
                                double xtol, double stpmin, double stpmax, int& task,
                                int* isave, double* dsave) {
  int brackt, stage;
  double finit, ftest, fm, fx, fxm, fy, fym, ginit, gtest, gm, gx, gxm, gy, gym;
  double stx, sty, stmin, stmax, width, width1;
  const double xtrapl = 1.1, xtrapu = 4.0;
  if (task == 0) {
    if (stp < stpmin) task = 2;
    if (stp > stpmax) task = 2;
    if (g >= 0) task = 2;
    if (ftol < 0) task = 2;
    if (gtol < 0) task = 2;
    if (xtol < 0) task = 2;
    if (stpmin < 0) task = 2;
    if (stpmax < stpmin) task = 2;
    if (task == 2) return;
    brackt = 0;
    stage = 1;
    finit = f;
    ginit = g;
    gtest = ftol * ginit;
    width = stpmax - stpmin;
    width1 = width / 0.5;
    stx = 0;
    fx = finit;
    gx = ginit;
    sty = 0;
    fy = finit;
    gy = ginit;
    stmin = 0;
    stmax = stp + xtrapu * stp;
    task = 1;
  } else {
    brackt = isave[0] == 1;
    stage = isave[1];
    ginit = dsave[0];
    gtest = dsave[1];
    gx = dsave[2];
    gy = dsave[3];
    finit = dsave[4];
    fx = dsave[5];
    fy = dsave[6];
    stx = dsave[7];
    sty = dsave[8];
    stmin = dsave[9];
    stmax = dsave[10];
    width = dsave[11];
    width1 = dsave[12];
    ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0) stage = 2;
    if (brackt && (stp <= stmin || stp >= stmax)) task = 3;
    if (brackt && stmax - stmin <= xtol * stmax) task = 3;
    if (stp == stpmax && f <= ftest && g <= gtest) task = 3;
    if (stp == stpmin && (f > ftest || g >= gtest)) task = 3;
    if (f <= ftest && fabs(g) <= gtol * (-ginit)) task = 4;
    if (!(task == 3 || task == 4)) {
      if (stage == 1 && f <= fx && f > ftest) {
        fm = f - stp * gtest;
        fxm = fx - stx * gtest;
        fym = fy - sty * gtest;
        gm = g - gtest;
        gxm = gx - gtest;
        gym = gy - gtest;
        d_dcstep(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
        fx = fxm + stx * gtest;
        fy = fym + sty * gtest;
        gx = gxm + gtest;
        gy = gym + gtest;
      } else {
        d_dcstep(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
      }
      if (brackt) {
        if (fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
        width1 = width;
        width = fabs(sty - stx);
      }
      if (brackt) {
        stmin = dmin_(stx, sty);
        stmax = dmax_(stx, sty);
      } else {
        stmin = stp + xtrapl * (stp - stx);
        stmax = stp + xtrapu * (stp - stx);
      }
      stp = dmax_(stp, stpmin);
      stp = dmin_(stp, stpmax);
      if ((brackt && (stp <= stmin || stp >= stmax)) || (brackt && stmax - stmin <= xtol * stmax))
        stp = stx;
      task = 1;
    }
  }
  isave[0] = brackt ? 1 : 0;
  isave[1] = stage;
  dsave[0] = ginit;
  dsave[1] = gtest;
  dsave[2] = gx;
  dsave[3] = gy;
  dsave[4] = finit;
  dsave[5] = fx;
  dsave[6] = fy;
  dsave[7] = stx;
  dsave[8] = sty;
  dsave[9] = stmin;
  dsave[10] = stmax;
  dsave[11] = width;
  dsave[12] = width1;
}

#undef A1

}  // namespace lbfgsb
