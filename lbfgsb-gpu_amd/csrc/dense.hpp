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

// LINPACK dpofa, upper Cholesky in place (lbfgsbcpu.cpp:1876-1908).
__device__ inline bool d_dpofa(double* a, int lda, int n) {
  for (int j = 1; j <= n; j++) {
    double s = 0.0;
    for (int k = 1; k <= j - 1; k++) {
      double v = ddot(&A1(a, lda, 1, k), lda, &A1(a, lda, 1, j), lda, k - 1);
      double t = A1(a, lda, k, j) - v;
      t = t / A1(a, lda, k, k);
      A1(a, lda, k, j) = t;
      s = s + t * t;
    }
    s = A1(a, lda, j, j) - s;
    if (s <= 0.0) return false;
    A1(a, lda, j, j) = sqrt(s);
  }
  return true;
}

// LINPACK dtrsl for upper-triangular t: job 1 solves T x = b, job 11 solves
// T' x = b (lbfgsbcpu.cpp:1910-1986).  Returns 0 or the zero-pivot index.
__device__ inline int d_dtrsl(const double* t, int ldt, int n, double* b, int job) {
  for (int j = 1; j <= n; j++)
    if (A1(t, ldt, j, j) == 0.0) return j;
  if (job == 1) {
    b[n - 1] = b[n - 1] / A1(t, ldt, n, n);
    for (int jj = 2; jj <= n; jj++) {
      int j = n - jj + 1;
      double temp = -b[j];
      for (int i = 1; i <= j; i++) b[i - 1] += temp * A1(t, ldt, i, j + 1);
      b[j - 1] = b[j - 1] / A1(t, ldt, j, j);
    }
  } else {  // job 11
    b[0] = b[0] / A1(t, ldt, 1, 1);
    for (int j = 2; j <= n; j++) {
      double v = ddot(&A1(t, ldt, 1, j), ldt, b, 1, j - 1);
      b[j - 1] = b[j - 1] - v;
      b[j - 1] = b[j - 1] / A1(t, ldt, j, j);
    }
  }
  return 0;
}

// p = M v with the 2col x 2col middle matrix through sy and the Cholesky
// factor of T in wt (lbfgsbcpu.cpp:517-565).  sy, wt have ld 8.
__device__ inline int d_bmv(const double* sy, const double* wt, int col, const double* v,
                            double* p) {
  if (col == 0) return 0;
  p[col] = v[col];
  for (int i = 2; i <= col; i++) {
    int i2 = col + i;
    double s = 0.0;
    for (int k = 1; k <= i - 1; k++) s = s + A1(sy, 8, i, k) * v[k - 1] / A1(sy, 8, k, k);
    p[i2 - 1] = v[i2 - 1] + s;
  }
  // the two triangular solves work in place on p[col .. 2col-1] (no
  // per-thread temporaries: those would live in scratch memory)
  int info = d_dtrsl(wt, 8, col, p + col, 11);
  if (info != 0) return info;
  for (int i = 1; i <= col; i++) p[i - 1] = v[i - 1] / sqrt(A1(sy, 8, i, i));
  info = d_dtrsl(wt, 8, col, p + col, 1);
  if (info != 0) return info;
  for (int i = 1; i <= col; i++) p[i - 1] = -p[i - 1] / sqrt(A1(sy, 8, i, i));
  for (int i = 1; i <= col; i++) {
    double s = 0.0;
    for (int k = i + 1; k <= col; k++) s = s + A1(sy, 8, k, i) * p[col + k - 1] / A1(sy, 8, i, i);
    p[i - 1] = p[i - 1] + s;
  }
  return 0;
}

// T = theta S'S + L D^{-1} L' and its Cholesky factor (lbfgsbcpu.cpp:1105-1133).
__device__ inline bool d_formt(double* wt, const double* sy, const double* ss, int col,
                               double theta) {
  for (int j = 1; j <= col; j++) A1(wt, 8, 1, j) = theta * A1(ss, 8, 1, j);
  for (int i = 2; i <= col; i++) {
    for (int j = i; j <= col; j++) {
      int k1 = (i < j ? i : j) - 1;
      double ddum = 0.0;
      for (int k = 1; k <= k1; k++)
        ddum = ddum + A1(sy, 8, i, k) * A1(sy, 8, j, k) / A1(sy, 8, k, k);
      A1(wt, 8, i, j) = ddum + theta * A1(ss, 8, i, j);
    }
  }
  return d_dpofa(wt, 8, col);
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
