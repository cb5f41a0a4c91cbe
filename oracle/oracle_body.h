/*
 * Body of the CPU oracle, instantiated twice by lbfgsb_oracle.c with
 *   R       the working real type (double or float)
 *   S(x)    suffixing macro (x##_f64 / x##_f32)
 *   R_IS_F32  1 for float
 * TEST INFRASTRUCTURE ONLY -- see lbfgsb_oracle.h.
 *
 * Indexing: the reference is 1-based throughout (ALGLIB arrays); the macros
 * below keep 1-based subscripts over 0-based storage so every loop reads
 * like the statement it restates.
 */

#define V1(a, i) ((a)[(i)-1])
#define M1(a, ld, i, j) ((a)[((i)-1) * (ld) + ((j)-1)])
#define WS(s, i, p) ((s)->ws[((i)-1) * (s)->m + ((p)-1)])
#define WY(s, i, p) ((s)->wy[((i)-1) * (s)->m + ((p)-1)])

#if R_IS_F32
#define RFABS(x) fabsf(x)
#define RSQRT(x) sqrtf(x)
#else
#define RFABS(x) fabs(x)
#define RSQRT(x) sqrt(x)
#endif

/* ap::maxreal / ap::minreal: `m1 > m2 ? m1 : m2` (culbfgsb/ap.h:508-516).
 * NaN propagation differs from fmax, so keep the comparison form. */
static inline R S(rmax)(R a, R b) { return a > b ? a : b; }
static inline R S(rmin)(R a, R b) { return a > b ? b : a; }

/* Machine epsilon used by the reference for both precisions
 * (culbfgsb/ap.cpp:38). */
#define ORC_MACHEPS 5e-16

/* Dot product in ALGLIB's blocked order (culbfgsb/apvt.h:86-131): sum of
 * 4-products added to the accumulator per block, then the remainder. */
R S(orc_dot)(const R* a, int sa, const R* b, int sb, int n) {
  R r = 0;
  int i, nb = n / 4;
  for (i = 0; i < nb; i++) {
    r += a[0] * b[0] + a[sa] * b[sb] + a[2 * sa] * b[2 * sb] +
         a[3 * sa] * b[3 * sb];
    a += 4 * sa;
    b += 4 * sb;
  }
  for (i = 0; i < n % 4; i++) {
    r += a[0] * b[0];
    a += sa;
    b += sb;
  }
  return r;
}

/* ---------------------------------------------------------------------- */
/* LINPACK dpofa: upper Cholesky A = R'R in place on the leading n x n block
 * of a row-major array with leading dimension lda.  Returns 1 on success,
 * 0 if a pivot is not positive (lbfgsbcpu.cpp:1876-1908). */
int S(orc_dpofa)(R* a, int lda, int n) {
  int j, k;
  for (j = 1; j <= n; j++) {
    R s = 0.0;
    for (k = 1; k <= j - 1; k++) {
      /* a(1..k-1, k) . a(1..k-1, j): column dots, stride lda */
      R v = S(orc_dot)(&M1(a, lda, 1, k), lda, &M1(a, lda, 1, j), lda, k - 1);
      R t = M1(a, lda, k, j) - v;
      t = t / M1(a, lda, k, k);
      M1(a, lda, k, j) = t;
      s = s + t * t;
    }
    s = M1(a, lda, j, j) - s;
    if (s <= 0.0) return 0;
    M1(a, lda, j, j) = RSQRT(s);
  }
  return 1;
}

/* LINPACK dtrsl for an upper-triangular t (lbfgsbcpu.cpp:1910-1986).
 * job 01: solve T x = b; job 11: solve T' x = b.  Returns 0, or the 1-based
 * index of a zero diagonal. */
int S(orc_dtrsl)(const R* t, int ldt, int n, R* b, int job) {
  int j, jj, cse;
  for (j = 1; j <= n; j++)
    if (M1(t, ldt, j, j) == 0.0) return j;
  cse = 1;
  if (job % 10 != 0) cse = 2;
  if (job % 100 / 10 != 0) cse = cse + 2;
  if (cse == 1) { /* T lower, T x = b */
    V1(b, 1) = V1(b, 1) / M1(t, ldt, 1, 1);
    for (j = 2; j <= n; j++) {
      R temp = -V1(b, j - 1);
      int i;
      for (i = j; i <= n; i++) V1(b, i) += temp * M1(t, ldt, i, j - 1);
      V1(b, j) = V1(b, j) / M1(t, ldt, j, j);
    }
    return 0;
  }
  if (cse == 2) { /* T upper, T x = b */
    V1(b, n) = V1(b, n) / M1(t, ldt, n, n);
    for (jj = 2; jj <= n; jj++) {
      int i;
      R temp;
      j = n - jj + 1;
      temp = -V1(b, j + 1);
      for (i = 1; i <= j; i++) V1(b, i) += temp * M1(t, ldt, i, j + 1);
      V1(b, j) = V1(b, j) / M1(t, ldt, j, j);
    }
    return 0;
  }
  if (cse == 3) { /* T lower, T' x = b */
    V1(b, n) = V1(b, n) / M1(t, ldt, n, n);
    for (jj = 2; jj <= n; jj++) {
      R v;
      j = n - jj + 1;
      v = S(orc_dot)(&M1(t, ldt, j + 1, j), ldt, &V1(b, j + 1), 1, jj - 1);
      V1(b, j) = V1(b, j) - v;
      V1(b, j) = V1(b, j) / M1(t, ldt, j, j);
    }
    return 0;
  }
  /* cse == 4: T upper, T' x = b */
  V1(b, 1) = V1(b, 1) / M1(t, ldt, 1, 1);
  for (j = 2; j <= n; j++) {
    R v = S(orc_dot)(&M1(t, ldt, 1, j), ldt, &V1(b, 1), 1, j - 1);
    V1(b, j) = V1(b, j) - v;
    V1(b, j) = V1(b, j) / M1(t, ldt, j, j);
  }
  return 0;
}

/* Product p = M v with the 2col x 2col middle matrix
 * M = [[-D, L'],[L, theta S'S]]^{-1}, through sy (m x m) and the Cholesky
 * factor of T = theta S'S + L D^{-1} L' held in wt (lbfgsbcpu.cpp:517-565).
 * Returns the dtrsl info (0 on success). */
int S(orc_bmv)(int m, const R* sy, const R* wt, int col, const R* v, R* p) {
  int i, k, info;
  R work[8];
  if (col == 0) return 0;
  V1(p, col + 1) = V1(v, col + 1);
  for (i = 2; i <= col; i++) {
    int i2 = col + i;
    R s = 0.0;
    for (k = 1; k <= i - 1; k++)
      s = s + M1(sy, m, i, k) * V1(v, k) / M1(sy, m, k, k);
    V1(p, i2) = V1(v, i2) + s;
  }
  for (i = 0; i < col; i++) work[i] = p[col + i];
  info = S(orc_dtrsl)(wt, m, col, work, 11);
  for (i = 0; i < col; i++) p[col + i] = work[i];
  if (info != 0) return info;
  for (i = 1; i <= col; i++) V1(p, i) = V1(v, i) / RSQRT(M1(sy, m, i, i));
  for (i = 0; i < col; i++) work[i] = p[col + i];
  info = S(orc_dtrsl)(wt, m, col, work, 1);
  for (i = 0; i < col; i++) p[col + i] = work[i];
  if (info != 0) return info;
  for (i = 1; i <= col; i++) V1(p, i) = -V1(p, i) / RSQRT(M1(sy, m, i, i));
  for (i = 1; i <= col; i++) {
    R s = 0;
    for (k = i + 1; k <= col; k++)
      s = s + M1(sy, m, k, i) * V1(p, col + k) / M1(sy, m, i, i);
    V1(p, i) = V1(p, i) + s;
  }
  return 0;
}

/* Heap helper for the breakpoint search (lbfgsbcpu.cpp:1175-1233).
 * iheap == 0: first build a min-heap of t(1..n).  Then move the least
 * element to position n and restore the heap on t(1..n-1). */
void S(orc_hpsolb)(int n, R* t, int* iorder, int iheap) {
  int i, j, k, indxin, indxou;
  R ddum, dout;
  if (iheap == 0) {
    for (k = 2; k <= n; k++) {
      ddum = V1(t, k);
      indxin = V1(iorder, k);
      i = k;
      while (i > 1) {
        j = i / 2;
        if (!(ddum < V1(t, j))) break;
        V1(t, i) = V1(t, j);
        V1(iorder, i) = V1(iorder, j);
        i = j;
      }
      V1(t, i) = ddum;
      V1(iorder, i) = indxin;
    }
  }
  if (n > 1) {
    i = 1;
    dout = V1(t, 1);
    indxou = V1(iorder, 1);
    ddum = V1(t, n);
    indxin = V1(iorder, n);
    for (;;) {
      j = i + i;
      if (j > n - 1) break;
      if (V1(t, j + 1) < V1(t, j)) j = j + 1;
      if (!(V1(t, j) < ddum)) break;
      V1(t, i) = V1(t, j);
      V1(iorder, i) = V1(iorder, j);
      i = j;
    }
    V1(t, i) = ddum;
    V1(iorder, i) = indxin;
    V1(t, n) = dout;
    V1(iorder, n) = indxou;
  }
}

/* ---------------------------------------------------------------------- */
/* Moré-Thuente step (MINPACK-2 dcstep) as in lbfgsbcpu.cpp:1737-1874.
 * The safeguard `stp + 0.66*(sty - stp)` is evaluated in double there
 * (a double literal), then narrowed by minreal/maxreal<real>. */
void S(orc_dcstep)(R* stx, R* fx, R* dx, R* sty, R* fy, R* dy, R* stp, R fp,
                   R dp, int* brackt, R stpmin, R stpmax) {
  R gamma, p, q, r, s, sgnd, stpc, stpf, stpq, theta;
  sgnd = dp * (*dx / RFABS(*dx));
  if (fp > *fx) {
    theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
    s = S(rmax)(RFABS(theta), S(rmax)(RFABS(*dx), RFABS(dp)));
    gamma = s * RSQRT((theta / s) * (theta / s) - *dx / s * (dp / s));
    if (*stp < *stx) gamma = -gamma;
    p = gamma - *dx + theta;
    q = gamma - *dx + gamma + dp;
    r = p / q;
    stpc = *stx + r * (*stp - *stx);
    stpq = *stx + *dx / ((*fx - fp) / (*stp - *stx) + *dx) / 2 * (*stp - *stx);
    if (RFABS(stpc - *stx) < RFABS(stpq - *stx))
      stpf = stpc;
    else
      stpf = stpc + (stpq - stpc) / 2;
    *brackt = 1;
  } else if (sgnd < 0) {
    theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
    s = S(rmax)(RFABS(theta), S(rmax)(RFABS(*dx), RFABS(dp)));
    gamma = s * RSQRT((theta / s) * (theta / s) - *dx / s * (dp / s));
    if (*stp > *stx) gamma = -gamma;
    p = gamma - dp + theta;
    q = gamma - dp + gamma + *dx;
    r = p / q;
    stpc = *stp + r * (*stx - *stp);
    stpq = *stp + dp / (dp - *dx) * (*stx - *stp);
    if (RFABS(stpc - *stp) > RFABS(stpq - *stp))
      stpf = stpc;
    else
      stpf = stpq;
    *brackt = 1;
  } else if (RFABS(dp) < RFABS(*dx)) {
    theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
    s = S(rmax)(RFABS(theta), S(rmax)(RFABS(*dx), RFABS(dp)));
    gamma = s * RSQRT(S(rmax)((R)0, (theta / s) * (theta / s) -
                                        *dx / s * (dp / s)));
    if (*stp > *stx) gamma = -gamma;
    p = gamma - dp + theta;
    q = gamma + (*dx - dp) + gamma;
    r = p / q;
    if (r < 0 && gamma != 0)
      stpc = *stp + r * (*stx - *stp);
    else if (*stp > *stx)
      stpc = stpmax;
    else
      stpc = stpmin;
    stpq = *stp + dp / (dp - *dx) * (*stx - *stp);
    if (*brackt) {
      double guard = (double)*stp + 0.66 * (double)(*sty - *stp);
      if (RFABS(stpc - *stp) < RFABS(stpq - *stp))
        stpf = stpc;
      else
        stpf = stpq;
      if (*stp > *stx)
        stpf = S(rmin)((R)guard, stpf);
      else
        stpf = S(rmax)((R)guard, stpf);
    } else {
      if (RFABS(stpc - *stp) > RFABS(stpq - *stp))
        stpf = stpc;
      else
        stpf = stpq;
      stpf = S(rmin)(stpmax, stpf);
      stpf = S(rmax)(stpmin, stpf);
    }
  } else {
    if (*brackt) {
      theta = 3 * (fp - *fy) / (*sty - *stp) + *dy + dp;
      s = S(rmax)(RFABS(theta), S(rmax)(RFABS(*dy), RFABS(dp)));
      gamma = s * RSQRT((theta / s) * (theta / s) - *dy / s * (dp / s));
      if (*stp > *sty) gamma = -gamma;
      p = gamma - dp + theta;
      q = gamma - dp + gamma + *dy;
      r = p / q;
      stpc = *stp + r * (*sty - *stp);
      stpf = stpc;
    } else if (*stp > *stx) {
      stpf = stpmax;
    } else {
      stpf = stpmin;
    }
  }
  if (fp > *fx) {
    *sty = *stp;
    *fy = fp;
    *dy = dp;
  } else {
    if (sgnd < 0) {
      *sty = *stx;
      *fy = *fx;
      *dy = *dx;
    }
    *stx = *stp;
    *fx = fp;
    *dx = dp;
  }
  *stp = stpf;
}

/* MINPACK-2 dcsrch as restated in lbfgsbcpu.cpp:1541-1735.
 * task: 0 = START, 1 = FG (evaluate at stp), 2 = ERROR, 3 = WARNING,
 * 4 = CONVERGENCE.  isave[0..1], dsave[0..12] carry the search state. */
void S(orc_dcsrch)(R f, R g, R* stp, R ftol, R gtol, R xtol, R stpmin,
                   R stpmax, int* task, int* isave, R* dsave) {
  int brackt, stage;
  R finit, ftest, fm, fx, fxm, fy, fym, ginit, gtest, gm, gx, gxm, gy, gym;
  R stx, sty, stmin, stmax, width, width1;
  const R xtrapl = (R)1.1, xtrapu = (R)4.0;

  if (*task == 0) {
    if (*stp < stpmin) *task = 2;
    if (*stp > stpmax) *task = 2;
    if (g >= 0) *task = 2;
    if (ftol < 0) *task = 2;
    if (gtol < 0) *task = 2;
    if (xtol < 0) *task = 2;
    if (stpmin < 0) *task = 2;
    if (stpmax < stpmin) *task = 2;
    if (*task == 2) return;
    brackt = 0;
    stage = 1;
    finit = f;
    ginit = g;
    gtest = ftol * ginit;
    width = stpmax - stpmin;
    width1 = width / (R)0.5;
    stx = 0;
    fx = finit;
    gx = ginit;
    sty = 0;
    fy = finit;
    gy = ginit;
    stmin = 0;
    stmax = *stp + xtrapu * *stp;
    *task = 1;
    goto save;
  }
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

  ftest = finit + *stp * gtest;
  if (stage == 1 && f <= ftest && g >= 0) stage = 2;
  if (brackt && (*stp <= stmin || *stp >= stmax)) *task = 3;
  if (brackt && stmax - stmin <= xtol * stmax) *task = 3;
  if (*stp == stpmax && f <= ftest && g <= gtest) *task = 3;
  if (*stp == stpmin && (f > ftest || g >= gtest)) *task = 3;
  if (f <= ftest && RFABS(g) <= gtol * (-ginit)) *task = 4;
  if (*task == 3 || *task == 4) goto save;

  if (stage == 1 && f <= fx && f > ftest) {
    fm = f - *stp * gtest;
    fxm = fx - stx * gtest;
    fym = fy - sty * gtest;
    gm = g - gtest;
    gxm = gx - gtest;
    gym = gy - gtest;
    S(orc_dcstep)(&stx, &fxm, &gxm, &sty, &fym, &gym, stp, fm, gm, &brackt,
                  stmin, stmax);
    fx = fxm + stx * gtest;
    fy = fym + sty * gtest;
    gx = gxm + gtest;
    gy = gym + gtest;
  } else {
    S(orc_dcstep)(&stx, &fx, &gx, &sty, &fy, &gy, stp, f, g, &brackt, stmin,
                  stmax);
  }
  if (brackt) {
    if (RFABS(sty - stx) >= (R)0.66 * width1) *stp = stx + (R)0.5 * (sty - stx);
    width1 = width;
    width = RFABS(sty - stx);
  }
  if (brackt) {
    stmin = S(rmin)(stx, sty);
    stmax = S(rmax)(stx, sty);
  } else {
    stmin = *stp + xtrapl * (*stp - stx);
    stmax = *stp + xtrapu * (*stp - stx);
  }
  *stp = S(rmax)(*stp, stpmin);
  *stp = S(rmin)(*stp, stpmax);
  if ((brackt && (*stp <= stmin || *stp >= stmax)) ||
      (brackt && stmax - stmin <= xtol * stmax))
    *stp = stx;
  *task = 1;
save:
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

/* ---------------------------------------------------------------------- */
/* Objectives (the callback side of the boundary; defined by this project,
 * not by the reference).  Energies are accumulated in double. */
int S(orc_evaluate)(const S(orc_objective)* o, int n, const R* x, R* f,
                    R* g) {
  int i;
  double acc = 0.0;
  switch (o->kind) {
    case ORC_OBJ_GRID_QUAD:
    case ORC_OBJ_GRID_COS: {
      int nx = o->nx, ny = o->ny, ix, iy;
      double h2 = o->h * o->h;
      for (iy = 0; iy < ny; iy++) {
        for (ix = 0; ix < nx; ix++) {
          int k = iy * nx + ix;
          double v = (double)x[k];
          double vl = ix > 0 ? (double)x[k - 1] : 0.0;
          double vr = ix < nx - 1 ? (double)x[k + 1] : 0.0;
          double vd = iy > 0 ? (double)x[k - nx] : 0.0;
          double vu = iy < ny - 1 ? (double)x[k + nx] : 0.0;
          double gk = (4.0 * v - vl - vr - vd - vu) - h2 * (double)o->b[k];
          /* each point owns its right and upper edge; the left and lower
           * boundary edges (to the zero ring) are owned by the first row and
           * column */
          double er = vr - v, eu = vu - v;
          double e = er * er + eu * eu;
          if (ix == 0) e += v * v;
          if (iy == 0) e += v * v;
          acc += 0.5 * e - h2 * (double)o->b[k] * v;
          if (o->kind == ORC_OBJ_GRID_COS) {
            double ck = (double)o->c[k];
            acc += h2 * ck * cos(v);
            gk -= h2 * ck * sin(v);
          }
          g[k] = (R)gk;
        }
      }
      break;
    }
    case ORC_OBJ_ROSEN: {
      for (i = 0; i + 1 < n; i += 2) {
        double a = (double)x[i], b = (double)x[i + 1];
        double t1 = b - a * a, t2 = 1.0 - a;
        acc += 100.0 * t1 * t1 + t2 * t2;
        g[i] = (R)(-400.0 * t1 * a - 2.0 * t2);
        g[i + 1] = (R)(200.0 * t1);
      }
      if (n % 2) g[n - 1] = 0;
      break;
    }
    case ORC_OBJ_PENTA: {
      for (i = 0; i < n; i++) {
        double ax = (double)o->a0[i] * (double)x[i];
        if (i + 1 < n) ax += (double)o->a1[i] * (double)x[i + 1];
        if (i + 2 < n) ax += (double)o->a2[i] * (double)x[i + 2];
        if (i >= 1) ax += (double)o->a1[i - 1] * (double)x[i - 1];
        if (i >= 2) ax += (double)o->a2[i - 2] * (double)x[i - 2];
        acc += 0.5 * ax * (double)x[i] - (double)o->b[i] * (double)x[i];
        g[i] = (R)(ax - (double)o->b[i]);
      }
      break;
    }
    case ORC_OBJ_DSSCFG: {
      /* MINPACK-2 dsscfg as the reference's example evaluates it
       * (examples/dsscfg/dsscfg_cpu.cpp:17-150): per grid point a lower
       * triangle (v, right, top) and an upper one (v, left, bottom), zero
       * outside the grid, trapezoidal rule for the exponential.  The
       * gradient is gathered per point from the five contributions the
       * reference scatters (its own, then bottom, left, right, top arrays),
       * in the reference's order of operations; in double. */
      int nx = o->nx, ny = o->ny, ix, iy;
      double lam = o->h;
      double hx = 1.0 / (double)(nx + 1), hy = 1.0 / (double)(ny + 1);
      double area = 0.5 * hx * hy;
      double fquad = 0.0, fexp = 0.0;
      for (iy = 0; iy < ny; iy++) {
        for (ix = 0; ix < nx; ix++) {
          int k = iy * nx + ix;
          double v = (double)x[k];
          double vr = ix != nx - 1 ? (double)x[k + 1] : 0.0;
          double vt = iy != ny - 1 ? (double)x[k + nx] : 0.0;
          double vl = ix != 0 ? (double)x[k - 1] : 0.0;
          double vb = iy != 0 ? (double)x[k - nx] : 0.0;
          double dxl = (vr - v) / hx, dyl = (vt - v) / hy;
          double dxu = (v - vl) / hx, dyu = (v - vb) / hy;
          double ev = exp(v), ler = lam * ev / 3.0;
          double own, gk;
          fquad += dxl * dxl + dyl * dyl;
          fexp = fexp - lam * (ev + exp(vr) + exp(vt)) / 3.0;
          fquad += dxu * dxu + dyu * dyu;
          fexp = fexp - lam * (exp(vb) + exp(vl) + ev) / 3.0;
          own = 0.0 + (-dxl / hx - dyl / hy - ler);
          own = own + (dxu / hx + dyu / hy - ler);
          gk = own;
          gk = gk + (iy != ny - 1 ? -dyl / hy - ler : 0.0); /* bottom */
          gk = gk + (ix != nx - 1 ? -dxl / hx - ler : 0.0); /* left */
          gk = gk + (ix != 0 ? dxu / hx - ler : 0.0);       /* right */
          gk = gk + (iy != 0 ? dyu / hy - ler : 0.0);       /* top */
          g[k] = (R)(gk * area);
        }
      }
      acc = area * (0.5 * fquad + fexp);
      break;
    }
    case ORC_OBJ_USER:
      return o->user(o->user_ctx, x, f, g);
    default:
      return -1;
  }
  *f = (R)acc;
  return 0;
}

/* ---------------------------------------------------------------------- */
S(orc_state)* S(orc_create)(int n, int m) {
  S(orc_state)* s = (S(orc_state)*)calloc(1, sizeof(S(orc_state)));
  size_t nn = (size_t)(n > 0 ? n : 1);
  int mm = m < 1 ? 1 : (m > 8 ? 8 : m);
  s->n = n;
  s->m = mm;
#define ALLOC_R(p, c) s->p = (R*)calloc((c), sizeof(R))
#define ALLOC_I(p, c) s->p = (int*)calloc((c), sizeof(int))
  ALLOC_R(x, nn); ALLOC_R(g, nn); ALLOC_R(l, nn); ALLOC_R(u, nn);
  ALLOC_R(xold, nn); ALLOC_R(xdiff, nn); ALLOC_R(z, nn); ALLOC_R(r, nn);
  ALLOC_R(d, nn); ALLOC_R(t, nn);
  ALLOC_R(ws, nn * mm); ALLOC_R(wy, nn * mm);
  ALLOC_R(sy, mm * mm); ALLOC_R(ss, mm * mm); ALLOC_R(wt, mm * mm);
  ALLOC_R(wn, 4 * mm * mm); ALLOC_R(snd, 4 * mm * mm);
  ALLOC_I(nbd, nn); ALLOC_I(index, nn); ALLOC_I(iwhere, nn); ALLOC_I(indx2, nn);
#undef ALLOC_R
#undef ALLOC_I
  return s;
}

void S(orc_destroy)(S(orc_state)* s) {
  if (!s) return;
  free(s->x); free(s->g); free(s->l); free(s->u); free(s->xold);
  free(s->xdiff); free(s->z); free(s->r); free(s->d); free(s->t);
  free(s->ws); free(s->wy); free(s->sy); free(s->ss); free(s->wt);
  free(s->wn); free(s->snd); free(s->nbd); free(s->index); free(s->iwhere);
  free(s->indx2);
  free(s);
}

/* Deep copy of all solver state (sizes must match). */
void S(orc_copy)(S(orc_state)* dst, const S(orc_state)* src) {
  size_t n = (size_t)src->n, m = (size_t)src->m;
  S(orc_state) keep = *dst;
  *dst = *src;
#define CP(p, c) do { dst->p = keep.p; memcpy(dst->p, src->p, (c) * sizeof(*src->p)); } while (0)
  CP(x, n); CP(g, n); CP(l, n); CP(u, n); CP(xold, n); CP(xdiff, n);
  CP(z, n); CP(r, n); CP(d, n); CP(t, n); CP(ws, n * m); CP(wy, n * m);
  CP(sy, m * m); CP(ss, m * m); CP(wt, m * m); CP(wn, 4 * m * m);
  CP(snd, 4 * m * m); CP(nbd, n); CP(index, n); CP(iwhere, n); CP(indx2, n);
#undef CP
}

/* lbfgsbactive (lbfgsbcpu.cpp:467-515): project x onto the box, initialise
 * iwhere and the prjctd / cnstnd / boxed flags. */
void S(orc_active)(S(orc_state)* s) {
  int i, n = s->n;
  s->prjctd = 0;
  s->cnstnd = 0;
  s->boxed = 1;
  for (i = 1; i <= n; i++) {
    int nb = V1(s->nbd, i);
    if (nb > 0) {
      if (nb <= 2 && V1(s->x, i) <= V1(s->l, i)) {
        if (V1(s->x, i) < V1(s->l, i)) {
          s->prjctd = 1;
          V1(s->x, i) = V1(s->l, i);
        }
      } else if (nb >= 2 && V1(s->x, i) >= V1(s->u, i)) {
        if (V1(s->x, i) > V1(s->u, i)) {
          s->prjctd = 1;
          V1(s->x, i) = V1(s->u, i);
        }
      }
    }
  }
  for (i = 1; i <= n; i++) {
    int nb = V1(s->nbd, i);
    if (nb != 2) s->boxed = 0;
    if (nb == 0) {
      V1(s->iwhere, i) = -1;
    } else {
      s->cnstnd = 1;
      if (nb == 2 && V1(s->u, i) - V1(s->l, i) <= 0)
        V1(s->iwhere, i) = 3;
      else
        V1(s->iwhere, i) = 0;
    }
  }
}

/* lbfgsbprojgr (lbfgsbcpu.cpp:1389-1414): infinity norm of the projected
 * gradient. */
void S(orc_projgr)(S(orc_state)* s) {
  int i;
  R sbgnrm = 0;
  for (i = 1; i <= s->n; i++) {
    R gi = V1(s->g, i);
    int nb = V1(s->nbd, i);
    if (nb != 0) {
      if (gi < 0) {
        if (nb >= 2) gi = S(rmax)(V1(s->x, i) - V1(s->u, i), gi);
      } else {
        if (nb <= 2) gi = S(rmin)(V1(s->x, i) - V1(s->l, i), gi);
      }
    }
    sbgnrm = S(rmax)(sbgnrm, RFABS(gi));
  }
  s->sbgnrm = sbgnrm;
}

/* lbfgsbcauchy (lbfgsbcpu.cpp:567-817): generalized Cauchy point along the
 * projected steepest-descent path, full breakpoint search with a heap.
 * Inputs x, l, u, nbd, g, iwhere, W, sy, wt, theta, col, head, sbgnrm.
 * Outputs xcp (in s->z), iwhere, p = wa[0..2m), c = wa[2m..4m), nint.
 * Scratch: t (breakpoints), indx2 (iorder), d (the path direction). */
void S(orc_cauchy)(S(orc_state)* s) {
  int n = s->n, m = s->m, col = s->col, head = s->head;
  R* x = s->x;
  R* xcp = s->z;
  R* t = s->t;
  R* d = s->d;
  int* iorder = s->indx2;
  R* p = &s->wa[0];
  R* c = &s->wa[2 * m];
  R* wbp = &s->wa[4 * m];
  R* v = &s->wa[6 * m];
  const R theta = s->theta;
  int i, j, col2, nfree, nbreak, pointr, ibp, nleft, ibkmin, it, bnded;
  R f1, f2, dt, dtm, tsum, dibp, zibp, dibp2, bkmin, tu, tl, wmc, wmp, wmw;
  R tj, tj0, neggi, f2org, tmpv;
  s->internalinfo = 0;

  if (s->sbgnrm <= 0) {
    memcpy(xcp, x, (size_t)n * sizeof(R));
    return;
  }
  bnded = 1;
  nfree = n + 1;
  nbreak = 0;
  ibkmin = 0;
  bkmin = 0;
  col2 = 2 * col;
  f1 = 0;
  for (i = 1; i <= col2; i++) V1(p, i) = 0;
  tl = tu = 0;
  for (i = 1; i <= n; i++) {
    int iw = V1(s->iwhere, i), nb = V1(s->nbd, i);
    neggi = -V1(s->g, i);
    if (iw != 3 && iw != -1) {
      int xlower, xupper;
      tl = 0;
      tu = 0;
      if (nb <= 2) tl = V1(x, i) - V1(s->l, i);
      if (nb >= 2) tu = V1(s->u, i) - V1(x, i);
      xlower = nb <= 2 && tl <= 0;
      xupper = nb >= 2 && tu <= 0;
      iw = 0;
      if (xlower) {
        if (neggi <= 0) iw = 1;
      } else if (xupper) {
        if (neggi >= 0) iw = 2;
      } else {
        if (RFABS(neggi) <= 0) iw = -3;
      }
      V1(s->iwhere, i) = iw;
    }
    pointr = head;
    if (iw != 0 && iw != -1) {
      V1(d, i) = 0;
    } else {
      V1(d, i) = neggi;
      f1 = f1 - neggi * neggi;
      for (j = 1; j <= col; j++) {
        V1(p, j) = V1(p, j) + WY(s, i, pointr) * neggi;
        V1(p, col + j) = V1(p, col + j) + WS(s, i, pointr) * neggi;
        pointr = pointr % m + 1;
      }
      if (nb <= 2 && nb != 0 && neggi < 0) {
        nbreak = nbreak + 1;
        V1(iorder, nbreak) = i;
        V1(t, nbreak) = tl / (-neggi);
        if (nbreak == 1 || V1(t, nbreak) < bkmin) {
          bkmin = V1(t, nbreak);
          ibkmin = nbreak;
        }
      } else if (nb >= 2 && neggi > 0) {
        nbreak = nbreak + 1;
        V1(iorder, nbreak) = i;
        V1(t, nbreak) = tu / neggi;
        if (nbreak == 1 || V1(t, nbreak) < bkmin) {
          bkmin = V1(t, nbreak);
          ibkmin = nbreak;
        }
      } else {
        nfree = nfree - 1;
        V1(iorder, nfree) = i;
        if (RFABS(neggi) > 0) bnded = 0;
      }
    }
  }
  if (theta != 1) {
    for (j = 1; j <= col; j++) V1(p, col + j) *= theta;
  }
  memcpy(xcp, x, (size_t)n * sizeof(R));
  if (nbreak == 0 && nfree == n + 1) return;
  for (j = 1; j <= col2; j++) V1(c, j) = 0;
  f2 = -theta * f1;
  f2org = f2;
  if (col > 0) {
    s->internalinfo = S(orc_bmv)(m, s->sy, s->wt, col, p, v);
    if (s->internalinfo != 0) return;
    tmpv = S(orc_dot)(v, 1, p, 1, col2);
    f2 = f2 - tmpv;
  }
  dtm = -f1 / f2;
  tsum = 0;
  s->nint = 1;
  if (nbreak != 0) {
    nleft = nbreak;
    it = 1;
    tj = 0;
    for (;;) {
      tj0 = tj;
      if (it == 1) {
        tj = bkmin;
        ibp = V1(iorder, ibkmin);
      } else {
        if (it == 2 && ibkmin != nbreak) {
          V1(t, ibkmin) = V1(t, nbreak);
          V1(iorder, ibkmin) = V1(iorder, nbreak);
        }
        S(orc_hpsolb)(nleft, t, iorder, it - 2);
        tj = V1(t, nleft);
        ibp = V1(iorder, nleft);
      }
      dt = tj - tj0;
      if (dtm < dt) break;
      tsum = tsum + dt;
      nleft = nleft - 1;
      it = it + 1;
      dibp = V1(d, ibp);
      V1(d, ibp) = 0;
      if (dibp > 0) {
        zibp = V1(s->u, ibp) - V1(x, ibp);
        V1(xcp, ibp) = V1(s->u, ibp);
        V1(s->iwhere, ibp) = 2;
      } else {
        zibp = V1(s->l, ibp) - V1(x, ibp);
        V1(xcp, ibp) = V1(s->l, ibp);
        V1(s->iwhere, ibp) = 1;
      }
      if (nleft == 0 && nbreak == n) {
        dtm = dt;
        if (col > 0)
          for (j = 0; j < col2; j++) c[j] += dtm * p[j];
        return;
      }
      s->nint = s->nint + 1;
      dibp2 = dibp * dibp;
      f1 = f1 + dt * f2 + dibp2 - theta * dibp * zibp;
      f2 = f2 - theta * dibp2;
      if (col > 0) {
        for (j = 0; j < col2; j++) c[j] += dt * p[j];
        pointr = head;
        for (j = 1; j <= col; j++) {
          V1(wbp, j) = WY(s, ibp, pointr);
          V1(wbp, col + j) = theta * WS(s, ibp, pointr);
          pointr = pointr % m + 1;
        }
        s->internalinfo = S(orc_bmv)(m, s->sy, s->wt, col, wbp, v);
        if (s->internalinfo != 0) return;
        wmc = S(orc_dot)(c, 1, v, 1, col2);
        wmp = S(orc_dot)(p, 1, v, 1, col2);
        wmw = S(orc_dot)(wbp, 1, v, 1, col2);
        for (j = 0; j < col2; j++) p[j] += (-dibp) * wbp[j];
        f1 = f1 + dibp * wmc;
        f2 = f2 + (R)2.0 * dibp * wmp - dibp2 * wmw;
      }
      f2 = S(rmax)((R)(ORC_MACHEPS * (double)f2org), f2);
      if (nleft > 0) {
        dtm = -f1 / f2;
        continue;
      }
      if (bnded) {
        f1 = 0;
        f2 = 0;
        dtm = 0;
      } else {
        dtm = -f1 / f2;
      }
      break;
    }
  }
  if (dtm <= 0) dtm = 0;
  tsum = tsum + dtm;
  for (i = 0; i < n; i++) xcp[i] += tsum * d[i];
  if (col > 0)
    for (j = 0; j < col2; j++) c[j] += dtm * p[j];
}

/* lbfgsbfreev (lbfgsbcpu.cpp:1135-1173): free/active partition of the GCP
 * and the entering/leaving sets against the previous partition.
 * index[0..nfree) = free (ascending), index[nfree..n) = active (descending
 * scan order); indx2[0..nenter) = entering (ascending), indx2[ileave-1..n)
 * = leaving.  Indices are 1-based values, as in the reference. */
void S(orc_freev)(S(orc_state)* s) {
  int n = s->n, i, k, iact;
  s->nenter = 0;
  s->ileave = n + 1;
  if (s->iter > 0 && s->cnstnd) {
    for (i = 1; i <= s->nfree; i++) {
      k = V1(s->index, i);
      if (V1(s->iwhere, k) > 0) {
        s->ileave = s->ileave - 1;
        V1(s->indx2, s->ileave) = k;
      }
    }
    for (i = 1 + s->nfree; i <= n; i++) {
      k = V1(s->index, i);
      if (V1(s->iwhere, k) <= 0) {
        s->nenter = s->nenter + 1;
        V1(s->indx2, s->nenter) = k;
      }
    }
  }
  s->wrk = s->ileave < n + 1 || s->nenter > 0 || s->updatd;
  s->nfree = 0;
  iact = n + 1;
  for (i = 1; i <= n; i++) {
    if (V1(s->iwhere, i) <= 0) {
      s->nfree = s->nfree + 1;
      V1(s->index, s->nfree) = i;
    } else {
      iact = iact - 1;
      V1(s->index, iact) = i;
    }
  }
}

/* lbfgsbformk (lbfgsbcpu.cpp:907-1103): the 2col x 2col matrix
 * K = [[-D - Y'ZZ'Y/theta, L_a' - R_z'], [L_a - R_z, theta S'AA'S]]
 * and its LEL' factorisation in wn; wn1 (s->snd) keeps the inner products
 * over the free (Z) and active (A) sets and is updated incrementally. */
void S(orc_formk)(S(orc_state)* s) {
  int n = s->n, m = s->m, col = s->col, head = s->head, nsub = s->nfree;
  int m2 = 2 * m;
  R* wn = s->wn;
  R* wn1 = s->snd;
  const int* ind = s->index;
  const int* indx2 = s->indx2;
  int ipntr, jpntr, iy, iis, jy, js, is1, js1, k1, i, k, col2, upcl;
  R temp1, temp2, temp3, temp4, v;
  R work[8], workmat[64];

  if (s->updatd) {
    if (s->iupdat > m) {
      /* shift the old part of wn1 up-left by one slot */
      for (jy = 1; jy <= m - 1; jy++) {
        js = m + jy;
        for (i = 0; i < m - jy; i++)
          M1(wn1, m2, jy + i, jy) = M1(wn1, m2, jy + 1 + i, jy + 1);
        for (i = 0; i < m - jy; i++)
          M1(wn1, m2, js + i, js) = M1(wn1, m2, js + 1 + i, js + 1);
        for (i = 0; i < m - 1; i++)
          M1(wn1, m2, m + 1 + i, jy) = M1(wn1, m2, m + 2 + i, jy + 1);
      }
    }
    /* new row and column of wn1 */
    iy = col;
    iis = m + col;
    ipntr = head + col - 1;
    if (ipntr > m) ipntr = ipntr - m;
    jpntr = head;
    for (jy = 1; jy <= col; jy++) {
      js = m + jy;
      temp1 = 0;
      temp2 = 0;
      temp3 = 0;
      for (k = 1; k <= nsub; k++) {
        k1 = V1(ind, k);
        temp1 = temp1 + WY(s, k1, ipntr) * WY(s, k1, jpntr);
      }
      for (k = nsub + 1; k <= n; k++) {
        k1 = V1(ind, k);
        temp2 = temp2 + WS(s, k1, ipntr) * WS(s, k1, jpntr);
        temp3 = temp3 + WS(s, k1, ipntr) * WY(s, k1, jpntr);
      }
      M1(wn1, m2, iy, jy) = temp1;
      M1(wn1, m2, iis, js) = temp2;
      M1(wn1, m2, iis, jy) = temp3;
      jpntr = jpntr % m + 1;
    }
    jy = col;
    jpntr = head + col - 1;
    if (jpntr > m) jpntr = jpntr - m;
    ipntr = head;
    for (i = 1; i <= col; i++) {
      iis = m + i;
      temp3 = 0;
      for (k = 1; k <= nsub; k++) {
        k1 = V1(ind, k);
        temp3 = temp3 + WS(s, k1, ipntr) * WY(s, k1, jpntr);
      }
      ipntr = ipntr % m + 1;
      M1(wn1, m2, iis, jy) = temp3;
    }
    upcl = col - 1;
  } else {
    upcl = col;
  }
  /* incremental corrections over entering (+) and leaving (-) variables */
  ipntr = head;
  for (iy = 1; iy <= upcl; iy++) {
    iis = m + iy;
    jpntr = head;
    for (jy = 1; jy <= iy; jy++) {
      js = m + jy;
      temp1 = temp2 = temp3 = temp4 = 0;
      for (k = 1; k <= s->nenter; k++) {
        k1 = V1(indx2, k);
        temp1 = temp1 + WY(s, k1, ipntr) * WY(s, k1, jpntr);
        temp2 = temp2 + WS(s, k1, ipntr) * WS(s, k1, jpntr);
      }
      for (k = s->ileave; k <= n; k++) {
        k1 = V1(indx2, k);
        temp3 = temp3 + WY(s, k1, ipntr) * WY(s, k1, jpntr);
        temp4 = temp4 + WS(s, k1, ipntr) * WS(s, k1, jpntr);
      }
      M1(wn1, m2, iy, jy) = M1(wn1, m2, iy, jy) + temp1 - temp3;
      M1(wn1, m2, iis, js) = M1(wn1, m2, iis, js) - temp2 + temp4;
      jpntr = jpntr % m + 1;
    }
    ipntr = ipntr % m + 1;
  }
  ipntr = head;
  for (iis = m + 1; iis <= m + upcl; iis++) {
    jpntr = head;
    for (jy = 1; jy <= upcl; jy++) {
      temp1 = temp3 = 0;
      for (k = 1; k <= s->nenter; k++) {
        k1 = V1(indx2, k);
        temp1 = temp1 + WS(s, k1, ipntr) * WY(s, k1, jpntr);
      }
      for (k = s->ileave; k <= n; k++) {
        k1 = V1(indx2, k);
        temp3 = temp3 + WS(s, k1, ipntr) * WY(s, k1, jpntr);
      }
      if (iis <= jy + m)
        M1(wn1, m2, iis, jy) = M1(wn1, m2, iis, jy) + temp1 - temp3;
      else
        M1(wn1, m2, iis, jy) = M1(wn1, m2, iis, jy) - temp1 + temp3;
      jpntr = jpntr % m + 1;
    }
    ipntr = ipntr % m + 1;
  }
  /* assemble the upper triangle of K in wn (2m leading dimension) */
  for (iy = 1; iy <= col; iy++) {
    iis = col + iy;
    is1 = m + iy;
    for (jy = 1; jy <= iy; jy++) {
      js = col + jy;
      js1 = m + jy;
      M1(wn, m2, jy, iy) = M1(wn1, m2, iy, jy) / s->theta;
      M1(wn, m2, js, iis) = M1(wn1, m2, is1, js1) * s->theta;
    }
    for (jy = 1; jy <= iy - 1; jy++) M1(wn, m2, jy, iis) = -M1(wn1, m2, is1, jy);
    for (jy = iy; jy <= col; jy++) M1(wn, m2, jy, iis) = M1(wn1, m2, is1, jy);
    M1(wn, m2, iy, iy) = M1(wn, m2, iy, iy) + M1(s->sy, m, iy, iy);
  }
  s->internalinfo = 0;
  if (!S(orc_dpofa)(wn, m2, col)) {
    s->internalinfo = -1;
    return;
  }
  col2 = 2 * col;
  for (js = col + 1; js <= col2; js++) {
    for (i = 1; i <= col; i++) work[i - 1] = M1(wn, m2, i, js);
    S(orc_dtrsl)(wn, m2, col, work, 11);
    for (i = 1; i <= col; i++) M1(wn, m2, i, js) = work[i - 1];
  }
  for (iis = col + 1; iis <= col2; iis++) {
    for (js = iis; js <= col2; js++) {
      v = S(orc_dot)(&M1(wn, m2, 1, iis), m2, &M1(wn, m2, 1, js), m2, col);
      M1(wn, m2, iis, js) = M1(wn, m2, iis, js) + v;
    }
  }
  for (i = 1; i <= col; i++)
    for (k = 1; k <= col; k++)
      M1(workmat, m, i, k) = M1(wn, m2, col + i, col + k);
  s->internalinfo = 0;
  if (!S(orc_dpofa)(workmat, m, col)) {
    s->internalinfo = -2;
    return;
  }
  for (i = 1; i <= col; i++)
    for (k = 1; k <= col; k++)
      M1(wn, m2, col + i, col + k) = M1(workmat, m, i, k);
}

/* lbfgsbformt (lbfgsbcpu.cpp:1105-1133): T = theta S'S + L D^{-1} L',
 * Cholesky factor into wt. */
void S(orc_formt)(S(orc_state)* s) {
  int m = s->m, col = s->col, i, j, k, k1;
  for (j = 1; j <= col; j++) M1(s->wt, m, 1, j) = s->theta * M1(s->ss, m, 1, j);
  for (i = 2; i <= col; i++) {
    for (j = i; j <= col; j++) {
      R ddum = 0;
      k1 = (i < j ? i : j) - 1;
      for (k = 1; k <= k1; k++)
        ddum = ddum + M1(s->sy, m, i, k) * M1(s->sy, m, j, k) / M1(s->sy, m, k, k);
      M1(s->wt, m, i, j) = ddum + s->theta * M1(s->ss, m, i, j);
    }
  }
  s->internalinfo = 0;
  if (!S(orc_dpofa)(s->wt, m, col)) s->internalinfo = -3;
}

/* lbfgsbcmprlb (lbfgsbcpu.cpp:819-869): reduced gradient
 * r = -theta (z - x) - g + W M c over the free variables (compacted by
 * index order).  Uses wa[2m..4m) = c, overwrites wa[0..2col) with M c. */
void S(orc_cmprlb)(S(orc_state)* s) {
  int n = s->n, m = s->m, col = s->col, i, j, k, pointr;
  R work2[16];
  s->internalinfo = 0;
  if (!s->cnstnd && col > 0) {
    for (i = 1; i <= n; i++) V1(s->r, i) = -V1(s->g, i);
    return;
  }
  for (i = 1; i <= s->nfree; i++) {
    k = V1(s->index, i);
    V1(s->r, i) = -s->theta * (V1(s->z, k) - V1(s->x, k)) - V1(s->g, k);
  }
  for (j = 0; j < 2 * m; j++) work2[j] = s->wa[2 * m + j];
  if (S(orc_bmv)(m, s->sy, s->wt, col, work2, s->wa) != 0) {
    s->internalinfo = -8;
    return;
  }
  pointr = s->head;
  for (j = 1; j <= col; j++) {
    R a1 = V1(s->wa, j);
    R a2 = s->theta * V1(s->wa, col + j);
    for (i = 1; i <= s->nfree; i++) {
      k = V1(s->index, i);
      V1(s->r, i) = V1(s->r, i) + WY(s, k, pointr) * a1 + WS(s, k, pointr) * a2;
    }
    pointr = pointr % m + 1;
  }
}

/* lbfgsbsubsm (lbfgsbcpu.cpp:1416-1539): subspace minimisation over the
 * free variables (2.1 form): Newton direction through the factorised K in
 * wn, then a backtrack to the box that pins the limiting variable.
 * x := s->z (the Cauchy point), d := s->r (reduced gradient, compacted). */
void S(orc_subsm)(S(orc_state)* s) {
  int m = s->m, col = s->col, nsub = s->nfree, pointr, col2, ibd = 0, jy, js;
  int i, j, k;
  R* x = s->z;
  R* d = s->r;
  R* wv = s->wa;
  R alpha, dk, temp1, temp2;
  const R theta = s->theta;
  s->internalinfo = 0;
  if (nsub <= 0) return;
  pointr = s->head;
  for (i = 1; i <= col; i++) {
    temp1 = 0;
    temp2 = 0;
    for (j = 1; j <= nsub; j++) {
      k = V1(s->index, j);
      temp1 = temp1 + WY(s, k, pointr) * V1(d, j);
      temp2 = temp2 + WS(s, k, pointr) * V1(d, j);
    }
    V1(wv, i) = temp1;
    V1(wv, col + i) = theta * temp2;
    pointr = pointr % m + 1;
  }
  col2 = 2 * col;
  s->internalinfo = S(orc_dtrsl)(s->wn, 2 * m, col2, wv, 11);
  if (s->internalinfo != 0) return;
  for (i = 1; i <= col; i++) V1(wv, i) = -V1(wv, i);
  s->internalinfo = S(orc_dtrsl)(s->wn, 2 * m, col2, wv, 1);
  if (s->internalinfo != 0) return;
  pointr = s->head;
  for (jy = 1; jy <= col; jy++) {
    js = col + jy;
    for (i = 1; i <= nsub; i++) {
      k = V1(s->index, i);
      V1(d, i) = V1(d, i) + WY(s, k, pointr) * V1(wv, jy) / theta +
                 WS(s, k, pointr) * V1(wv, js);
    }
    pointr = pointr % m + 1;
  }
  for (i = 1; i <= nsub; i++) V1(d, i) = V1(d, i) / theta;
  alpha = 1;
  temp1 = alpha;
  for (i = 1; i <= nsub; i++) {
    int nb;
    k = V1(s->index, i);
    dk = V1(d, i);
    nb = V1(s->nbd, k);
    if (nb != 0) {
      if (dk < 0 && nb <= 2) {
        temp2 = V1(s->l, k) - V1(x, k);
        if (temp2 >= 0)
          temp1 = 0;
        else if (dk * alpha < temp2)
          temp1 = temp2 / dk;
      } else if (dk > 0 && nb >= 2) {
        temp2 = V1(s->u, k) - V1(x, k);
        if (temp2 <= 0)
          temp1 = 0;
        else if (dk * alpha > temp2)
          temp1 = temp2 / dk;
      }
      if (temp1 < alpha) {
        alpha = temp1;
        ibd = i;
      }
    }
  }
  if (alpha < 1) {
    dk = V1(d, ibd);
    k = V1(s->index, ibd);
    if (dk > 0) {
      V1(x, k) = V1(s->u, k);
      V1(d, ibd) = 0;
    } else if (dk < 0) {
      V1(x, k) = V1(s->l, k);
      V1(d, ibd) = 0;
    }
  }
  for (i = 1; i <= nsub; i++) {
    k = V1(s->index, i);
    V1(x, k) = V1(x, k) + alpha * V1(d, i);
  }
  s->iword = alpha < 1 ? 1 : 0;
}

/* lbfgsblnsrlb (lbfgsbcpu.cpp:1235-1339): one step of the line search.
 * On entry with task != 1 it initialises (dtd, stpmx, the first stp, the
 * copies t = x, r = g, fold = f); then it calls dcsrch and either asks for a
 * new evaluation (task = 1, x set to the trial point) or finishes (task = 5).
 * Sets internalinfo = -4 if d is not a descent direction. */
static void S(orc_lnsrlb)(S(orc_state)* s) {
  int n = s->n, i;
  R big = (R)1.0E10, ftol = (R)1.0E-3, gtol = (R)0.9E0, xtol = (R)0.1E0;
  if (s->task != 1) {
    s->dtd = S(orc_dot)(s->d, 1, s->d, 1, n);
    s->dnrm = RSQRT(s->dtd);
    s->stpmx = big;
    if (s->cnstnd) {
      if (s->iter == 0) {
        s->stpmx = 1;
      } else {
        for (i = 1; i <= n; i++) {
          R a1 = V1(s->d, i), a2;
          int nb = V1(s->nbd, i);
          if (nb != 0) {
            if (a1 < 0 && nb <= 2) {
              a2 = V1(s->l, i) - V1(s->x, i);
              if (a2 >= 0)
                s->stpmx = 0;
              else if (a1 * s->stpmx < a2)
                s->stpmx = a2 / a1;
            } else if (a1 > 0 && nb >= 2) {
              a2 = V1(s->u, i) - V1(s->x, i);
              if (a2 <= 0)
                s->stpmx = 0;
              else if (a1 * s->stpmx > a2)
                s->stpmx = a2 / a1;
            }
          }
        }
      }
    }
    if (s->iter == 0 && !s->boxed)
      s->stp = S(rmin)(1 / s->dnrm, s->stpmx);
    else
      s->stp = 1;
    memcpy(s->t, s->x, (size_t)n * sizeof(R));
    memcpy(s->r, s->g, (size_t)n * sizeof(R));
    s->fold = s->f;
    s->ifun = 0;
    s->iback = 0;
    s->csave = 0;
  }
  s->gd = S(orc_dot)(s->g, 1, s->d, 1, n);
  if (s->ifun == 0) {
    s->gdold = s->gd;
    if (s->gd >= 0) {
      s->internalinfo = -4;
      return;
    }
  }
  S(orc_dcsrch)(s->f, s->gd, &s->stp, ftol, gtol, xtol, (R)0, s->stpmx,
                &s->csave, s->isave, s->dsave);
  s->xstep = s->stp * s->dnrm;
  if (s->csave != 4 && s->csave != 3) {
    s->task = 1;
    s->ifun = s->ifun + 1;
    s->nfgv = s->nfgv + 1;
    s->iback = s->ifun - 1;
    if (s->stp == 1) {
      memcpy(s->x, s->z, (size_t)n * sizeof(R));
    } else {
      for (i = 0; i < n; i++) s->x[i] = s->stp * s->d[i] + s->t[i];
    }
  } else {
    s->task = 5;
  }
}

/* One lnsrlb call on a state (task, d, x, g as the driver leaves them);
 * returns the new task. */
int S(orc_lnsrlb_step)(S(orc_state)* s) {
  S(orc_lnsrlb)(s);
  return s->task;
}

/* lbfgsbmatupd (lbfgsbcpu.cpp:1341-1387): append s = d, y = r to the ring,
 * update theta, shift sy/ss when the ring is full, new row of sy and new
 * column of ss. */
void S(orc_matupd)(S(orc_state)* s) {
  int n = s->n, m = s->m, j, i, pointr;
  if (s->iupdat <= m) {
    s->col = s->iupdat;
    s->itail = (s->head + s->iupdat - 2) % m + 1;
  } else {
    s->itail = s->itail % m + 1;
    s->head = s->head % m + 1;
  }
  for (i = 1; i <= n; i++) {
    WS(s, i, s->itail) = V1(s->d, i);
    WY(s, i, s->itail) = V1(s->r, i);
  }
  s->theta = s->rr / s->dr;
  if (s->iupdat > m) {
    for (j = 1; j <= s->col - 1; j++) {
      for (i = 1; i <= j; i++) M1(s->ss, m, i, j) = M1(s->ss, m, i + 1, j + 1);
      for (i = j; i <= s->col - 1; i++) M1(s->sy, m, i, j) = M1(s->sy, m, i + 1, j + 1);
    }
  }
  pointr = s->head;
  for (j = 1; j <= s->col - 1; j++) {
    M1(s->sy, m, s->col, j) = S(orc_dot)(s->d, 1, &WY(s, 1, pointr), m, n);
    M1(s->ss, m, j, s->col) = S(orc_dot)(&WS(s, 1, pointr), m, s->d, 1, n);
    pointr = pointr % m + 1;
  }
  if (s->stp == 1)
    M1(s->ss, m, s->col, s->col) = s->dtd;
  else
    M1(s->ss, m, s->col, s->col) = s->stp * s->stp * s->dtd;
  M1(s->sy, m, s->col, s->col) = s->dr;
}

/* lbfgsberrclb (lbfgsbcpu.cpp:871-905) */
static int S(orc_errclb)(S(orc_state)* s, double factr) {
  int task = 0, i;
  if (s->n <= 0) task = 2;
  if (s->m <= 0) task = 2;
  if (s->m > s->n) task = 2;
  if (factr < 0) task = 2;
  for (i = 1; i <= s->n; i++) {
    if (V1(s->nbd, i) < 0 || V1(s->nbd, i) > 3) {
      task = 2;
      s->info = -6;
      s->k_err = i;
    }
    if (V1(s->nbd, i) == 2 && V1(s->l, i) > V1(s->u, i)) {
      task = 2;
      s->info = -7;
      s->k_err = i;
    }
  }
  return task;
}

static void S(orc_reset_memory)(S(orc_state)* s) {
  s->internalinfo = 0;
  s->col = 0;
  s->head = 1;
  s->theta = 1;
  s->iupdat = 0;
  s->updatd = 0;
  s->restarts++;
}

static void S(orc_eval_state)(S(orc_state)* s) {
  S(orc_evaluate)(&s->obj, s->n, s->x, &s->f, s->g);
}

/* Driver set-up (lbfgsbcpu.cpp:117-278).  Returns 1 if iterations should
 * follow, 0 if the solve already ended (info set in s->info). */
int S(orc_setup)(S(orc_state)* s, const R* x0, const int* nbd, const R* l,
                 const R* u, const S(orc_options)* opt,
                 const S(orc_objective)* obj) {
  int n = s->n;
  s->opt = *opt;
  s->obj = *obj;
  s->m = opt->m < 8 ? opt->m : 8;
  memcpy(s->x, x0, (size_t)n * sizeof(R));
  memcpy(s->nbd, nbd, (size_t)n * sizeof(int));
  memcpy(s->l, l, (size_t)n * sizeof(R));
  memcpy(s->u, u, (size_t)n * sizeof(R));
  s->col = 0;
  s->head = 1;
  s->theta = 1;
  s->iupdat = 0;
  s->updatd = 0;
  s->iter = 0;
  s->nfgv = 0;
  s->nint = 0;
  s->nintol = 0;
  s->nskip = 0;
  s->nfree = n;
  s->internalinfo = 0;
  s->restarts = 0;
  s->info = 0;
  s->task = 0;
  if (S(orc_errclb)(s, opt->eps_f) == 2 || opt->max_iteration < 0 ||
      opt->eps_g < 0 || opt->eps_x < 0) {
    s->info = -1;
    return 0;
  }
  s->info = 0;
  S(orc_active)(s);
  memcpy(s->xold, s->x, (size_t)n * sizeof(R));
  S(orc_eval_state)(s);
  s->nfgv = 1;
  S(orc_projgr)(s);
  s->residual_g = s->sbgnrm;
  if (s->sbgnrm <= (R)opt->eps_g) {
    s->info = 4;
    return 0;
  }
  return 1;
}

/* Head of an outer iteration (lbfgsbcpu.cpp:280-345): Cauchy point, free
 * set, formk, cmprlb and subsm, each failure resetting the memory and
 * starting the head again (:297-305, 317-325, 333-341); ends with
 * d = z - x and task = 0. */
void S(orc_iter_head)(S(orc_state)* s) {
  int n = s->n, i;
  for (;;) {
    s->iword = -1;
    if (!s->cnstnd && s->col > 0) {
      memcpy(s->z, s->x, (size_t)n * sizeof(R));
      s->wrk = s->updatd;
      s->nint = 0;
    } else {
      S(orc_cauchy)(s);
      if (s->internalinfo != 0) {
        S(orc_reset_memory)(s);
        continue;
      }
      s->nintol = s->nintol + s->nint;
      S(orc_freev)(s);
      s->nact = n - s->nfree;
    }
    if (s->nfree != 0 && s->col != 0) {
      if (s->wrk) S(orc_formk)(s);
      if (s->internalinfo != 0) {
        S(orc_reset_memory)(s);
        continue;
      }
      S(orc_cmprlb)(s);
      if (s->internalinfo == 0) S(orc_subsm)(s);
      if (s->internalinfo != 0) {
        S(orc_reset_memory)(s);
        continue;
      }
    }
    for (i = 0; i < n; i++) s->d[i] = s->z[i] - s->x[i];
    s->task = 0;
    return;
  }
}

/* Line-search loop (lbfgsbcpu.cpp:346-358): lnsrlb and an evaluation per
 * trial, left after an error, 20 backtracks or a finished search. */
void S(orc_iter_search)(S(orc_state)* s) {
  for (;;) {
    S(orc_lnsrlb)(s);
    if (s->internalinfo != 0 || s->iback >= 20 || s->task != 1) break;
    S(orc_eval_state)(s);
  }
}

/* The rest of an iteration (lbfgsbcpu.cpp:359-460): line-search failure,
 * stopping tests in the reference's order (projected gradient, step,
 * decrease, iteration count, customised test) and the memory update. */
int S(orc_iter_tail)(S(orc_state)* s) {
  int n = s->n, i;
  const S(orc_options)* opt = &s->opt;
  if (s->internalinfo != 0) {
    memcpy(s->x, s->t, (size_t)n * sizeof(R));
    memcpy(s->g, s->r, (size_t)n * sizeof(R));
    s->f = s->fold;
    if (s->col == 0) {
      s->task = 2;
      s->iter = s->iter + 1;
      s->info = -2;
      return ORC_TAIL_END;
    }
    S(orc_reset_memory)(s);
    return ORC_TAIL_RESTART;
  }
  s->iter = s->iter + 1;
  S(orc_projgr)(s);
  s->residual_g = s->sbgnrm;
  if (s->sbgnrm <= (R)opt->eps_g) {
    s->info = 4;
    return ORC_TAIL_END;
  }
  for (i = 0; i < n; i++) s->xdiff[i] = s->xold[i] - s->x[i];
  s->tf = S(orc_dot)(s->xdiff, 1, s->xdiff, 1, n);
  s->tf = RSQRT(s->tf);
  s->residual_x = s->tf;
  if (s->tf <= (R)opt->eps_x) {
    s->info = 2;
    return ORC_TAIL_END;
  }
  s->ddum = S(rmax)(RFABS(s->fold), S(rmax)(RFABS(s->f), (R)1));
  s->residual_f = (s->fold - s->f) / s->ddum;
  if (s->fold - s->f <= (R)opt->eps_f * s->ddum) {
    s->info = 1;
    return ORC_TAIL_END;
  }
  if (s->iter > opt->max_iteration && opt->max_iteration > 0) {
    s->info = 5;
    return ORC_TAIL_END;
  }
  if (opt->stop_at_iter > 0 && s->iter >= opt->stop_at_iter) {
    s->info = 0;
    return ORC_TAIL_END;
  }
  memcpy(s->xold, s->x, (size_t)n * sizeof(R));
  for (i = 0; i < n; i++) s->r[i] = s->g[i] - s->r[i];
  s->rr = S(orc_dot)(s->r, 1, s->r, 1, n);
  if (s->stp == 1) {
    s->dr = s->gd - s->gdold;
    s->ddum = -s->gdold;
  } else {
    s->dr = (s->gd - s->gdold) * s->stp;
    for (i = 0; i < n; i++) s->d[i] *= s->stp;
    s->ddum = -s->gdold * s->stp;
  }
  if ((double)s->dr <= ORC_MACHEPS * (double)s->ddum) {
    s->nskip = s->nskip + 1;
    s->updatd = 0;
  } else {
    s->updatd = 1;
    s->iupdat = s->iupdat + 1;
    S(orc_matupd)(s);
    S(orc_formt)(s);
    /* a failed factorisation discards the memory; the reference's loop
     * then goes on with the next iteration (:450-458) */
    if (s->internalinfo != 0) S(orc_reset_memory)(s);
  }
  return ORC_TAIL_NEXT;
}

static void S(orc_fill_record)(const S(orc_state)* s, S(orc_record)* rec,
                               int restarts0, int info) {
  rec->f = s->f; rec->sbgnrm = s->sbgnrm; rec->stp = s->stp;
  rec->theta = s->theta; rec->dtd = s->dtd; rec->gd = s->gd;
  rec->residual_f = s->residual_f; rec->residual_x = s->residual_x;
  rec->iter = s->iter; rec->nint = s->nint; rec->nfree = s->nfree;
  rec->nenter = s->nenter; rec->ileave = s->ileave; rec->wrk = s->wrk;
  rec->iword = s->iword; rec->iback = s->iback; rec->nfgv = s->nfgv;
  rec->col = s->col; rec->updatd = s->updatd;
  rec->restarts = s->restarts - restarts0; rec->info = info;
}

/* One outer iteration (lbfgsbcpu.cpp:279-462), including the memory resets
 * that restart it.  Returns 1 to continue, 0 when the solve has ended. */
int S(orc_iterate)(S(orc_state)* s, S(orc_record)* rec) {
  int restarts0 = s->restarts, r;
  for (;;) {
    S(orc_iter_head)(s);
    S(orc_iter_search)(s);
    r = S(orc_iter_tail)(s);
    if (r != ORC_TAIL_RESTART) break;
  }
  if (rec) S(orc_fill_record)(s, rec, restarts0, r == ORC_TAIL_END ? s->info : 0);
  return r == ORC_TAIL_NEXT;
}

/* Whole solve: lbfgsbminimize in LCM_NO_ACCELERATION mode
 * (lbfgsb.cpp:31-34 -> lbfgsbcpu.cpp:117-465).  Writes x back on exit. */
int S(orc_minimize)(int n, const S(orc_options)* opt,
                    const S(orc_objective)* obj, R* x, const int* nbd,
                    const R* l, const R* u, S(orc_record)* records,
                    int max_records, int* out_iters, R* out_f) {
  S(orc_state)* s = S(orc_create)(n, opt->m);
  int k = 0, info;
  S(orc_record) tmp;
  if (S(orc_setup)(s, x, nbd, l, u, opt, obj)) {
    for (;;) {
      S(orc_record)* rec = (records && k < max_records) ? &records[k] : &tmp;
      int more = S(orc_iterate)(s, rec);
      k++;
      if (!more) break;
    }
  }
  if (s->info != -1) memcpy(x, s->x, (size_t)n * sizeof(R));
  if (out_iters) *out_iters = s->iter;
  if (out_f) *out_f = s->f;
  info = s->info;
  S(orc_destroy)(s);
  return info;
}

#undef V1
#undef M1
#undef WS
#undef WY
#undef RFABS
#undef RSQRT
