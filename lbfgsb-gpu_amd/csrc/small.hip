// Single-workgroup FP64 control kernels: they finish the per-block partial
// sums of the bulk passes and run the scalar logic of the reference driver
// on the device (lbfgsbcpu.cpp:279-462), so that a whole iteration can be
// enqueued without a host round trip.  One launch = one block of 256
// threads; partial reductions are spread over the four waves, the scalar
// logic runs on thread 0.
#include "common.hpp"
#include "dense.hpp"
#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

#define A1(a, ld, i, j) (a)[((i)-1) * (ld) + ((j)-1)]

static constexpr double kMachEps = 5e-16;  // culbfgsb/ap.cpp:38, both precisions
static constexpr double kBig = 1.0e10;     // lbfgsbcpu.cpp:1259
static constexpr double kFtol = 1.0e-3, kGtol = 0.9, kXtol = 0.1;

// Compensated reduction of nv value rows (each G long) into out[0..nv).
__device__ void reduce_rows(const double* part, int G, int v0, int nv, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int v = w; v < nv; v += kWavesPerBlock) {
    const double* row = part + (size_t)(v0 + v) * G;
    double s = 0.0, c = 0.0;
    for (int b = lane; b < G; b += kWave) {
      double e;
      two_sum(s, row[b], s, e);
      c += e;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double os = __shfl_xor(s, o, kWave), oc = __shfl_xor(c, o, kWave), e;
      two_sum(s, os, s, e);
      c += oc + e;
    }
    if (lane == 0) out[v] = s + c;
  }
  __syncthreads();
}

__device__ double reduce_max(const double* pm, int G, double* sh) {
  double m = -__builtin_inf();
  for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmax(m, pm[b]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < kWavesPerBlock; w++) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}

__device__ void reduce_argmin(const double* pv, const long long* pi, int G, double* shv,
                              long long* shi, double& v, long long& idx) {
  double bv = __builtin_inf();
  long long bi = 0x7fffffffffffffffLL;
  for (int b = threadIdx.x; b < G; b += blockDim.x) {
    if (pv[b] < bv || (pv[b] == bv && pi[b] < bi)) {
      bv = pv[b];
      bi = pi[b];
    }
  }
  wave_argmin(bv, bi);
  if ((threadIdx.x & 63) == 0) {
    shv[threadIdx.x >> 6] = bv;
    shi[threadIdx.x >> 6] = bi;
  }
  __syncthreads();
  v = shv[0];
  idx = shi[0];
  for (int w = 1; w < kWavesPerBlock; w++)
    if (shv[w] < v || (shv[w] == v && shi[w] < idx)) {
      v = shv[w];
      idx = shi[w];
    }
  __syncthreads();
}

__device__ void reduce_min(const double* pv, int G, double* sh, double& v) {
  double m = __builtin_inf();
  for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmin(m, pv[b]);
  m = -wave_max(-m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  v = sh[0];
  for (int w = 1; w < kWavesPerBlock; w++) v = fmin(v, sh[w]);
  __syncthreads();
}

// BFGS memory reset used by every restart of the reference driver
// (lbfgsbcpu.cpp:297-304, 317-324, 333-340, 380-385, 452-459).
__device__ void reset_memory(DevScalars* ds) {
  ds->internalinfo = 0;
  ds->col = 0;
  ds->head = 0;
  ds->theta = 1.0;
  ds->iupdat = 0;
  ds->updatd = 0;
  ds->restarts++;
}

__device__ void set_done(DevScalars* ds, int info) {
  ds->info = info;
  ds->status = ST_DONE;
  ds->halt = 1;
}

// ---------------------------------------------------------------------------
__global__ void s_resume(DevScalars* ds, int reset_bp) {
  if (threadIdx.x == 0) {
    ds->halt = 0;
    ds->status = ST_RUNNING;
    ds->write_trial = 0;
    if (reset_bp) ds->bp_count = 0;
  }
}

__global__ void s_setup_errclb(DevScalars* ds, const double* part, int G) {
  __shared__ double r[2];
  reduce_rows(part, G, 0, 2, r);
  if (threadIdx.x == 0 && (r[0] > 0 || r[1] > 0)) set_done(ds, -1);
}

__global__ void s_setup_active(DevScalars* ds, const double* part, int G) {
  __shared__ double r[3];
  reduce_rows(part, G, 0, 3, r);
  if (threadIdx.x == 0) {
    ds->prjctd = r[0] > 0;
    ds->cnstnd = r[1] > 0;
    ds->boxed = !(r[2] > 0);
  }
}

// First projected-gradient test after the initial evaluation
// (lbfgsbcpu.cpp:271-278).
__global__ void s_setup_projgr(DevScalars* ds, const double* pmax, int G, double f) {
  __shared__ double sh[kWavesPerBlock];
  double m = reduce_max(pmax, G, sh);
  if (threadIdx.x == 0) {
    ds->f = f;
    ds->nfgv = 1;
    ds->sbgnrm = m;
    ds->residual_g = m;
    if (m <= ds->eps_g) set_done(ds, 4);
  }
}

// Stopping tests after an accepted step (lbfgsbcpu.cpp:394-429).
__global__ void s_stop_tests(DevScalars* ds, const double* part, const double* pmax,
                             const double* pmin, const long long* pmini, int G) {
  __shared__ double r[PA_COUNT];
  __shared__ double sh[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if (ds->halt) {
    if (threadIdx.x == 0) ds->write_trial = 0;
    return;
  }
  reduce_rows(part, G, 0, PA_COUNT, r);
  if (threadIdx.x == 0) ds->write_trial = 0;
  double sbg = reduce_max(pmax, G, sh);
  double bk;
  long long ibk;
  reduce_argmin(pmin, pmini, G, sh, shi, bk, ibk);
  if (threadIdx.x != 0) return;
  for (int v = 0; v < PA_COUNT; v++) ds->pa[v] = r[v];
  ds->gcp_bkmin = bk;
  ds->gcp_ibkmin = ibk;
  ds->sbgnrm = sbg;
  ds->residual_g = sbg;
  if (sbg <= ds->eps_g) return set_done(ds, 4);
  double tf = sqrt(r[PA_TF]);
  ds->tf = tf;
  ds->residual_x = tf;
  if (tf <= ds->eps_x) return set_done(ds, 2);
  double ddum = dmax_(fabs(ds->fold), dmax_(fabs(ds->f), 1.0));
  ds->ddum = ddum;
  ds->residual_f = (ds->fold - ds->f) / ddum;
  if (ds->fold - ds->f <= ds->eps_f * ddum) return set_done(ds, 1);
  if (ds->iter > ds->max_iteration && ds->max_iteration > 0) return set_done(ds, 5);
  if (ds->has_custom_stop) {
    ds->status = ST_CHECKED;
    ds->halt = 1;
  }
}

// c <- M c for the reduced gradient (lbfgsbcpu.cpp:851-857); the same
// factor already passed in the Cauchy search, so this cannot fail there.
__device__ bool gcp_finish(DevScalars* ds) {
  const int col = ds->col;
  if (col > 0) {
    if (d_bmv(ds->sy, ds->wt, col, ds->c, ds->mc) != 0) return false;
  }
  ds->nintol += ds->nint;
  return true;
}

// Cauchy-search controller (lbfgsbcpu.cpp:620-724 and the first step of the
// breakpoint loop :725-747).  Returns true if the sorted breakpoint rounds
// are needed.
__device__ void gcp_controller(DevScalars* ds) {
  for (;;) {
    const int col = ds->col;
    ds->iword = -1;
    ds->shortcut = (!ds->cnstnd && col > 0);
    if (ds->shortcut) {
      ds->wrk = ds->updatd;
      ds->nint = 0;
      ds->gcp_mode = 2;
      ds->status = ST_RUNNING;
      return;
    }
    const double theta = ds->theta;
    const int nbreak = (int)ds->pa[PA_NBREAK];
    const int nfreegcp = (int)ds->pa[PA_NFREEGCP];
    ds->gcp_nbreak = nbreak;
    ds->gcp_nfreegcp = nfreegcp;
    ds->gcp_unbnd = ds->pa[PA_UNBND] > 0;
    const double f1 = -ds->pa[PA_F1];
    ds->gcp_D0 = ds->pa[PA_F1];
    for (int j = 0; j < col; j++) {
      ds->p[j] = ds->pa[PA_P + j];
      ds->p[col + j] = theta != 1.0 ? ds->pa[PA_P + 8 + j] * theta : ds->pa[PA_P + 8 + j];
    }
    if (nbreak == 0 && nfreegcp == 0) {
      // nothing moves: x_cp = x, c and nint keep their previous values
      ds->gcp_mode = 2;
      if (!gcp_finish(ds)) {
        reset_memory(ds);
        continue;
      }
      ds->status = ST_RUNNING;
      return;
    }
    for (int j = 0; j < 2 * col; j++) ds->c[j] = 0.0;
    double f2 = -theta * f1;
    const double f2org = f2;
    if (col > 0) {
      if (d_bmv(ds->sy, ds->wt, col, ds->p, ds->v) != 0) {
        reset_memory(ds);
        continue;
      }
      f2 = f2 - ddot(ds->v, 1, ds->p, 1, 2 * col);
    }
    double dtm = -f1 / f2;
    ds->gcp_f1 = f1;
    ds->gcp_f2 = f2;
    ds->gcp_f2org = f2org;
    ds->nint = 1;
    ds->gcp_kpassed = 0;
    ds->gcp_t_prev = 0.0;
    ds->gcp_idx_prev = -1;
    if (nbreak == 0 || dtm < ds->gcp_bkmin) {
      if (dtm <= 0) dtm = 0;
      ds->gcp_dtm = dtm;
      ds->gcp_tsum = dtm;
      for (int j = 0; j < 2 * col; j++) ds->c[j] += dtm * ds->p[j];
      ds->gcp_mode = 0;
      if (!gcp_finish(ds)) {
        reset_memory(ds);
        continue;
      }
      ds->status = ST_RUNNING;
      return;
    }
    ds->status = ST_GCP_ROUNDS;
    ds->halt = 1;
    return;
  }
}

// matupd's small part + formt (lbfgsbcpu.cpp:1365-1386, 1105-1133), then
// the next iteration's Cauchy controller.  `reduce` selects the standalone
// path (iteration start / restart), which reduces the pass-A partials here.
__global__ void s_update_gcp(DevScalars* ds, const double* part, const double* pmin,
                             const long long* pmini, int G, int reduce) {
  __shared__ double r[PA_COUNT];
  __shared__ double sh[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if (ds->halt) {
    if (threadIdx.x == 0) ds->write_trial = 0;
    return;
  }
  if (reduce) {
    reduce_rows(part, G, 0, PA_COUNT, r);
    double bk;
    long long ibk;
    reduce_argmin(pmin, pmini, G, sh, shi, bk, ibk);
    if (threadIdx.x == 0) {
      for (int v = 0; v < PA_COUNT; v++) ds->pa[v] = r[v];
      ds->gcp_bkmin = bk;
      ds->gcp_ibkmin = ibk;
    }
  }
  if (threadIdx.x != 0) return;
  ds->write_trial = 0;
  if (!reduce && ds->updatd) {
    const int m = ds->m, col = ds->col;
    ds->rr = ds->pa[PA_RR];
    ds->theta = ds->rr / ds->dr;
    if (ds->iupdat > m) {
      for (int j = 1; j <= col - 1; j++) {
        for (int i = 1; i <= j; i++) A1(ds->ss, 8, i, j) = A1(ds->ss, 8, i + 1, j + 1);
        for (int i = j; i <= col - 1; i++) A1(ds->sy, 8, i, j) = A1(ds->sy, 8, i + 1, j + 1);
      }
    }
    for (int j = 1; j <= col - 1; j++) {
      A1(ds->sy, 8, col, j) = ds->pa[PA_SY + j - 1];
      A1(ds->ss, 8, j, col) = ds->pa[PA_SS + j - 1];
    }
    A1(ds->ss, 8, col, col) = ds->stp == 1.0 ? ds->dtd : ds->stp * ds->stp * ds->dtd;
    A1(ds->sy, 8, col, col) = ds->dr;
    if (!d_formt(ds->wt, ds->sy, ds->ss, col, ds->theta)) reset_memory(ds);
  }
  gcp_controller(ds);
}

// formk (lbfgsbcpu.cpp:907-1103) from the inner products of k_p4.
__device__ bool formk_device(DevScalars* ds, const double* nc, const double* corr) {
  const int m = ds->m, col = ds->col;
  double* wn1 = ds->wn1;
  double* wn = ds->wn;
  const double theta = ds->theta;
  int upcl;
  if (ds->updatd) {
    if (ds->iupdat > m) {
      for (int jy = 1; jy <= m - 1; jy++) {
        const int js = m + jy;
        for (int i = 0; i < m - jy; i++) A1(wn1, 16, jy + i, jy) = A1(wn1, 16, jy + 1 + i, jy + 1);
        for (int i = 0; i < m - jy; i++) A1(wn1, 16, js + i, js) = A1(wn1, 16, js + 1 + i, js + 1);
        for (int i = 0; i < m - 1; i++) A1(wn1, 16, m + 1 + i, jy) = A1(wn1, 16, m + 2 + i, jy + 1);
      }
    }
    for (int jy = 1; jy <= col; jy++) {
      A1(wn1, 16, col, jy) = nc[jy - 1];
      A1(wn1, 16, m + col, m + jy) = nc[8 + jy - 1];
      A1(wn1, 16, m + col, jy) = nc[16 + jy - 1];
    }
    for (int i = 1; i <= col; i++) A1(wn1, 16, m + i, col) = nc[24 + i - 1];
    upcl = col - 1;
  } else {
    upcl = col;
  }
  for (int iy = 1; iy <= upcl; iy++) {
    for (int jy = 1; jy <= iy; jy++) {
      const int q = (iy - 1) * iy / 2 + (jy - 1);
      A1(wn1, 16, iy, jy) = A1(wn1, 16, iy, jy) + corr[q];
      A1(wn1, 16, m + iy, m + jy) = A1(wn1, 16, m + iy, m + jy) - corr[36 + q];
    }
  }
  for (int i = 1; i <= upcl; i++) {
    for (int jy = 1; jy <= upcl; jy++) {
      const double v = corr[72 + (i - 1) * 8 + (jy - 1)];
      if (i <= jy)
        A1(wn1, 16, m + i, jy) = A1(wn1, 16, m + i, jy) + v;
      else
        A1(wn1, 16, m + i, jy) = A1(wn1, 16, m + i, jy) - v;
    }
  }
  for (int iy = 1; iy <= col; iy++) {
    const int iis = col + iy, is1 = m + iy;
    for (int jy = 1; jy <= iy; jy++) {
      const int js = col + jy, js1 = m + jy;
      A1(wn, 16, jy, iy) = A1(wn1, 16, iy, jy) / theta;
      A1(wn, 16, js, iis) = A1(wn1, 16, is1, js1) * theta;
    }
    for (int jy = 1; jy <= iy - 1; jy++) A1(wn, 16, jy, iis) = -A1(wn1, 16, is1, jy);
    for (int jy = iy; jy <= col; jy++) A1(wn, 16, jy, iis) = A1(wn1, 16, is1, jy);
    A1(wn, 16, iy, iy) = A1(wn, 16, iy, iy) + A1(ds->sy, 8, iy, iy);
  }
  if (!d_dpofa(wn, 16, col)) {
    ds->internalinfo = -1;
    return false;
  }
  const int col2 = 2 * col;
  double work[8];
  for (int js = col + 1; js <= col2; js++) {
    for (int i = 1; i <= col; i++) work[i - 1] = A1(wn, 16, i, js);
    d_dtrsl(wn, 16, col, work, 11);
    for (int i = 1; i <= col; i++) A1(wn, 16, i, js) = work[i - 1];
  }
  for (int iis = col + 1; iis <= col2; iis++)
    for (int js = iis; js <= col2; js++)
      A1(wn, 16, iis, js) = A1(wn, 16, iis, js) +
                            ddot(&A1(wn, 16, 1, iis), 16, &A1(wn, 16, 1, js), 16, col);
  double wm[64];
  for (int i = 1; i <= col; i++)
    for (int k = 1; k <= col; k++) A1(wm, 8, i, k) = A1(wn, 16, col + i, col + k);
  if (!d_dpofa(wm, 8, col)) {
    ds->internalinfo = -2;
    return false;
  }
  for (int i = 1; i <= col; i++)
    for (int k = 1; k <= col; k++) A1(wn, 16, col + i, col + k) = A1(wm, 8, i, k);
  return true;
}

// After k_p4: free/active bookkeeping (lbfgsbcpu.cpp:1143-1161), formk,
// and the two triangular solves of subsm (lbfgsbcpu.cpp:1459-1471).
__global__ void s_formk_subsm(DevScalars* ds, const double* part, int G) {
  __shared__ double r[P4_COUNT];
  if (ds->halt) return;
  reduce_rows(part, G, 0, P4_COUNT, r);
  if (threadIdx.x != 0) return;
  const long long n = ds->n;
  const int col = ds->col;
  if (!ds->shortcut) {
    ds->nfree = (int)r[P4_NFREE];
    if (ds->iter > 0 && ds->cnstnd) {
      ds->nenter = (int)r[P4_NENTER];
      ds->nleave = (int)r[P4_NLEAVE];
    } else {
      ds->nenter = 0;
      ds->nleave = 0;
    }
    ds->ileave = (int)(n + 1 - ds->nleave);
    ds->wrk = ds->nleave > 0 || ds->nenter > 0 || ds->updatd;
  }
  ds->skip_subsm = (ds->nfree == 0 || col == 0);
  if (ds->skip_subsm) return;
  if (ds->wrk) {
    if (!formk_device(ds, r + P4_NC, r + P4_CORR)) {
      reset_memory(ds);
      ds->status = ST_RESTART;
      ds->halt = 1;
      return;
    }
  }
  const double theta = ds->theta;
  double* wv = ds->wv;
  for (int i = 0; i < col; i++) {
    wv[i] = r[P4_WV + i];
    wv[col + i] = theta * r[P4_WV + 8 + i];
  }
  int info = d_dtrsl(ds->wn, 16, 2 * col, wv, 11);
  if (info == 0) {
    for (int i = 0; i < col; i++) wv[i] = -wv[i];
    info = d_dtrsl(ds->wn, 16, 2 * col, wv, 1);
  }
  if (info != 0) {
    reset_memory(ds);
    ds->status = ST_RESTART;
    ds->halt = 1;
  }
}

// Backtrack length of subsm (lbfgsbcpu.cpp:1484-1538).
__global__ void s_alpha(DevScalars* ds, const double* pmin, const long long* pmini, int G) {
  __shared__ double sh[kWavesPerBlock];
  __shared__ long long shi[kWavesPerBlock];
  if (ds->halt || ds->skip_subsm) return;
  double a;
  long long i;
  reduce_argmin(pmin, pmini, G, sh, shi, a, i);
  if (threadIdx.x == 0) {
    if (a < 1.0) {
      ds->alpha = a;
      ds->ibd = i;
      ds->iword = 1;
    } else {
      ds->alpha = 1.0;
      ds->ibd = -1;
      ds->iword = 0;
    }
  }
}

__device__ void request_trial(DevScalars* ds) {
  ds->task = 1;
  ds->ifun = ds->ifun + 1;
  ds->nfgv = ds->nfgv + 1;
  ds->iback = ds->ifun - 1;
  ds->write_trial = 1;
  ds->trial_src = ds->phase & 1;
  ds->trial_dst = (ds->phase & 1) ^ 1;
}

// Step accepted: the reference increments iter and, after its stopping
// tests, decides the update (lbfgsbcpu.cpp:388-449).  The decision only
// depends on line-search scalars, so it is taken here and the ring slot is
// chosen before the fused tail pass writes it.
__device__ void accept_step(DevScalars* ds) {
  ds->phase ^= 1;
  ds->iter = ds->iter + 1;
  ds->bp_count = 0;
  const double stp = ds->stp;
  if (stp == 1.0) {
    ds->dr = ds->gd - ds->gdold;
    ds->ddum = -ds->gdold;
  } else {
    ds->dr = (ds->gd - ds->gdold) * stp;
    ds->ddum = -ds->gdold * stp;
  }
  if (ds->dr <= kMachEps * ds->ddum) {
    ds->nskip = ds->nskip + 1;
    ds->updatd = 0;
  } else {
    const int m = ds->m;
    ds->updatd = 1;
    ds->iupdat = ds->iupdat + 1;
    if (ds->iupdat <= m) {
      ds->col = ds->iupdat;
      ds->itail = (ds->head + ds->iupdat - 1) % m;
    } else {
      ds->itail = (ds->itail + 1) % m;
      ds->head = (ds->head + 1) % m;
    }
  }
  if (ds->has_after_iteration) {
    ds->status = ST_ACCEPTED;
    ds->halt = 1;
  } else {
    ds->status = ST_RUNNING;
  }
}

// Line-search failure (lbfgsbcpu.cpp:359-387): the iterate and gradient are
// still the accepted ones (buffers were never swapped), f is restored.
__device__ void ls_failure(DevScalars* ds) {
  ds->f = ds->fold;
  ds->write_trial = 0;
  if (ds->col == 0) {
    ds->task = 2;
    ds->iter = ds->iter + 1;
    set_done(ds, -2);
    return;
  }
  reset_memory(ds);
  ds->status = ST_RESTART;
  ds->halt = 1;
}

// First call of lnsrlb for an iteration (lbfgsbcpu.cpp:1263-1339).
__global__ void s_ls_init(DevScalars* ds, const double* part, const double* pmin, int G) {
  __shared__ double r[P5_COUNT];
  __shared__ double sh[kWavesPerBlock];
  if (ds->halt) return;
  reduce_rows(part, G, 0, P5_COUNT, r);
  double smin;
  reduce_min(pmin, G, sh, smin);
  if (threadIdx.x != 0) return;
  ds->dtd = r[P5_DTD];
  ds->dnrm = sqrt(ds->dtd);
  ds->stpmx = kBig;
  if (ds->cnstnd) {
    if (ds->iter == 0)
      ds->stpmx = 1.0;
    else
      ds->stpmx = dmin_(kBig, smin);
  }
  if (ds->iter == 0 && !ds->boxed)
    ds->stp = dmin_(1.0 / ds->dnrm, ds->stpmx);
  else
    ds->stp = 1.0;
  ds->fold = ds->f;
  ds->ifun = 0;
  ds->iback = 0;
  ds->csave = 0;
  ds->gd = r[P5_GD];
  ds->gdold = ds->gd;
  if (ds->gd >= 0) {
    ds->internalinfo = -4;
    return ls_failure(ds);
  }
  d_dcsrch(ds->f, ds->gd, ds->stp, kFtol, kGtol, kXtol, 0.0, ds->stpmx, ds->csave, ds->isave,
           ds->dsave);
  ds->xstep = ds->stp * ds->dnrm;
  if (ds->csave != 4 && ds->csave != 3) {
    request_trial(ds);
    ds->status = ST_NEED_EVAL;
    ds->halt = 1;
  } else {
    ds->task = 5;
    accept_step(ds);
  }
}

// Later calls of lnsrlb, after an evaluation with value f
// (lbfgsbcpu.cpp:1312-1339 and the loop exit at :352).
__global__ void s_ls_step(DevScalars* ds, const double* part, int G, double f) {
  __shared__ double r[1];
  reduce_rows(part, G, 0, 1, r);
  if (threadIdx.x != 0) return;
  ds->halt = 0;
  ds->status = ST_RUNNING;
  ds->f = f;
  ds->gd = r[0];
  ds->write_trial = 0;
  d_dcsrch(ds->f, ds->gd, ds->stp, kFtol, kGtol, kXtol, 0.0, ds->stpmx, ds->csave, ds->isave,
           ds->dsave);
  ds->xstep = ds->stp * ds->dnrm;
  if (ds->csave != 4 && ds->csave != 3) {
    request_trial(ds);
    if (ds->iback >= 20) {
      // the reference leaves the loop before evaluating this trial point
      accept_step(ds);
    } else {
      ds->status = ST_NEED_EVAL;
      ds->halt = 1;
    }
  } else {
    ds->task = 5;
    accept_step(ds);
  }
}

// ---------------------------------------------------------------------------
void launch_resume(hipStream_t s, DevScalars* ds, int reset_bp) {
  hipLaunchKernelGGL(s_resume, dim3(1), dim3(64), 0, s, ds, reset_bp);
}
void launch_setup_errclb(hipStream_t s, DevScalars* ds, const double* part, int G) {
  hipLaunchKernelGGL(s_setup_errclb, dim3(1), dim3(kBlock), 0, s, ds, part, G);
}
void launch_setup_active(hipStream_t s, DevScalars* ds, const double* part, int G) {
  hipLaunchKernelGGL(s_setup_active, dim3(1), dim3(kBlock), 0, s, ds, part, G);
}
void launch_setup_projgr(hipStream_t s, DevScalars* ds, const double* pmax, int G, double f) {
  hipLaunchKernelGGL(s_setup_projgr, dim3(1), dim3(kBlock), 0, s, ds, pmax, G, f);
}
void launch_stop_tests(hipStream_t s, DevScalars* ds, const double* part, const double* pmax,
                       const double* pmin, const long long* pmini, int G) {
  hipLaunchKernelGGL(s_stop_tests, dim3(1), dim3(kBlock), 0, s, ds, part, pmax, pmin, pmini, G);
}
void launch_update_gcp(hipStream_t s, DevScalars* ds, const double* part, const double* pmin,
                       const long long* pmini, int G, int reduce) {
  hipLaunchKernelGGL(s_update_gcp, dim3(1), dim3(kBlock), 0, s, ds, part, pmin, pmini, G,
                     reduce);
}
void launch_formk_subsm(hipStream_t s, DevScalars* ds, const double* part, int G) {
  hipLaunchKernelGGL(s_formk_subsm, dim3(1), dim3(kBlock), 0, s, ds, part, G);
}
void launch_alpha(hipStream_t s, DevScalars* ds, const double* pmin, const long long* pmini,
                  int G) {
  hipLaunchKernelGGL(s_alpha, dim3(1), dim3(kBlock), 0, s, ds, pmin, pmini, G);
}
void launch_ls_init(hipStream_t s, DevScalars* ds, const double* part, const double* pmin,
                    int G) {
  hipLaunchKernelGGL(s_ls_init, dim3(1), dim3(kBlock), 0, s, ds, part, pmin, G);
}
void launch_ls_step(hipStream_t s, DevScalars* ds, const double* part, int G, double f) {
  hipLaunchKernelGGL(s_ls_step, dim3(1), dim3(kBlock), 0, s, ds, part, G, f);
}

#undef A1

}  // namespace lbfgsb
