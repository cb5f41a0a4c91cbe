// Scalar control logic of the reference driver as device functions over a
// DevScalars block (normally staged in LDS).  The single-workgroup control
// kernels of the plain engine and the persistent per-problem kernel of the
// batched solver both run these, so the two forms take the same decisions.
// Thread 0 runs everything except build_middle (threads 0 .. 2col-1).
#pragma once

#include "common.hpp"
#include "dense.hpp"
#include "state.hpp"

namespace lbfgsb {

// a(i,j) of a row-major matrix with leading dimension ld, 1-based
#define A1(a, ld, i, j) (a)[((i)-1) * (ld) + ((j)-1)]

static constexpr double kMachEps = 5e-16;  // culbfgsb/ap.cpp:38, both precisions
static constexpr double kBig = 1.0e10;     // lbfgsbcpu.cpp:1259
static constexpr double kFtol = 1.0e-3, kGtol = 0.9, kXtol = 0.1;

// BFGS memory reset used by every restart of the reference driver
// (lbfgsbcpu.cpp:297-304, 317-324, 333-340, 380-385, 452-459).
__device__ inline void reset_memory(DevScalars* ds) {
  ds->internalinfo = 0;
  ds->col = 0;
  ds->head = 0;
  ds->theta = 1.0;
  ds->iupdat = 0;
  ds->updatd = 0;
  ds->restarts++;
}

__device__ inline void set_done(DevScalars* ds, int info) {
  ds->info = info;
  ds->status = ST_DONE;
  ds->halt = 1;
}

// Cauchy-search controller (lbfgsbcpu.cpp:620-724 and the first step of the
// breakpoint loop :725-747).  Thread 0.  Sets gcp_pending when the sorted
// breakpoint rounds must run.
__device__ __forceinline__ void gcp_controller(DevScalars* ds) {
  ds->gcp_pending = 0;
  ds->p4_flagged = 0;
  for (;;) {
    const int col = ds->col;
    ds->iword = -1;
    ds->shortcut = (!ds->cnstnd && col > 0);
    if (ds->shortcut) {
      ds->wrk = ds->updatd;
      ds->nint = 0;
      ds->gcp_mode = 2;
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
      if (col > 0 && d_bmv(ds->sy, ds->wt, col, ds->c, ds->mc) != 0) {
        reset_memory(ds);
        continue;
      }
      ds->nintol += ds->nint;
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
      if (col > 0 && d_bmv(ds->sy, ds->wt, col, ds->c, ds->mc) != 0) {
        reset_memory(ds);
        continue;
      }
      ds->nintol += ds->nint;
      return;
    }
    // at least the smallest breakpoint is passed: sorted rounds, the first
    // over t in [0, 2 dtm) (dtm >= bkmin > 0 here)
    ds->gcp_pending = 1;
    ds->gcp_need_select = 1;
    ds->gcp_rounds_run++;
    ds->gcp_round = 0;
    ds->gcp_floor_hits = 0;
    for (int q = 0; q < 40; q++) ds->gcp_carry[q] = 0.0;
    ds->gcp_tau_lo = -__builtin_inf();
    const double t0 = 2.0 * dtm;
    ds->gcp_tau_hi = t0 < 1e300 ? t0 : __builtin_inf();
    ds->cand_count = 0;
    return;
  }
}

// Columns M e_k of the middle matrix for the breakpoint rounds; threads
// 0 .. 2col-1 (sy, wt in LDS).
__device__ inline void build_middle(DevScalars* ds) {
  const int col = ds->col;
  // per-thread unit vector and result rows in LDS (not scratch memory)
  __shared__ double s_e[16][17], s_out[16][17];
  if (!ds->gcp_pending) return;
  // Mx is stored in the padded layout of the breakpoint sums: index a of
  // the packed 2col vector [y-part ; s-part] goes to a (a < col) or
  // 8 + a - col; the other entries are zero
  for (int q = threadIdx.x; q < 256; q += blockDim.x) ds->Mx[q] = 0.0;
  __syncthreads();
  if ((int)threadIdx.x < 2 * col) {
    double* e = s_e[threadIdx.x];
    double* out = s_out[threadIdx.x];
    for (int i = 0; i < 16; i++) e[i] = 0.0;
    e[threadIdx.x] = 1.0;
    d_bmv(ds->sy, ds->wt, col, e, out);
    const int k = (int)threadIdx.x;
    const int kp = k < col ? k : 8 + k - col;
    for (int r = 0; r < 2 * col; r++) ds->Mx[(r < col ? r : 8 + r - col) * 16 + kp] = out[r];
  }
}

// matupd's small part (lbfgsbcpu.cpp:1365-1386); formt follows in
// formt_block.  Returns whether a pair was added.
__device__ inline bool memory_update_pre(DevScalars* ds) {
  if (!ds->updatd) return false;
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
  return true;
}

// memory_update_pre by the whole workgroup (at least 64 threads; the caller
// synchronises before, and it returns synchronised).  Each entry (i, j) of
// the shift takes (i + 1, j + 1), which the serial loops read before they
// overwrite it, so every thread reads its source before any writes.
__device__ inline bool memory_update_block(DevScalars* ds) {
  if (!ds->updatd) return false;
  const int m = ds->m, col = ds->col;
  const int t = threadIdx.x;
  if (ds->iupdat > m) {
    const int i = t / 8 + 1, j = t % 8 + 1;
    const bool in = t < 64 && i <= col - 1 && j <= col - 1;
    double vs = 0.0, vy = 0.0;
    if (in) {
      if (i <= j) vs = A1(ds->ss, 8, i + 1, j + 1);
      if (i >= j) vy = A1(ds->sy, 8, i + 1, j + 1);
    }
    __syncthreads();
    if (in) {
      if (i <= j) A1(ds->ss, 8, i, j) = vs;
      if (i >= j) A1(ds->sy, 8, i, j) = vy;
    }
    __syncthreads();
  }
  if (t < col - 1) {
    A1(ds->sy, 8, col, t + 1) = ds->pa[PA_SY + t];
    A1(ds->ss, 8, t + 1, col) = ds->pa[PA_SS + t];
  }
  if (t == 0) {
    ds->rr = ds->pa[PA_RR];
    ds->theta = ds->rr / ds->dr;
    A1(ds->ss, 8, col, col) = ds->stp == 1.0 ? ds->dtd : ds->stp * ds->stp * ds->dtd;
    A1(ds->sy, 8, col, col) = ds->dr;
  }
  __syncthreads();
  return true;
}

// formt (lbfgsbcpu.cpp:1105-1133) by the whole workgroup: the entries of
// T = theta S'S + L D^{-1} L' one per thread (each summed in the reference
// order), then the Cholesky factor by thread 0 in registers.  A failed
// factorisation resets the memory (thread 0).
__device__ inline void formt_block(DevScalars* ds) {
  const int col = ds->col;
  const double theta = ds->theta;
  const int t = threadIdx.x;
  if (t < 64) {
    const int i = t / 8 + 1, j = t % 8 + 1;
    if (i <= j && j <= col) {
      if (i == 1) {
        A1(ds->wt, 8, 1, j) = theta * A1(ds->ss, 8, 1, j);
      } else {
        const int k1 = i - 1;  // min(i, j) - 1
        double ddum = 0.0;
        for (int k = 1; k <= k1; k++)
          ddum = ddum + A1(ds->sy, 8, i, k) * A1(ds->sy, 8, j, k) / A1(ds->sy, 8, k, k);
        A1(ds->wt, 8, i, j) = ddum + theta * A1(ds->ss, 8, i, j);
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    double a[8][8];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 8; c++) a[r][c] = (r <= c && c < col) ? A1(ds->wt, 8, r + 1, c + 1) : 0.0;
    const bool ok = d_dpofa_r<8>(a, col);
    if (ok) {
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 8; c++)
          if (r <= c && c < col) A1(ds->wt, 8, r + 1, c + 1) = a[r][c];
    } else {
      // dpofa leaves the factor partly written; the memory is discarded
      reset_memory(ds);
    }
  }
  __syncthreads();
}

// formk (lbfgsbcpu.cpp:907-1103) from the inner products of k_p4, by the
// whole workgroup.  Every entry of wn1 and wn is formed by one thread with
// the reference's operations in the reference's order; the Cholesky
// factorisations run on thread 0 in registers and the triangular solves of
// the off-diagonal block one column per thread.  Returns false (internalinfo
// set) when a factorisation fails.
__device__ inline bool formk_block(DevScalars* ds, const double* nc, const double* corr) {
  const int m = ds->m, col = ds->col;
  double* wn1 = ds->wn1;
  double* wn = ds->wn;
  const double theta = ds->theta;
  const int t = threadIdx.x;
  const bool upd = ds->updatd != 0;
  const int upcl = upd ? col - 1 : col;
  if (upd && ds->iupdat > m) {
    // the ring dropped its oldest pair: the three lower blocks of wn1 move
    // one step up the diagonal (every destination takes its source's value
    // from before the move, as the reference's loop order gives)
    double v = 0.0;
    bool dest = false;
    int r = 0, c = 0;
    if (t < 256) {
      r = t / 16 + 1;
      c = t % 16 + 1;
      const bool yy = c <= m - 1 && r >= c && r <= m - 1;
      const bool ssb = c >= m + 1 && c <= 2 * m - 1 && r >= c && r <= 2 * m - 1;
      const bool sy = c <= m - 1 && r >= m + 1 && r <= 2 * m - 1;
      dest = yy || ssb || sy;
      if (dest) v = A1(wn1, 16, r + 1, c + 1);
    }
    __syncthreads();
    if (dest) A1(wn1, 16, r, c) = v;
    __syncthreads();
  }
  if (upd) {
    if (t < col) {
      const int jy = t + 1;
      A1(wn1, 16, col, jy) = nc[jy - 1];
      A1(wn1, 16, m + col, m + jy) = nc[8 + jy - 1];
      A1(wn1, 16, m + col, jy) = nc[16 + jy - 1];
    }
    __syncthreads();
    if (t < col) A1(wn1, 16, m + t + 1, col) = nc[24 + t];
    __syncthreads();
  }
  if (corr) {
    if (t < 64) {
      const int iy = t / 8 + 1, jy = t % 8 + 1;
      if (iy <= upcl && jy <= iy) {
        const int q = (iy - 1) * iy / 2 + (jy - 1);
        A1(wn1, 16, iy, jy) = A1(wn1, 16, iy, jy) + corr[q];
        A1(wn1, 16, m + iy, m + jy) = A1(wn1, 16, m + iy, m + jy) - corr[36 + q];
      }
      if (iy <= upcl && jy <= upcl) {
        const double v = corr[72 + (iy - 1) * 8 + (jy - 1)];
        if (iy <= jy)
          A1(wn1, 16, m + iy, jy) = A1(wn1, 16, m + iy, jy) + v;
        else
          A1(wn1, 16, m + iy, jy) = A1(wn1, 16, m + iy, jy) - v;
      }
    }
    __syncthreads();
  }
  // wn from wn1, entry by entry
  if (t < 64) {
    const int iy = t / 8 + 1, jy = t % 8 + 1;
    if (iy <= col && jy <= col) {
      const int iis = col + iy, is1 = m + iy;
      if (jy <= iy) {
        double v = A1(wn1, 16, iy, jy) / theta;
        if (jy == iy) v = v + A1(ds->sy, 8, iy, iy);
        A1(wn, 16, jy, iy) = v;
        A1(wn, 16, col + jy, iis) = A1(wn1, 16, is1, m + jy) * theta;
      }
      A1(wn, 16, jy, iis) = jy < iy ? -A1(wn1, 16, is1, jy) : A1(wn1, 16, is1, jy);
    }
  }
  __syncthreads();
  // LEL': the upper-left block's factor
  if (t == 0) {
    double a[8][8];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 8; c++) a[r][c] = (r <= c && c < col) ? A1(wn, 16, r + 1, c + 1) : 0.0;
    const bool ok = d_dpofa_r<8>(a, col);
    if (ok) {
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 8; c++)
          if (r <= c && c < col) A1(wn, 16, r + 1, c + 1) = a[r][c];
    } else {
      ds->internalinfo = -1;
    }
    ds->blk_flag = ok ? 1 : 0;
  }
  __syncthreads();
  if (!ds->blk_flag) return false;
  // L^{-1} (-L_a'+R_z') one column per thread
  if (t < col) {
    const int js = col + 1 + t;
    double b[8];
#pragma unroll
    for (int i = 0; i < 8; i++) b[i] = i < col ? A1(wn, 16, i + 1, js) : 0.0;
    d_dtrsl_r<8>(wn, 16, col, b, 11);
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (i < col) A1(wn, 16, i + 1, js) = b[i];
  }
  __syncthreads();
  if (t < 64) {
    const int a = t / 8, b = t % 8;
    if (a <= b && b < col) {
      const int iis = col + 1 + a, js = col + 1 + b;
      A1(wn, 16, iis, js) = A1(wn, 16, iis, js) +
                            ddot(&A1(wn, 16, 1, iis), 16, &A1(wn, 16, 1, js), 16, col);
    }
  }
  __syncthreads();
  // the lower-right block's factor
  if (t == 0) {
    double a[8][8];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 8; c++)
        a[r][c] = (r <= c && c < col) ? A1(wn, 16, col + r + 1, col + c + 1) : 0.0;
    const bool ok = d_dpofa_r<8>(a, col);
    if (ok) {
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 8; c++)
          if (r <= c && c < col) A1(wn, 16, col + r + 1, col + c + 1) = a[r][c];
    } else {
      ds->internalinfo = -2;
    }
    ds->blk_flag = ok ? 1 : 0;
  }
  __syncthreads();
  return ds->blk_flag != 0;
}

// The trial point x = z for a unit step, x = stp d + x otherwise
// (lbfgsbcpu.cpp:1329-1335).  The subspace pass has left z in the trial
// buffer: a unit step while it is still there needs no write.
__device__ inline void request_trial(DevScalars* ds) {
  ds->task = 1;
  ds->ifun = ds->ifun + 1;
  ds->nfgv = ds->nfgv + 1;
  ds->iback = ds->ifun - 1;
  ds->trial_src = ds->phase & 1;
  ds->trial_dst = (ds->phase & 1) ^ 1;
  if (ds->stp == 1.0 && ds->trial_has_z) {
    ds->write_trial = 0;
  } else {
    ds->write_trial = 1;
    ds->trial_has_z = 0;  // k_trial overwrites it
    // the first trial that leaves z stores d = z - x before it does
    ds->trial_d = !ds->d_stored;
    ds->d_stored = 1;
  }
}

// Step accepted: the reference increments iter and, after its stopping
// tests, decides the update (lbfgsbcpu.cpp:388-449).  The decision only
// depends on line-search scalars, so it is taken here and the ring slot is
// chosen before the fused tail pass writes it.
__device__ inline void accept_step(DevScalars* ds) {
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
__device__ inline void ls_failure(DevScalars* ds) {
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


// ---------------------------------------------------------------------------
// Per-step logic, thread 0, on reduced values.  Each control kernel and the
// batched kernel call these after finishing their reductions.

// errclb outcome (lbfgsbcpu.cpp:871-905): counts of bad nbd and of l > u.
__device__ inline void logic_setup_errclb(DevScalars* ds, double bad_nbd, double bad_box) {
  if (bad_nbd > 0 || bad_box > 0) set_done(ds, -1);
}

// active's flags (lbfgsbcpu.cpp:467-515): projected, constrained, not-boxed
// counts.
__device__ inline void logic_setup_active(DevScalars* ds, const double (&r)[3]) {
  ds->prjctd = r[0] > 0;
  ds->cnstnd = r[1] > 0;
  ds->boxed = !(r[2] > 0);
}

// First projected-gradient test after the initial evaluation
// (lbfgsbcpu.cpp:271-278).
__device__ inline void logic_setup_projgr(DevScalars* ds, double sbgnrm, double f) {
  ds->f = f;
  ds->nfgv = 1;
  ds->sbgnrm = sbgnrm;
  ds->residual_g = sbgnrm;
  if (sbgnrm <= ds->eps_g) set_done(ds, 4);
}

__device__ inline void logic_resume(DevScalars* ds, int reset_bp) {
  ds->halt = 0;
  ds->status = ST_RUNNING;
  ds->write_trial = 0;
  if (reset_bp) ds->bp_count = 0;
}

// Stopping tests after an accepted step (lbfgsbcpu.cpp:394-429) on the
// pass-A sums `pa`, the projected-gradient norm and the smallest breakpoint;
// without a customised stopping test it continues with the memory update
// and the next Cauchy controller.
__device__ inline void logic_stop_tests(DevScalars* s, const double* pa, double sbg, double bk,
                                        long long ibk) {
  s->write_trial = 0;
  s->upd_done = 0;
  for (int v = 0; v < PA_COUNT; v++) s->pa[v] = pa[v];
  s->gcp_bkmin = bk;
  s->gcp_ibkmin = ibk;
  s->sbgnrm = sbg;
  s->residual_g = sbg;
  const double tf = sqrt(pa[PA_TF]);
  const double ddum = dmax_(fabs(s->fold), dmax_(fabs(s->f), 1.0));
  if (sbg <= s->eps_g) {
    set_done(s, 4);
    return;
  }
  s->tf = tf;
  s->residual_x = tf;
  if (tf <= s->eps_x) {
    set_done(s, 2);
    return;
  }
  s->ddum = ddum;
  s->residual_f = (s->fold - s->f) / ddum;
  if (s->fold - s->f <= s->eps_f * ddum) {
    set_done(s, 1);
  } else if (s->iter > s->max_iteration && s->max_iteration > 0) {
    set_done(s, 5);
  } else if (s->has_custom_stop) {
    s->status = ST_CHECKED;
    s->halt = 1;
  } else {
    s->post_update = 1;  // memory update + Cauchy controller (post_update_block)
  }
}

// Memory update + Cauchy controller.  reduce: iteration start / restart,
// where the classification sums `pa` and the smallest breakpoint come from
// a standalone pass A (no update then).
__device__ inline void logic_update_gcp(DevScalars* s, int reduce, const double* pa, double bk,
                                        long long ibk) {
  s->write_trial = 0;
  if (reduce) {
    for (int v = 0; v < PA_COUNT; v++) s->pa[v] = pa[v];
    s->gcp_bkmin = bk;
    s->gcp_ibkmin = ibk;
    s->post_update = 2;  // the Cauchy controller only
  } else {
    s->post_update = 1;  // memory update + Cauchy controller
  }
}

// The step a one-thread decision requested in post_update, by the whole
// workgroup (the stopping tests and logic_update_gcp end with it; the
// caller synchronises before): the memory update, formt, and the Cauchy
// controller.
__device__ inline void post_update_block(DevScalars* s) {
  const int mode = s->post_update;
  if (!mode) return;
  __syncthreads();
  if (mode == 1) {
    if (threadIdx.x == 0) s->blk_flag = memory_update_pre(s) ? 1 : 0;
    __syncthreads();
    if (s->blk_flag) formt_block(s);
  }
  if (threadIdx.x == 0) {
    gcp_controller(s);
    s->upd_done = 1;
    s->post_update = 0;
  }
  __syncthreads();
}

// post_update_block followed by build_middle, with the columns M e_k formed
// by the second wave while thread 0 runs the Cauchy controller (both only
// read sy and wt; the controller changes col only when it discards the
// memory, and then the columns are formed again for the new col).  The same
// results as the two calls in sequence; needs at least 64 + 16 threads.
__device__ __forceinline__ void post_update_middle(DevScalars* s) {
  const int mode = s->post_update;
  if (!mode) {
    build_middle(s);
    return;
  }
  __shared__ double s_e[16][17], s_out[16][17];
  __syncthreads();
  if (mode == 1 && memory_update_block(s)) formt_block(s);
  const int col = s->col;
  const int k = (int)threadIdx.x - kWave;
  if (threadIdx.x == 0) {
    gcp_controller(s);
    s->upd_done = 1;
    s->post_update = 0;
  } else if (k >= 0 && k < 2 * col) {
    double* e = s_e[k];
    for (int i = 0; i < 16; i++) e[i] = 0.0;
    e[k] = 1.0;
    d_bmv(s->sy, s->wt, col, e, s_out[k]);
  }
  __syncthreads();
  if (!s->gcp_pending) return;
  if (s->col != col) {
    build_middle(s);
    return;
  }
  for (int q = threadIdx.x; q < 256; q += blockDim.x) s->Mx[q] = 0.0;
  __syncthreads();
  if (k >= 0 && k < 2 * col) {
    // the padded layout of build_middle
    const int kp = k < col ? k : 8 + k - col;
    for (int r = 0; r < 2 * col; r++) s->Mx[(r < col ? r : 8 + r - col) * 16 + kp] = s_out[k][r];
  }
}

// After the Cauchy-point pass, by the whole workgroup: free/active
// bookkeeping (lbfgsbcpu.cpp:1143-1161) on thread 0, formk, and the two
// triangular solves of subsm (lbfgsbcpu.cpp:1459-1471) on thread 0 with the
// vector in registers.  r = the pass's sums (P4_* layout).  The caller
// synchronises before.
__device__ LBFGSB_COLD void logic_formk_subsm(DevScalars* s, const double* r, bool flagged) {
  const int col = s->col;
  if (threadIdx.x == 0) {
    const long long n = s->n;
    if (!s->shortcut) {
      s->nfree = (int)r[P4_NFREE];
      if (s->iter > 0 && s->cnstnd) {
        s->nenter = (int)r[P4_NENTER];
        s->nleave = (int)r[P4_NLEAVE];
      } else {
        s->nenter = 0;
        s->nleave = 0;
      }
      s->ileave = (int)(n + 1 - s->nleave);
      s->wrk = s->nleave > 0 || s->nenter > 0 || s->updatd;
    }
    s->skip_subsm = (s->nfree == 0 || col == 0);
  }
  __syncthreads();
  if (s->skip_subsm) return;
  bool fail = false;
  if (s->wrk) fail = !formk_block(s, r + P4_NC, flagged ? r + P4_CORR : nullptr);
  if (threadIdx.x == 0) {
    if (!fail) {
      const double theta = s->theta;
      double b[16];
#pragma unroll
      for (int q = 0; q < 16; q++)
        b[q] = q < col ? r[P4_WV + q] : (q < 2 * col ? theta * r[P4_WV + 8 + q - col] : 0.0);
      int info = d_dtrsl_r<16>(s->wn, 16, 2 * col, b, 11);
      if (info == 0) {
#pragma unroll
        for (int q = 0; q < 16; q++)
          if (q < col) b[q] = -b[q];
        info = d_dtrsl_r<16>(s->wn, 16, 2 * col, b, 1);
      }
#pragma unroll
      for (int q = 0; q < 16; q++)
        if (q < 2 * col) s->wv[q] = b[q];
      fail = info != 0;
    }
    if (fail) {
      reset_memory(s);
      s->status = ST_RESTART;
      s->halt = 1;
    }
  }
  __syncthreads();
}

// Backtrack length of subsm (lbfgsbcpu.cpp:1484-1538): the smallest
// step-to-box ratio a and its variable.
__device__ inline void logic_alpha(DevScalars* ds, double a, long long i) {
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

// First call of lnsrlb for an iteration (lbfgsbcpu.cpp:1263-1339) on d'd,
// g'd and the smallest step to the box.
__device__ inline void logic_ls_init(DevScalars* s, double dtd, double gd, double smin) {
  s->dtd = dtd;
  s->dnrm = sqrt(s->dtd);
  s->stpmx = kBig;
  if (s->cnstnd) {
    if (s->iter == 0)
      s->stpmx = 1.0;
    else
      s->stpmx = dmin_(kBig, smin);
  }
  if (s->iter == 0 && !s->boxed)
    s->stp = dmin_(1.0 / s->dnrm, s->stpmx);
  else
    s->stp = 1.0;
  s->fold = s->f;
  s->ifun = 0;
  s->iback = 0;
  s->csave = 0;
  s->trial_has_z = 1;  // the subspace pass wrote z into the trial buffer
  s->d_stored = 0;     // and did not store d
  s->gd = gd;
  s->gdold = s->gd;
  if (s->gd >= 0) {
    s->internalinfo = -4;
    ls_failure(s);
    return;
  }
  d_dcsrch(s->f, s->gd, s->stp, kFtol, kGtol, kXtol, 0.0, s->stpmx, s->csave, s->isave,
           s->dsave);
  s->xstep = s->stp * s->dnrm;
  if (s->csave != 4 && s->csave != 3) {
    request_trial(s);
    s->status = ST_NEED_EVAL;
    s->halt = 1;
  } else {
    s->task = 5;
    accept_step(s);
  }
}

// Later calls of lnsrlb, after an evaluation with value f and g'd
// (lbfgsbcpu.cpp:1312-1339 and the loop exit at :352).
__device__ LBFGSB_COLD void logic_ls_step(DevScalars* s, double f, double gd) {
  s->halt = 0;
  s->status = ST_RUNNING;
  s->f = f;
  s->gd = gd;
  s->write_trial = 0;
  d_dcsrch(s->f, s->gd, s->stp, kFtol, kGtol, kXtol, 0.0, s->stpmx, s->csave, s->isave,
           s->dsave);
  s->xstep = s->stp * s->dnrm;
  if (s->csave != 4 && s->csave != 3) {
    request_trial(s);
    if (s->iback >= 20) {
      // the reference leaves the loop before evaluating this trial point
      accept_step(s);
    } else {
      s->status = ST_NEED_EVAL;
      s->halt = 1;
    }
  } else {
    s->task = 5;
    accept_step(s);
  }
}

}  // namespace lbfgsb
