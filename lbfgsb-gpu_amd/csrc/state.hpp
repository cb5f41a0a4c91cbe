// Device-resident solver state shared by the host driver and the kernels.
//
// Everything the per-iteration control flow needs lives in one DevScalars
// block in HBM, so the small single-workgroup kernels can make every
// decision of the reference driver (lbfgsbcpu.cpp:279-462) on the device and
// the host only has to read a status word once per function evaluation.
// Small dense algebra and all line-search scalars are FP64 regardless of the
// bulk type.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace lbfgsb {

// Steps of the iteration chain: kernels of the plain engine, block steps of
// the batched solver.
enum Step : int {
  K_RESUME_RESET = 0,
  K_RESUME,
  K_PASSA_SA,   // standalone Cauchy classification pass
  K_UPD_REDUCE, // controller after the standalone pass
  K_PASSA_FUSED,
  K_STOP,
  K_UPD,
  K_P4,
  K_FORMK,
  K_P5A,
  K_ALPHA,
  K_P5B,
  K_LSINIT,
  K_TRIAL,
  K_GD,
  K_LSSTEP,
  K_END,        // also the timing slot of host-sized breakpoint rounds
  K_GCP = 17,   // device-sized breakpoint round (select + one-workgroup round)
  K_PASSA_SPEC = 18,  // pass A on the evaluated trial point before its acceptance
  K_LSSTOP = 19       // line-search step, then on acceptance the stopping tests
};
constexpr int kNumSteps = 20;

// Programs (sequences of steps).
enum Program : int {
  PROG_ITER_START = 0,  // iteration 0 and after every memory reset
  PROG_AFTER_EVAL = 1,  // a function value for the pending trial is known
  PROG_AFTER_ACCEPT = 2,
  PROG_AFTER_CHECK = 3,
  PROG_TAIL = 4
};

// Host-visible status (DevScalars::status).
enum Status : int {
  ST_RUNNING = 0,     // chain continues on the device
  ST_NEED_EVAL = 1,   // a trial point is ready: call funcgrad(xb, f, gb)
  ST_ACCEPTED = 2,    // line search accepted a step (host may run callbacks)
  ST_CHECKED = 3,     // stopping tests passed, custom stop may be asked
  ST_GCP_ROUNDS = 4,  // breakpoint search needs sorted rounds
  ST_RESTART = 5,     // memory reset: the iteration restarts at the Cauchy step
  ST_DONE = 6,        // solve ended, info holds the reason
  ST_REDO = 7         // step accepted, but not as pass A assumed: run it again
};

// Slot layout of per-block partial sums (value-major: part[v * G + block]).
enum PartA : int {
  PA_TF = 0,        // sum (xold - x)^2
  PA_RR = 1,        // sum y^2
  PA_F1 = 2,        // sum over GCP-free variables of g^2
  PA_NBREAK = 3,    // number of breakpoints
  PA_NFREEGCP = 4,  // free variables that are not breakpoints
  PA_UNBND = 5,     // free non-breakpoint with g != 0 (reference `bnded` false)
  PA_SY = 6,        // 8: s . y_j for the retained slots
  PA_SS = 14,       // 8: s_j . s for the retained slots
  PA_P = 22,        // 16: p = [Y'd ; S'd] over GCP-free variables
  PA_GD = 38,       // g'd at the evaluated trial point (pass A ahead of acceptance)
  PA_COUNT = 39
};
enum Part4 : int {
  P4_NFREE = 0,
  P4_NENTER = 1,
  P4_NLEAVE = 2,
  P4_NC = 3,      // 32: new-column inner products (yy, ss, sy_active, sy_free)
  P4_WV = 35,     // 16: [Y'r ; S'r] over free variables
  P4_CORR = 51,   // 136: signed entering/leaving inner products
  P4_COUNT = 187
};
enum Part5 : int {
  P5_DTD = 0,
  P5_GD = 1,
  P5_COUNT = 2
};

struct DevScalars {
  // ---- control / host interface
  int status;
  int halt;          // when set, bulk kernels in the speculative chain no-op
  int phase;
  int info;          // summary info code
  int internalinfo;

  // ---- sizes and ring state (head/itail are 0-based slots)
  long long n;
  int m, col, head, itail, iupdat, iter, nfgv, nint, nintol, nskip;
  int nfree, nenter, ileave, nleave, wrk, updatd, iword, ifun, iback, task, csave;
  int cnstnd, boxed, prjctd;
  int isave[2];
  int max_iteration;
  int restarts;
  int skip_subsm;    // nfree == 0 || col == 0 for this iteration
  int shortcut;      // unconstrained shortcut (no Cauchy search)

  // ---- line-search / driver scalars
  double f, fold, theta, dr, rr, dnrm, xstep, sbgnrm, ddum, dtd, gd, gdold;
  double stp, stpmx, tf;
  double eps_g, eps_f, eps_x;
  double residual_g, residual_f, residual_x;
  double dsave[13];

  // ---- generalized Cauchy point
  int gcp_mode;      // 0 none passed, 1 threshold (passed set), 2 trivial, 3 all-n passed
  int gcp_nbreak;
  int gcp_nfreegcp;
  int gcp_unbnd;
  long long gcp_kpassed;   // breakpoints passed so far
  long long gcp_last_idx;  // (t, idx) of the last passed breakpoint
  double gcp_last_t;
  double gcp_f1, gcp_f2, gcp_f2org, gcp_dtm, gcp_tsum, gcp_bkmin;
  long long gcp_ibkmin;
  double gcp_D0;           // sum of d^2 over GCP-free variables
  double gcp_carry[40];    // running [P(2m) ; Q(2m) ; D] over passed breakpoints
  double gcp_tau_lo, gcp_tau_hi;
  int gcp_round;
  int gcp_floor_hits;      // f2 floor engaged (closed form may deviate)
  long long gcp_near_listed;  // statistics: entries in pass A's lists at the last round they fed
  long long gcp_stop;      // first stop position in the current round (or -1)
  int gcp_last_round;      // the current round covers all remaining breakpoints
  int gcp_cand_in_b;       // sorted candidates ended in the second buffer
  double gcp_t_prev;       // t of the last passed breakpoint (0 before any)
  long long gcp_idx_prev;
  // Breakpoints pass A lists: every variable with t below gcp_near_tau (0:
  // none listed), per block of pass A (plain form) or per problem (batched
  // form); a round whose t-range lies below it takes its candidates from
  // that list instead of classifying every variable again (gcp_sel_near:
  // the current round did)
  double gcp_near_tau;
  double gcp_near_grow;    // near_tau over the t-range the last search ended in
  int gcp_sel_near;
  int gcp_near_misses;     // statistics: rounds the list did not cover
  long long gcp_list_n;    // batched: entries pass A counted into the problem's list (-1: none)

  // ---- small dense algebra (row-major, leading dimension 8 or 16)
  double sy[64], ss[64], wt[64];
  double wn[256], wn1[256];
  double Mx[256];          // explicit middle matrix M (2col x 2col, ld 16)
  double p[16], c[16], v[16], wv[16], mc[16];
  double newcol[32];
  double pending_s_ss_diag;
  // workspace of the small dense routines; the control kernels stage this
  // struct in LDS, so temporaries placed here avoid per-thread scratch memory
  double tmp[80];

  // ---- reduced pass-A values (see PartA)
  double pa[PA_COUNT];

  // ---- subspace step
  double alpha;
  long long ibd;

  // ---- trial point bookkeeping: k_trial writes X[trial_dst] from
  // X[trial_src] when write_trial is set.  The subspace pass leaves the
  // subspace point z in the trial buffer (trial_has_z), which is the trial
  // point of a unit step; the direction d = z - x is stored only by the
  // first trial that overwrites z (d_stored; trial_d: this k_trial does it),
  // and is otherwise the difference of the two x buffers.
  int write_trial, trial_src, trial_dst, trial_has_z, d_stored, trial_d;
  // ---- physical W slots: ring slot s is physical slot s ^ pslot_x[s], the
  // spare is kMaxM ^ pspare_x (zero-initialised state = the identity)
  int pslot_x[kMaxM];
  int pspare_x;
  int pad_ring;
  int has_after_iteration, has_custom_stop;

  // ---- append counters (breakpoint list, round candidates)
  unsigned long long bp_count;
  unsigned long long cand_count;

  int gcp_pending;   // sorted breakpoint rounds are in progress
  int gcp_need_select;  // the current t-range's candidates are not selected yet
  int gcp_ranges;       // statistics: t-ranges opened after the first one
  int upd_done;      // the memory update of this iteration already ran
  int p4_flagged;    // k_p4 saw entering/leaving variables
  int gcp_rounds_run;  // statistics: Cauchy searches that needed sorted rounds
  int gcp_handoffs;    // rounds too large for the one-workgroup round
  int pad_handoffs;
  long long gcp_passed_total;  // statistics: breakpoints passed in those
  // workgroup-wide steps requested by the one-thread decisions: 1 = memory
  // update then the Cauchy controller, 2 = the controller only
  int post_update;
  int blk_flag;  // a workgroup-wide step's outcome, handed between its phases

  // ---- resumable solve of the batched form: set up done, next program of
  // the chain and the function value that program takes
  int b_started;
  int b_next_prog;
  double b_next_f;

  // ---- finished row sums of the last multi-block reduction (k_rowsum)
  double red[P4_COUNT + 1];
  // pass A's projected-gradient norm and smallest breakpoint (value, index),
  // finished by the extra block of k_rowsum
  double red_sbg, red_bk;
  long long red_ibk;
};

// The device state staged into LDS by a single-workgroup kernel (all threads
// call; both end with the workgroup synchronised on the copy).
static_assert(sizeof(DevScalars) % 8 == 0, "DevScalars is copied as 8-byte words");
constexpr int kDsWords = sizeof(DevScalars) / 8;

// Every word a thread copies is loaded before its first LDS store, so the
// loads are in flight together (sized for kBlock threads; other block sizes
// finish in the remainder loop).
__device__ inline void ds_load(DevScalars* sds, const DevScalars* ds) {
  const unsigned long long* src = (const unsigned long long*)ds;
  unsigned long long* dst = (unsigned long long*)sds;
  constexpr int kPer = (kDsWords + kBlock - 1) / kBlock;
  const int nt = blockDim.x;
  unsigned long long v[kPer];
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int i = threadIdx.x + k * nt;
    v[k] = i < kDsWords ? src[i] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int i = threadIdx.x + k * nt;
    if (i < kDsWords) dst[i] = v[k];
  }
  for (int i = threadIdx.x + kPer * nt; i < kDsWords; i += nt) dst[i] = src[i];
  __syncthreads();
}
__device__ inline void ds_store(DevScalars* ds, const DevScalars* sds) {
  __syncthreads();
  const unsigned long long* src = (const unsigned long long*)sds;
  unsigned long long* dst = (unsigned long long*)ds;
  for (int i = threadIdx.x; i < kDsWords; i += blockDim.x) dst[i] = src[i];
}

// Physical W slot of ring slot s, the spare, and the exchange that makes the
// spare (holding a new pair) ring slot s.
__host__ __device__ inline int ring_phys(const DevScalars* ds, int s) {
  return ring_map(s, ds->pslot_x);
}
__host__ __device__ inline int ring_spare(const DevScalars* ds) { return kMaxM ^ ds->pspare_x; }
#if defined(LBFGSB_RING_IDENTITY) && defined(__HIP_DEVICE_COMPILE__)
__device__ void ring_take_spare(DevScalars* ds, int s) = delete;  // the identity stays
#else
__host__ __device__ inline void ring_take_spare(DevScalars* ds, int s) {
  const int old = ring_phys(ds, s);
  ds->pslot_x[s] = ring_spare(ds) ^ s;
  ds->pspare_x = old ^ kMaxM;
}
#endif

}  // namespace lbfgsb
