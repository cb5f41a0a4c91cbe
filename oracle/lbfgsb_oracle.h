/*
 * CPU oracle for the L-BFGS-B hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a 0-based C restatement of the reference's CPU path
 * (culbfgsb/lbfgsbcpu.cpp, an ALGLIB translation of L-BFGS-B 2.1) used as
 * the checker for the HIP engine.  Only tests, the smoke check and the CPU
 * baseline leg of the benchmark link it; the product never does.
 *
 * Each function cites the reference lines it follows.  Summation order of
 * every dot product follows ALGLIB's 4-way blocked form
 * (culbfgsb/apvt.h:86-148), and the fixed machine epsilon 5e-16
 * (culbfgsb/ap.cpp:38) is used for both precisions, as in the reference.
 *
 * Ring indexing: the reference keeps a 1-based ring pointer advanced by
 * `p = p % m + 1` (lbfgsbcpu.cpp:673); here slots are 0-based and advance by
 * `p = (p + 1) % m`.  Matrices are row-major: a(i,j) (1-based) is
 * a[(i-1)*ld + (j-1)].
 *
 * Pinning status (see tests/test_oracle_*.py): every component is checked
 * against an independent statement of the mathematics (brute force, closed
 * forms, long-double factorizations, MINPACK-2 invariants); the driver's
 * parts (active's projection and classes, lnsrlb's largest feasible step and
 * first step, the order of the stopping tests and their info codes) against
 * their definitions (tests/test_oracle_pins_driver.py).  The reference
 * itself cannot be compiled in this image (its CPU path includes CUDA and
 * cuBLAS headers through culbfgsb/lbfgsbcuda.h:13 and cutil_inline.h:13), so
 * no function is pinned against reference outputs.
 */
#ifndef LBFGSB_ORACLE_H_
#define LBFGSB_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Objective kinds evaluated by the oracle's built-in callbacks. */
enum orc_objective_kind {
  ORC_OBJ_NONE = 0,
  ORC_OBJ_GRID_QUAD = 1, /* 0.5*sum_edges (dv)^2 - h^2 b.v, zero Dirichlet */
  ORC_OBJ_GRID_COS = 2,  /* 0.5*sum_edges (dv)^2 + h^2 sum(c cos v - b v) */
  ORC_OBJ_ROSEN = 3,     /* extended Rosenbrock over pairs */
  ORC_OBJ_PENTA = 4,     /* 0.5 x'Ax - b'x, A symmetric pentadiagonal */
  ORC_OBJ_USER = 5,      /* caller-supplied function pointer */
  ORC_OBJ_DSSCFG = 6     /* MINPACK-2 steady-state combustion (the
                            reference's example problem); h holds lambda */
};

/* Outcome of the part of an iteration after the line search. */
enum orc_tail_result {
  ORC_TAIL_NEXT = 0,    /* iteration done, the next one follows */
  ORC_TAIL_RESTART = 1, /* memory reset: the iteration starts again */
  ORC_TAIL_END = 2      /* solve ended, info holds the reason */
};

#define ORC_DEFINE_TYPES(REAL, SFX)                                            \
  typedef int (*orc_fg_##SFX)(void* ctx, const REAL* x, REAL* f, REAL* g);     \
  typedef struct {                                                             \
    int kind;                                                                  \
    int nx, ny;                                                                \
    double h;                                                                  \
    const REAL* b;  /* linear term */                                          \
    const REAL* c;  /* cos weights (GRID_COS) */                               \
    const REAL* a0; /* PENTA diagonal */                                       \
    const REAL* a1; /* PENTA first off-diagonal, a1[i] = A(i,i+1) */           \
    const REAL* a2; /* PENTA second off-diagonal, a2[i] = A(i,i+2) */          \
    orc_fg_##SFX user;                                                         \
    void* user_ctx;                                                            \
  } orc_objective_##SFX;                                                       \
  /* stop_at_iter > 0 stands for a customised stopping callback that       \
   * returns non-zero once iter >= stop_at_iter (culbfgsb.h:64-66). */         \
  typedef struct {                                                             \
    double eps_g, eps_f, eps_x;                                                \
    int m;                                                                     \
    int max_iteration;                                                         \
    int stop_at_iter;                                                          \
    int pad_;                                                                  \
  } orc_options_##SFX;                                                         \
  /* Per-iteration record written by orc_iterate (one per outer iteration). */ \
  typedef struct {                                                             \
    double f, sbgnrm, stp, theta, dtd, gd, residual_f, residual_x;             \
    int iter, nint, nfree, nenter, ileave, wrk, iword, iback, nfgv, col;       \
    int updatd, restarts, info, pad;                                           \
  } orc_record_##SFX;                                                          \
  typedef struct {                                                             \
    /* scalars */                                                              \
    int n, m, col, head, itail, iupdat, iter, nfgv, nint, nintol, nskip;       \
    int nfree, nact, ileave, nenter, iword, ifun, iback, info, task, csave;    \
    int internalinfo, prjctd, cnstnd, boxed, updatd, wrk, k_err, restarts;     \
    int isave[2];                                                              \
    REAL f, fold, theta, dr, rr, dnrm, xstep, sbgnrm, ddum, dtd, gd, gdold;   \
    REAL stp, stpmx, tf;                                                       \
    REAL residual_g, residual_f, residual_x;                                   \
    REAL dsave[13];                                                            \
    REAL wa[64];                                                               \
    /* n-vectors */                                                            \
    REAL *x, *g, *l, *u, *xold, *xdiff, *z, *r, *d, *t;                       \
    /* n x m, row-major, slot-indexed */                                       \
    REAL *ws, *wy;                                                             \
    /* m x m and 2m x 2m, row-major */                                         \
    REAL *sy, *ss, *wt, *wn, *snd;                                             \
    int *nbd, *index, *iwhere, *indx2;                                         \
    orc_objective_##SFX obj;                                                   \
    orc_options_##SFX opt;                                                     \
  } orc_state_##SFX;

ORC_DEFINE_TYPES(double, f64)
ORC_DEFINE_TYPES(float, f32)

#define ORC_DECLARE_API(REAL, SFX)                                             \
  orc_state_##SFX* orc_create_##SFX(int n, int m);                             \
  void orc_destroy_##SFX(orc_state_##SFX* s);                                  \
  void orc_copy_##SFX(orc_state_##SFX* dst, const orc_state_##SFX* src);       \
  int orc_evaluate_##SFX(const orc_objective_##SFX* obj, int n, const REAL* x, \
                         REAL* f, REAL* g);                                    \
  int orc_setup_##SFX(orc_state_##SFX* s, const REAL* x0, const int* nbd,     \
                      const REAL* l, const REAL* u,                            \
                      const orc_options_##SFX* opt,                            \
                      const orc_objective_##SFX* obj);                         \
  int orc_iterate_##SFX(orc_state_##SFX* s, orc_record_##SFX* rec);           \
  /* the three parts of orc_iterate, for stepping tests */                     \
  void orc_iter_head_##SFX(orc_state_##SFX* s);                                \
  void orc_iter_search_##SFX(orc_state_##SFX* s);                              \
  int orc_iter_tail_##SFX(orc_state_##SFX* s);                                 \
  int orc_minimize_##SFX(int n, const orc_options_##SFX* opt,                  \
                         const orc_objective_##SFX* obj, REAL* x,              \
                         const int* nbd, const REAL* l, const REAL* u,         \
                         orc_record_##SFX* records, int max_records,           \
                         int* out_iters, REAL* out_f);                         \
  /* components on a state (reference component semantics) */                 \
  void orc_cauchy_##SFX(orc_state_##SFX* s);                                   \
  void orc_freev_##SFX(orc_state_##SFX* s);                                    \
  void orc_formk_##SFX(orc_state_##SFX* s);                                    \
  void orc_cmprlb_##SFX(orc_state_##SFX* s);                                   \
  void orc_subsm_##SFX(orc_state_##SFX* s);                                    \
  void orc_matupd_##SFX(orc_state_##SFX* s);                                   \
  void orc_formt_##SFX(orc_state_##SFX* s);                                    \
  void orc_projgr_##SFX(orc_state_##SFX* s);                                   \
  void orc_active_##SFX(orc_state_##SFX* s);                                   \
  int orc_lnsrlb_step_##SFX(orc_state_##SFX* s);                               \
  /* pure kernels with explicit arguments */                                   \
  int orc_dpofa_##SFX(REAL* a, int lda, int n);                                \
  int orc_dtrsl_##SFX(const REAL* t, int ldt, int n, REAL* b, int job);        \
  int orc_bmv_##SFX(int m, const REAL* sy, const REAL* wt, int col,            \
                    const REAL* v, REAL* p);                                   \
  void orc_hpsolb_##SFX(int n, REAL* t, int* iorder, int iheap);               \
  void orc_dcsrch_##SFX(REAL f, REAL g, REAL* stp, REAL ftol, REAL gtol,       \
                        REAL xtol, REAL stpmin, REAL stpmax, int* task,        \
                        int* isave, REAL* dsave);                              \
  void orc_dcstep_##SFX(REAL* stx, REAL* fx, REAL* dx, REAL* sty, REAL* fy,    \
                        REAL* dy, REAL* stp, REAL fp, REAL dp, int* brackt,    \
                        REAL stpmin, REAL stpmax);                             \
  REAL orc_dot_##SFX(const REAL* a, int sa, const REAL* b, int sb, int n);

ORC_DECLARE_API(double, f64)
ORC_DECLARE_API(float, f32)

#ifdef __cplusplus
}
#endif

#endif /* LBFGSB_ORACLE_H_ */
