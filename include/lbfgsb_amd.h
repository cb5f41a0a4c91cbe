/*
 * C ABI of the MI355X L-BFGS-B engine (liblbfgsb_amd.so).
 *
 * Plain pointers and sizes only.  "Device" pointers are HIP device memory
 * (hipMalloc); `stream` is a hipStream_t passed as void* (NULL = the null
 * stream).  The C++ drop-in API in culbfgsb.h is a header-only layer over
 * these entry points.
 *
 * Precision: the _f64 entry points compute all n-length work in FP64; the
 * _f32 entry points keep the n-length vectors and the S/Y memory in FP32
 * with FP64-accumulated (compensated) reductions, while the 2m x 2m matrix
 * algebra, the Cauchy-search scalars and the line search stay in FP64.
 */
#ifndef LBFGSB_AMD_H_
#define LBFGSB_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Modes, as LBFGSB_CUDA_MODE (culbfgsb/culbfgsb.h:31).  Only the GPU mode
 * is provided by this library; the CPU mode returns info = -1. */
enum { LBFGSB_AMD_MODE_CPU = 0, LBFGSB_AMD_MODE_GPU = 1 };

/* LBFGSB_CUDA_OPTION<real> (culbfgsb/culbfgsb.h:33-44), same field order.
 * machine_epsilon, machine_maximum and step_scaling are accepted and
 * ignored: the solver follows the reference's CPU path, whose update-skip
 * test, formt and formk use the fixed machine epsilon 5e-16 for both
 * precisions (culbfgsb/lbfgsbcpu.cpp:443, culbfgsb/ap.cpp:38), not the
 * option the reference's CUDA path reads (culbfgsb/lbfgsbcuda.cpp:352). */
typedef struct {
  double machine_epsilon, machine_maximum, step_scaling, eps_g, eps_f, eps_x;
  int hessian_approximate_dimension, max_iteration, mode;
} lbfgsb_options_f64;
typedef struct {
  float machine_epsilon, machine_maximum, step_scaling, eps_g, eps_f, eps_x;
  int hessian_approximate_dimension, max_iteration, mode;
} lbfgsb_options_f32;

/* LBFGSB_CUDA_SUMMARY<real> (culbfgsb/culbfgsb.h:46-53).
 * info: -2 internal error, -1 bad input, 0 stopped by the user callback,
 *        1 eps_f, 2 eps_x, 4 eps_g, 5 max_iteration. */
typedef struct {
  double residual_g, residual_f, residual_x;
  int num_iteration, info;
} lbfgsb_summary_f64;
typedef struct {
  float residual_g, residual_f, residual_x;
  int num_iteration, info;
} lbfgsb_summary_f32;

/* Callback of LBFGSB_CUDA_STATE (culbfgsb/culbfgsb.h:55-67): x and g are
 * device pointers, f is a host scalar the callback writes; stream is the
 * solve's hipStream_t.  funcgrad must have finished its device work when it
 * returns.  The return value of funcgrad and after_iteration is ignored; a
 * non-zero return of customized_stopping stops with info = 0. */
typedef int (*lbfgsb_callback_f64)(void* ctx, double* x, double* f, double* g, void* stream,
                                   const lbfgsb_summary_f64* summary);
typedef int (*lbfgsb_callback_f32)(void* ctx, float* x, float* f, float* g, void* stream,
                                   const lbfgsb_summary_f32* summary);

typedef struct {
  lbfgsb_callback_f64 funcgrad;
  void* funcgrad_ctx;
  lbfgsb_callback_f64 after_iteration;
  void* after_iteration_ctx;
  lbfgsb_callback_f64 customized_stopping;
  void* customized_stopping_ctx;
} lbfgsb_callbacks_f64;
typedef struct {
  lbfgsb_callback_f32 funcgrad;
  void* funcgrad_ctx;
  lbfgsb_callback_f32 after_iteration;
  void* after_iteration_ctx;
  lbfgsb_callback_f32 customized_stopping;
  void* customized_stopping_ctx;
} lbfgsb_callbacks_f32;

/* lbfgsbdefaultoption<double|float> (culbfgsb/lbfgsb.cpp:44-66). */
void lbfgsb_default_options_f64(lbfgsb_options_f64* opt);
void lbfgsb_default_options_f32(lbfgsb_options_f32* opt);

/* lbfgsbminimize<double|float> (culbfgsb/culbfgsb.h:157-163,
 * lbfgsb.cpp:24-42).  x (in/out), nbd, l, u are device arrays of length n.
 * Blocking; returns summary->info. */
int lbfgsb_minimize_f64(int n, const lbfgsb_callbacks_f64* cb, const lbfgsb_options_f64* opt,
                        double* x, const int* nbd, const double* l, const double* u,
                        lbfgsb_summary_f64* summary, void* stream);
int lbfgsb_minimize_f32(int n, const lbfgsb_callbacks_f32* cb, const lbfgsb_options_f32* opt,
                        float* x, const int* nbd, const float* l, const float* u,
                        lbfgsb_summary_f32* summary, void* stream);

/* lbfgsb_minimize_{f64,f32} without any allocation (no counterpart in the
 * reference, which allocates and frees its workspace on every call,
 * culbfgsb/lbfgsbcuda.cpp:127-179, 399-440): for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same arguments, results and info codes as
 * lbfgsb_minimize; the workspace is a fixed 64 MiB region of device memory
 * per GPU, held until the process ends, which admits fp64 solves up to
 * 291,584 variables and fp32 up to 485,888 (m = 8).  A larger n returns -3 (info
 * -3) with a message, leaving x untouched.  Every call on one GPU shares that
 * GPU's region: calls that overlap on one GPU race. */
int lbfgsb_minimize_noalloc_f64(int n, const lbfgsb_callbacks_f64* cb,
                                const lbfgsb_options_f64* opt, double* x, const int* nbd,
                                const double* l, const double* u, lbfgsb_summary_f64* summary,
                                void* stream);
int lbfgsb_minimize_noalloc_f32(int n, const lbfgsb_callbacks_f32* cb,
                                const lbfgsb_options_f32* opt, float* x, const int* nbd,
                                const float* l, const float* u, lbfgsb_summary_f32* summary,
                                void* stream);

/* ---- Stepping engine: the per-step wrappers of the reference GPU driver
 * (culbfgsb/lbfgsbcuda.cpp:443-770: lbfgsbactive, projgr, cauchy, freev,
 * formk, cmprlb, subsm, lnsrlb, matupd, formt) become programs of fused
 * kernels over a device-resident state.  dtype: 0 = fp32 bulk, 1 = fp64.
 * Used by the parity tests, the batched form and the serving loop. */
typedef struct lbfgsb_engine lbfgsb_engine;

lbfgsb_engine* lbfgsb_engine_create(int dtype, long long n, int m, void* x, const int* nbd,
                                    const void* l, const void* u, void* stream);
void lbfgsb_engine_destroy(lbfgsb_engine* e);
/* errclb + active projection (lbfgsbcpu.cpp:258-266); returns the status. */
int lbfgsb_engine_setup_begin(lbfgsb_engine* e, double eps_g, double eps_f, double eps_x,
                              int max_iteration);
/* projgr after the first evaluation at x (lbfgsbcpu.cpp:271-278). */
int lbfgsb_engine_setup_finish(lbfgsb_engine* e, double f);
/* Enqueue program `program` from step `start_step` (-1 = first) up to, not
 * including, `stop_step` (16 = to the end), then read the status.
 * Programs: 0 iteration start, 1 after an evaluation (f = its value),
 * 2 after an accepted step, 3 after the stopping check, 4 Cauchy-point tail.
 * Statuses: 0 running, 1 evaluate the trial point (x/g which = 1), 2 step
 * accepted, 3 stopping tests passed, 4 breakpoint rounds, 5 restart (run
 * program 0), 6 done, 7 step accepted but the tail pass the whole program 1
 * ran ahead of the line search does not apply (run program 2). */
int lbfgsb_engine_run(lbfgsb_engine* e, int program, int start_step, int stop_step, double f);
/* Device pointers of the accepted (which = 0) or pending trial (1) x / g. */
void* lbfgsb_engine_x(lbfgsb_engine* e, int which);
void* lbfgsb_engine_g(lbfgsb_engine* e, int which);
/* Named scalar of the device state (see DevScalars), as double; also
 * "max_sorted_round", the most breakpoint candidates one host-sized round
 * of the Cauchy search sorted so far (a statistic of the engine). */
int lbfgsb_engine_get_scalar(lbfgsb_engine* e, const char* name, double* value);
int lbfgsb_engine_set_scalar(lbfgsb_engine* e, const char* name, double value);
/* Named array copied to / from host memory.  Names: x, g, xb, gb, dsub
 * (bulk type, n); xcp, z, d (bulk type, n; read only: the Cauchy point,
 * the subspace point and the line-search direction computed from the
 * current state, which the iteration does not keep as such);
 * iwhere (int32, n); iwbits (uint8, n: bits 0-3 iwhere + 4, bit 4 free in the
 * last partition, bit 5 free in the one before, bits 6-7 nbd); ws, wy (bulk type, n x m
 * row-major); index, indx2 (int32, n, 1-based as the reference); sy, ss,
 * wt (m x m), wn, snd (2m x 2m), p, c, mc, wv, pa (double); dsave (13) and
 * isave (2, as double): the line search's saved state. */
int lbfgsb_engine_get_array(lbfgsb_engine* e, const char* name, void* out, long long count);
int lbfgsb_engine_set_array(lbfgsb_engine* e, const char* name, const void* in,
                            long long count);
/* Leave the accepted iterate in the caller's x buffer. */
int lbfgsb_engine_finalize(lbfgsb_engine* e);
/* A whole solve on an existing engine: the host loop of lbfgsb_minimize
 * (lbfgsbcpu.cpp:117-465 control flow).  iter_clock (may be NULL) receives,
 * for k < clock_cap, the host time in seconds at which the acceptance of
 * iteration k+1 was observed. */
int lbfgsb_engine_solve_f64(lbfgsb_engine* e, const lbfgsb_callbacks_f64* cb,
                            const lbfgsb_options_f64* opt, lbfgsb_summary_f64* summary,
                            double* iter_clock, int clock_cap);
int lbfgsb_engine_solve_f32(lbfgsb_engine* e, const lbfgsb_callbacks_f32* cb,
                            const lbfgsb_options_f32* opt, lbfgsb_summary_f32* summary,
                            double* iter_clock, int clock_cap);
/* Per-step kernel timing with HIP events on the engine's stream (no
 * counterpart in the reference; used by the benchmark).  Launches are timed
 * once the iteration count reaches from_iter (-1 = off; counters reset).
 * step 16 = the sorted breakpoint rounds of the Cauchy search. */
void lbfgsb_engine_set_profiling(lbfgsb_engine* e, int from_iter);
/* The same for the steps whose bit (1 << step) is set in step_mask only.
 * Each timed step costs two event records on the stream (a few us each),
 * so timing every step slows the iteration; timing one step does not. */
void lbfgsb_engine_set_profiling_steps(lbfgsb_engine* e, int from_iter, unsigned step_mask);
int lbfgsb_engine_step_time(lbfgsb_engine* e, int step, double* total_ms, long long* count);

/* ---- Built-in objectives (the callback side; no counterpart in the
 * reference library -- they restate its example callback shape,
 * examples/dsscfg/dsscfg_cuda.cu:43-154, for the benchmark workloads).
 * kind: 1 grid quadratic, 2 grid cos-energy, 3 extended Rosenbrock,
 * 4 pentadiagonal QP, 6 the reference example's MINPACK-2 dsscfg energy
 * (examples/dsscfg/dsscfg_cpu.cpp:17-150) on an nx x ny grid with h = lambda
 * and no arrays.  Arrays are device arrays of the bulk type.  Returns NULL
 * on a bad kind or shape (grids need nx*ny = n, Rosenbrock an even n). */
typedef struct lbfgsb_problem lbfgsb_problem;
lbfgsb_problem* lbfgsb_problem_create(int dtype, int kind, long long n, long long nx,
                                      long long ny, double h, const void* b, const void* c,
                                      const void* a0, const void* a1, const void* a2);
void lbfgsb_problem_destroy(lbfgsb_problem* p);
/* Usable directly as funcgrad (ctx = the problem).  Passed by this address,
 * lbfgsb_minimize_* and lbfgsb_engine_solve_* read each trial's value from
 * the problem's device double instead of waiting for it on the host (the
 * same value, rounded to the solve's precision). */
int lbfgsb_problem_funcgrad_f64(void* ctx, double* x, double* f, double* g, void* stream,
                                const lbfgsb_summary_f64* summary);
int lbfgsb_problem_funcgrad_f32(void* ctx, float* x, float* f, float* g, void* stream,
                                const lbfgsb_summary_f32* summary);
/* Device-side evaluation: writes g and the value to f_dev (double, device). */
int lbfgsb_problem_eval_device(lbfgsb_problem* p, const void* x, void* g, double* f_dev,
                               void* stream);

/* ---- Batched form (no counterpart in the reference, which solves one
 * problem per call).  Many independent problems with built-in objectives in
 * one arena of HBM slots, solved by a persistent kernel in which each
 * workgroup runs whole problems with the iteration chain of the plain form
 * (the same decisions, block-level reductions).  Used for batches of small
 * problems and by the serving loop.
 *   capacity      elements of all admitted problems together (the arena
 *                 holds about 160 B (fp32) / 290 B (fp64) per element)
 *   max_problems  slots
 * _add copies the problem's device arrays (x0, nbd, l, u and the objective
 * data its kind uses: kind 1 b; 2 b, c; 3 none; 4 a0, a1, a2, b) into a free
 * slot and returns the slot, -1 on bad input (the checks of lbfgsbminimize),
 * -3 when no slot or range is free.  _run advances every unfinished slot,
 * largest first, by at most iter_budget iterations (0 = to completion) and
 * returns how many remain unfinished (negative on a device error); threads
 * is 256 or 512 per workgroup, blocks 0 = the device's concurrent capacity
 * (a launch never exceeds what can be resident at once).
 * Results follow the summary conventions (info as in lbfgsb_summary). */
typedef struct lbfgsb_batch lbfgsb_batch;
lbfgsb_batch* lbfgsb_batch_create(int dtype, long long capacity, int max_problems, int m,
                                  void* stream);
void lbfgsb_batch_destroy(lbfgsb_batch* b);
int lbfgsb_batch_add(lbfgsb_batch* b, long long n, const void* x0, const int* nbd, const void* l,
                     const void* u, int kind, long long nx, long long ny, double h, const void* bv,
                     const void* c, const void* a0, const void* a1, const void* a2, double eps_g,
                     double eps_f, double eps_x, int max_iteration);
int lbfgsb_batch_run(lbfgsb_batch* b, int threads, int blocks, int iter_budget);
int lbfgsb_batch_result(lbfgsb_batch* b, int slot, double* f, int* iter, int* info,
                        double* residual_g, int* done);
/* f, iter, info, done for all max_problems slots (host arrays). */
int lbfgsb_batch_results(lbfgsb_batch* b, double* f, int* iter, int* info, int* done);
/* Device wall-clock ticks at which each slot's solve started and finished
 * (0 while unfinished), and the tick rate. */
int lbfgsb_batch_clocks(lbfgsb_batch* b, long long* start, long long* end,
                        double* ticks_per_second);
/* The slot's current iterate into a device array of n elements. */
int lbfgsb_batch_get_x(lbfgsb_batch* b, int slot, void* out);
/* Start slot dst_slot of dst (added, not yet run) from the current iterate
 * of slot src_slot of src (same n; the batches may differ in precision, the
 * values are converted).  The serving loop's FP64 polish of fp32 answers.
 * Returns 0, -1 on bad slots, -2 on a device error. */
int lbfgsb_batch_warm_start(lbfgsb_batch* dst, int dst_slot, lbfgsb_batch* src, int src_slot);
int lbfgsb_batch_remove(lbfgsb_batch* b, int slot);
long long lbfgsb_batch_free_elems(lbfgsb_batch* b);
/* Device bytes an arena of this dtype, capacity and slot count holds. */
long long lbfgsb_batch_bytes(int dtype, long long capacity, int max_problems);
/* Gangs: with segment > 0 (at least 4096 elements) a problem of n elements
 * is solved by min(max_gang, ceil(n / segment)) workgroups together, each
 * on one segment of its elements, reductions combined in member order; the
 * result depends on n and this configuration only.  segment 0 (the
 * default) = one workgroup per problem.  _run returns -4 when a gang is
 * wider than the launch.  Returns 0, or -1 on a bad argument. */
int lbfgsb_batch_configure(lbfgsb_batch* b, long long segment, int max_gang);
/* Workgroups the slot's problem is solved by under the current
 * configuration (-1 for a free slot). */
int lbfgsb_batch_gang(lbfgsb_batch* b, int slot);
/* on != 0: every _run replays a captured HIP graph of its launches (a
 * queue-clearing kernel and the two batch kernels at fixed grids; queue
 * sizes are read on the device), captured again when threads, blocks or
 * iter_budget change.  Results are identical with and without. */
int lbfgsb_batch_use_graph(lbfgsb_batch* b, int on);

/* ---- Generated problems (no counterpart in the reference): the benchmark
 * objectives built on the device from a seed -- kind 1 grid quadratic,
 * 2 grid cos (nx * ny == n), 3 extended Rosenbrock (even n),
 * 4 pentadiagonal box QP -- with the arrays of lbfgsb_amd/problems.py
 * (uniform draws and the arithmetic on them bit-identical; the normal
 * draws through the device's log and cos).
 * _add_generated admits such a problem into a slot (returns as _add);
 * lbfgsb_generate writes one into device arrays of n elements (the
 * objective arrays its kind uses; others may be NULL), returns 0, -1 on a
 * bad spec, -2 on a device error. */
int lbfgsb_batch_add_generated(lbfgsb_batch* b, int kind, long long n, long long nx, long long ny,
                               unsigned long long seed, double eps_g, double eps_f, double eps_x,
                               int max_iteration);
int lbfgsb_generate(int dtype, int kind, long long n, long long nx, long long ny,
                    unsigned long long seed, void* x0, int* nbd, void* l, void* u, void* b,
                    void* c, void* a0, void* a1, void* a2, void* stream);

/* ---- Serving loop (no counterpart in the reference).  A stream of solve
 * requests with arrival times is served in real time: a request is admitted
 * (its problem generated into a slot of the arena of its precision) as soon
 * as it has arrived and the arena has a slot and a range free -- first come,
 * first served per precision -- and every resident problem advances by
 * `quantum` iterations per launch of the batch kernels (graph-replayed when
 * use_graph), so requests that arrive while others run join the next
 * launch.  The two arenas are sized so that all the server holds on the
 * device stays within hbm_budget_bytes (fp32_share of it for fp32).  A
 * problem's result is the one it has alone (the batched form's results do
 * not depend on the other problems).  Times are host clock seconds from the
 * start of the stream; requests are taken in arrival order.
 * FP64 polish (polish = 1): an fp32 request that finishes is admitted again
 * into the fp64 arena, ahead of waiting fp64 requests, with the same
 * problem, the tolerance eps * polish_eps_scale (a scale of 0 means 1) and
 * the iterations its fp32 phase left of the cap, started from its fp32
 * answer (lbfgsb_batch_warm_start); its answer is that FP64 solve's (info,
 * f), so it stops on the FP64 form's own tests.  The polish restarts the
 * memory, and a restarted solve that stops on the same decrease test as a
 * cold FP64 solve can stop above the cold solve's objective; a scale below
 * 1 carries it past that point, within the request's iteration cap.  The
 * fp32 slot is held until the polish is admitted.  A request larger than
 * the fp64 arena keeps its fp32 answer (polished = 0). */
typedef struct {
  int kind;                 /* objective kind, as lbfgsb_generate */
  int dtype;                /* 0 fp32 bulk, 1 fp64 */
  long long n, nx, ny;      /* size (grid kinds: nx * ny == n) */
  unsigned long long seed;  /* problem seed, as lbfgsb_generate */
  double eps;               /* eps_g = eps_f = eps_x */
  int max_iteration;
  int pad;
  double arrival;           /* seconds after the start of the stream */
} lbfgsb_request;

typedef struct {
  double f;
  int iter;                 /* iterations, both phases of a polished answer */
  int info;                 /* summary info; -1 bad request, -3 larger than its arena */
  int gang;                 /* workgroups that solved it (the last phase) */
  int dtype;
  double t_arrival, t_admit, t_done;
  int polished;             /* 1: an fp32 solve finished by an FP64 polish */
  int iter_fp32;            /* iterations of the fp32 phase of a polished answer */
} lbfgsb_response;

typedef struct {
  double hbm_budget_bytes;  /* device memory the two arenas may hold together */
  double fp32_share;        /* fraction of it for the fp32 arena (0..1) */
  int max_slots;            /* problems resident at once per arena */
  int threads, blocks;      /* batch workgroup shape (0 = 256 and the device's capacity) */
  long long segment;        /* gang segment (0 = one workgroup per problem) */
  int max_gang;
  int quantum;              /* iterations per launch (> 0) */
  int use_graph;
  int polish;               /* 1: fp32 answers are finished in FP64 (below) */
  double polish_eps_scale;  /* tolerance of the polish over the request's (0: 1) */
} lbfgsb_serve_config;

typedef struct {
  double seconds;           /* stream start to last completion */
  long long launches;       /* batch launches (one per arena per quantum) */
  double arena_bytes;       /* device memory the arenas hold */
  int max_waiting;          /* longest admission queue seen */
  int max_resident;         /* most problems resident at once */
} lbfgsb_serve_stats;

/* Returns 0, -1 on a bad configuration, -2 on a device error. */
int lbfgsb_serve_stream(const lbfgsb_serve_config* cfg, const lbfgsb_request* reqs, int count,
                        lbfgsb_response* out, lbfgsb_serve_stats* stats, void* stream);

/* Library version string. */
const char* lbfgsb_amd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LBFGSB_AMD_H_ */
