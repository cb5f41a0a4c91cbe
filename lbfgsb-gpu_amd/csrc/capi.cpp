// extern "C" boundary of liblbfgsb_amd.so (declared in include/lbfgsb_amd.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <limits>
#include <string>

#include "lbfgsb_amd.h"
#include "common.hpp"
#include "engine.hpp"
#include "objectives.hpp"

using namespace lbfgsb;

struct lbfgsb_engine {
  EngineBase* e;
};

struct lbfgsb_problem {
  int dtype;
  long long n;
  ObjDesc desc;
  double* part;
  double* fdev;
  HostF* fhost;  // fine-grained pinned: the objective publishes f here
  unsigned fseq;
  int G;
};

namespace {

struct ScalarField {
  const char* name;
  size_t off;
  int type;  // 0 int, 1 long long, 2 double
};

#define SF_I(f) {#f, offsetof(DevScalars, f), 0}
#define SF_L(f) {#f, offsetof(DevScalars, f), 1}
#define SF_D(f) {#f, offsetof(DevScalars, f), 2}
const ScalarField kFields[] = {
    SF_I(status), SF_I(halt), SF_I(phase), SF_I(info), SF_I(internalinfo), SF_L(n), SF_I(m),
    SF_I(col), SF_I(head), SF_I(itail), SF_I(iupdat), SF_I(iter), SF_I(nfgv), SF_I(nint),
    SF_I(nintol), SF_I(nskip), SF_I(nfree), SF_I(nenter), SF_I(ileave), SF_I(nleave),
    SF_I(wrk), SF_I(updatd), SF_I(iword), SF_I(ifun), SF_I(iback), SF_I(task), SF_I(csave),
    SF_I(cnstnd), SF_I(boxed), SF_I(prjctd), SF_I(max_iteration), SF_I(restarts),
    SF_I(skip_subsm), SF_I(shortcut), SF_D(f), SF_D(fold), SF_D(theta), SF_D(dr), SF_D(rr),
    SF_D(dnrm), SF_D(xstep), SF_D(sbgnrm), SF_D(ddum), SF_D(dtd), SF_D(gd), SF_D(gdold),
    SF_D(stp), SF_D(stpmx), SF_D(tf), SF_D(eps_g), SF_D(eps_f), SF_D(eps_x),
    SF_D(residual_g), SF_D(residual_f), SF_D(residual_x), SF_I(gcp_mode), SF_I(gcp_nbreak),
    SF_I(gcp_nfreegcp), SF_I(gcp_unbnd), SF_L(gcp_kpassed), SF_L(gcp_last_idx),
    SF_D(gcp_last_t), SF_D(gcp_f1), SF_D(gcp_f2), SF_D(gcp_f2org), SF_D(gcp_dtm),
    SF_D(gcp_tsum), SF_D(gcp_bkmin), SF_L(gcp_ibkmin), SF_D(gcp_D0), SF_I(gcp_round),
    SF_I(gcp_floor_hits), SF_D(alpha), SF_L(ibd), SF_I(write_trial), SF_I(trial_src), SF_I(trial_has_z), SF_I(d_stored), SF_I(trial_d),
    SF_I(trial_dst), SF_I(has_after_iteration), SF_I(has_custom_stop), SF_I(gcp_pending),
    SF_I(upd_done), SF_I(p4_flagged), SF_I(gcp_rounds_run), SF_L(gcp_passed_total),
    SF_L(cand_count), SF_I(gcp_need_select), SF_I(gcp_ranges), SF_D(gcp_near_tau),
    SF_D(gcp_near_grow), SF_I(gcp_sel_near), SF_I(gcp_near_misses), SF_L(gcp_near_listed)};

const ScalarField* find_field(const char* name) {
  for (const auto& f : kFields)
    if (std::strcmp(f.name, name) == 0) return &f;
  return nullptr;
}

template <typename Summary>
void fill_summary(const DevScalars& h, Summary* s) {
  using R = decltype(s->residual_g);
  s->residual_g = (R)h.residual_g;
  s->residual_f = (R)h.residual_f;
  s->residual_x = (R)h.residual_x;
  s->num_iteration = h.iter;
  s->info = h.info;
}

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Launches the library objective at x with no wait: the gradient into g,
// the value into p->fdev (and published to p->fhost under a new sequence
// number).
template <typename R>
int problem_launch(lbfgsb_problem* p, R* x, R* g, hipStream_t st) {
  launch_objective<R>(st, p->desc, p->n, x, g, p->part, p->G, p->fdev, p->fhost, ++p->fseq);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The problem when funcgrad is the library's own objective of this
// precision (lbfgsb_problem_funcgrad_*), else null.
template <typename R, typename Cb>
lbfgsb_problem* own_objective(const Cb* cb) {
  const void* own = sizeof(R) == 8 ? reinterpret_cast<const void*>(&lbfgsb_problem_funcgrad_f64)
                                   : reinterpret_cast<const void*>(&lbfgsb_problem_funcgrad_f32);
  lbfgsb_problem* p = (lbfgsb_problem*)cb->funcgrad_ctx;
  if (reinterpret_cast<const void*>(cb->funcgrad) != own || !p) return nullptr;
  return p->dtype == (sizeof(R) == 8 ? 1 : 0) ? p : nullptr;
}

// Host loop of a plain solve on an existing engine (lbfgsbcpu.cpp:117-465
// control flow; the per-iteration decisions run on the device, see
// small.hip).  iter_clock[k], if given, receives the host time at which the
// acceptance of iteration k+1 was observed.
template <typename R, typename Opt, typename Cb, typename Summary>
int drive(EngineBase* e, const Cb* cb, const Opt* opt, Summary* summary, void* stream,
          double* iter_clock, int clock_cap) {
  DevScalars& h = e->host_scalars();
  if (e->read_scalars()) return -2;
  const long long n = h.n;
  const int m = h.m;
  std::memset(&h, 0, sizeof(DevScalars));
  h.n = n;
  h.m = m;
  h.theta = 1.0;
  h.nfree = (int)n;
  h.eps_g = (double)opt->eps_g;
  h.eps_f = (double)opt->eps_f;
  h.eps_x = (double)opt->eps_x;
  h.max_iteration = opt->max_iteration;
  h.has_after_iteration = cb->after_iteration != nullptr;
  h.has_custom_stop = cb->customized_stopping != nullptr;
  h.ibd = -1;
  int seen_iter = 0;
  auto clock_iter = [&]() {
    if (!iter_clock) return;
    const double t = now_seconds();
    while (seen_iter < h.iter && seen_iter < clock_cap) iter_clock[seen_iter++] = t;
  };
  int status = e->write_scalars();
  if (status == 0) status = e->setup_begin();
  R f = 0;
  if (status >= 0 && status != ST_DONE) {
    fill_summary(h, summary);
    cb->funcgrad(cb->funcgrad_ctx, (R*)e->x_ptr(0), &f, (R*)e->g_ptr(0), stream, summary);
    status = e->setup_finish((double)f);
  }
  int prog = PROG_ITER_START;
  double fval = 0.0;
  // With the library's own objective the line search reads each trial's
  // value from device memory, so the chain after an evaluation is launched
  // right behind it, with no host wait for f in between (the same values:
  // the callback returns that double rounded to R).  The first evaluation
  // above goes through the callback.
  lbfgsb_problem* own = own_objective<R>(cb);
  e->set_value_source(own ? own->fdev : nullptr);
  // every host round trip either evaluates f (at most 20 per iteration) or
  // restarts; this bound is never reached by a correct chain
  long long budget = 64LL * ((long long)opt->max_iteration + 2) + 4096;
  while (status >= 0 && status != ST_DONE) {
    if (--budget < 0) {
      std::fprintf(stderr, "lbfgsb_amd: iteration chain did not terminate\n");
      status = -1;
      break;
    }
    status = e->run(prog, -1, K_END, fval);
    if (status < 0) break;
    clock_iter();
    fill_summary(h, summary);
    if (status == ST_NEED_EVAL) {
      if (own) {
        if (problem_launch<R>(own, (R*)e->x_ptr(1), (R*)e->g_ptr(1), e->stream())) {
          status = -1;
          break;
        }
      } else {
        f = 0;
        cb->funcgrad(cb->funcgrad_ctx, (R*)e->x_ptr(1), &f, (R*)e->g_ptr(1), stream, summary);
        fval = (double)f;
      }
      prog = PROG_AFTER_EVAL;
    } else if (status == ST_ACCEPTED) {
      R fc = (R)h.f;
      cb->after_iteration(cb->after_iteration_ctx, (R*)e->x_ptr(0), &fc, (R*)e->g_ptr(0),
                          stream, summary);
      prog = PROG_AFTER_ACCEPT;
    } else if (status == ST_CHECKED) {
      R fc = (R)h.f;
      if (cb->customized_stopping(cb->customized_stopping_ctx, (R*)e->x_ptr(0), &fc,
                                  (R*)e->g_ptr(0), stream, summary)) {
        h.info = 0;
        h.status = ST_DONE;
        break;
      }
      prog = PROG_AFTER_CHECK;
    } else if (status == ST_RESTART) {
      prog = PROG_ITER_START;
    } else if (status == ST_REDO) {
      // the step was accepted other than pass A assumed (no memory update,
      // or a step the line search took without evaluating it): pass A again
      prog = PROG_AFTER_ACCEPT;
    } else if (status != ST_DONE) {
      status = -1;
    }
  }
  e->set_value_source(nullptr);
  int info;
  if (status < 0) {
    std::fprintf(stderr, "lbfgsb_amd: device error %s\n", hipGetErrorString(e->last_error));
    info = -2;
    h.info = -2;
  } else {
    info = h.info;
  }
  e->finalize_x();
  fill_summary(h, summary);
  summary->info = info;
  return info;
}

// The argument checks of lbfgsbminimize; returns m, or 0 after setting
// summary->info = -1.
template <typename Opt, typename Cb, typename Summary>
int check_minimize_args(int n, const Cb* cb, const Opt* opt, Summary* summary) {
  std::memset(summary, 0, sizeof(Summary));
  if (!cb || !cb->funcgrad || !opt) {
    summary->info = -1;
    return 0;
  }
  if (opt->mode != LBFGSB_AMD_MODE_GPU) {
    std::fprintf(stderr,
                 "lbfgsb_amd: mode %d is not provided (this library implements the GPU mode "
                 "only)\n",
                 opt->mode);
    summary->info = -1;
    return 0;
  }
  const int m = opt->hessian_approximate_dimension < 8 ? opt->hessian_approximate_dimension : 8;
  // errclb's scalar checks (lbfgsbcpu.cpp:879-890) and the driver's option
  // checks (lbfgsbcpu.cpp:259-264)
  if (n <= 0 || m <= 0 || m > n || opt->eps_f < 0 || opt->max_iteration < 0 ||
      opt->eps_g < 0 || opt->eps_x < 0) {
    summary->info = -1;
    return 0;
  }
  return m;
}

template <typename R, typename Opt, typename Cb, typename Summary>
int solve_on(EngineBase* e, int n, int m, const Cb* cb, const Opt* opt, Summary* summary,
             void* stream) {
  DevScalars& h = e->host_scalars();
  std::memset(&h, 0, sizeof(DevScalars));
  h.n = n;
  h.m = m;
  e->write_scalars();
  return drive<R>(e, cb, opt, summary, stream, nullptr, 0);
}

template <typename R, typename Opt, typename Cb, typename Summary>
int minimize_impl(int dtype, int n, const Cb* cb, const Opt* opt, R* x, const int* nbd,
                  const R* l, const R* u, Summary* summary, void* stream) {
  const int m = check_minimize_args(n, cb, opt, summary);
  if (!m) return -1;
  EngineBase* e = make_engine(dtype, n, m, x, nbd, l, u, (hipStream_t)stream);
  if (!e) {
    summary->info = -2;
    return -2;
  }
  const int info = solve_on<R>(e, n, m, cb, opt, summary, stream);
  delete e;
  return info;
}

// The same solve with no allocation: the engine object on this stack, its
// device buffers carved from this GPU's scratch region (noalloc.hip).
template <typename R, typename Opt, typename Cb, typename Summary>
int minimize_noalloc_impl(int dtype, int n, const Cb* cb, const Opt* opt, R* x,
                          const int* nbd, const R* l, const R* u, Summary* summary,
                          void* stream) {
  const int m = check_minimize_args(n, cb, opt, summary);
  if (!m) return -1;
  void* region = nullptr;
  size_t cap = 0;
  if (noalloc_region(&region, &cap) != 0) {
    std::fprintf(stderr, "lbfgsb_amd: no scratch region on the current device\n");
    summary->info = -2;
    return -2;
  }
  const size_t need = engine_region_bytes(dtype, n, m);
  if (need > cap) {
    std::fprintf(stderr,
                 "lbfgsb_amd: lbfgsb_minimize_noalloc: n = %d needs %zu bytes of device "
                 "scratch, the allocation-free region holds %zu; use lbfgsb_minimize\n",
                 n, need, cap);
    summary->info = -3;
    return -3;
  }
  alignas(64) unsigned char storage[kEngineObjectBytes];
  EngineBase* e = make_engine_in(storage, dtype, n, m, x, nbd, l, u, (hipStream_t)stream,
                                 region, cap);
  if (!e) {
    summary->info = -2;
    return -2;
  }
  const int info = solve_on<R>(e, n, m, cb, opt, summary, stream);
  e->~EngineBase();
  return info;
}

template <typename R>
int problem_funcgrad(lbfgsb_problem* p, R* x, R* f, R* g, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (problem_launch<R>(p, x, g, st)) return -1;
  // f is final on return, as the reference's callback contract requires
  if (mail_wait(st, &p->fhost->s, p->fseq) != hipSuccess) return -1;
  *f = (R)p->fhost->f;
  return 0;
}

}  // namespace

extern "C" {

void lbfgsb_default_options_f64(lbfgsb_options_f64* o) {
  o->eps_f = o->eps_g = o->eps_x = 1e-15;
  o->hessian_approximate_dimension = 8;
  o->machine_epsilon = 1e-15;
  o->machine_maximum = std::numeric_limits<double>::max();
  o->max_iteration = 1000;
  o->mode = LBFGSB_AMD_MODE_GPU;
  o->step_scaling = 1.0;
}

void lbfgsb_default_options_f32(lbfgsb_options_f32* o) {
  o->eps_f = o->eps_g = o->eps_x = 1e-15f;
  o->hessian_approximate_dimension = 8;
  o->machine_epsilon = 1e-8f;
  o->machine_maximum = std::numeric_limits<float>::max();
  o->max_iteration = 1000;
  o->mode = LBFGSB_AMD_MODE_GPU;
  o->step_scaling = 1.0f;
}

int lbfgsb_minimize_f64(int n, const lbfgsb_callbacks_f64* cb, const lbfgsb_options_f64* opt,
                        double* x, const int* nbd, const double* l, const double* u,
                        lbfgsb_summary_f64* summary, void* stream) {
  return minimize_impl<double>(1, n, cb, opt, x, nbd, l, u, summary, stream);
}

int lbfgsb_minimize_f32(int n, const lbfgsb_callbacks_f32* cb, const lbfgsb_options_f32* opt,
                        float* x, const int* nbd, const float* l, const float* u,
                        lbfgsb_summary_f32* summary, void* stream) {
  return minimize_impl<float>(0, n, cb, opt, x, nbd, l, u, summary, stream);
}

int lbfgsb_minimize_noalloc_f64(int n, const lbfgsb_callbacks_f64* cb,
                                const lbfgsb_options_f64* opt, double* x, const int* nbd,
                                const double* l, const double* u, lbfgsb_summary_f64* summary,
                                void* stream) {
  return minimize_noalloc_impl<double>(1, n, cb, opt, x, nbd, l, u, summary, stream);
}

int lbfgsb_minimize_noalloc_f32(int n, const lbfgsb_callbacks_f32* cb,
                                const lbfgsb_options_f32* opt, float* x, const int* nbd,
                                const float* l, const float* u, lbfgsb_summary_f32* summary,
                                void* stream) {
  return minimize_noalloc_impl<float>(0, n, cb, opt, x, nbd, l, u, summary, stream);
}

lbfgsb_engine* lbfgsb_engine_create(int dtype, long long n, int m, void* x, const int* nbd,
                                    const void* l, const void* u, void* stream) {
  if (n <= 0 || m <= 0 || m > 8 || (dtype != 0 && dtype != 1)) return nullptr;
  EngineBase* e = make_engine(dtype, n, m, x, nbd, l, u, (hipStream_t)stream);
  if (!e) return nullptr;
  DevScalars& h = e->host_scalars();
  std::memset(&h, 0, sizeof(DevScalars));
  h.n = n;
  h.m = m;
  h.theta = 1.0;
  h.nfree = (int)n;
  h.ibd = -1;
  if (e->write_scalars()) {
    delete e;
    return nullptr;
  }
  return new lbfgsb_engine{e};
}

void lbfgsb_engine_destroy(lbfgsb_engine* e) {
  if (!e) return;
  delete e->e;
  delete e;
}

int lbfgsb_engine_setup_begin(lbfgsb_engine* e, double eps_g, double eps_f, double eps_x,
                              int max_iteration) {
  DevScalars& h = e->e->host_scalars();
  if (e->e->read_scalars()) return -1;
  h.eps_g = eps_g;
  h.eps_f = eps_f;
  h.eps_x = eps_x;
  h.max_iteration = max_iteration;
  if (e->e->write_scalars()) return -1;
  return e->e->setup_begin();
}

int lbfgsb_engine_setup_finish(lbfgsb_engine* e, double f) { return e->e->setup_finish(f); }

int lbfgsb_engine_run(lbfgsb_engine* e, int program, int start_step, int stop_step, double f) {
  return e->e->run(program, start_step, stop_step, f);
}

void* lbfgsb_engine_x(lbfgsb_engine* e, int which) {
  e->e->read_scalars();
  return e->e->x_ptr(which);
}
void* lbfgsb_engine_g(lbfgsb_engine* e, int which) {
  e->e->read_scalars();
  return e->e->g_ptr(which);
}

int lbfgsb_engine_get_scalar(lbfgsb_engine* e, const char* name, double* value) {
  if (std::strcmp(name, "max_sorted_round") == 0) {
    *value = (double)e->e->max_sorted_round;
    return 0;
  }
  const ScalarField* f = find_field(name);
  if (!f) return -2;
  if (e->e->read_scalars()) return -1;
  const char* base = (const char*)&e->e->host_scalars();
  if (f->type == 0)
    *value = *(const int*)(base + f->off);
  else if (f->type == 1)
    *value = (double)*(const long long*)(base + f->off);
  else
    *value = *(const double*)(base + f->off);
  return 0;
}

int lbfgsb_engine_set_scalar(lbfgsb_engine* e, const char* name, double value) {
  const ScalarField* f = find_field(name);
  if (!f) return -2;
  if (e->e->read_scalars()) return -1;
  char* base = (char*)&e->e->host_scalars();
  if (f->type == 0)
    *(int*)(base + f->off) = (int)value;
  else if (f->type == 1)
    *(long long*)(base + f->off) = (long long)value;
  else
    *(double*)(base + f->off) = value;
  return e->e->write_scalars();
}

int lbfgsb_engine_get_array(lbfgsb_engine* e, const char* name, void* out, long long count) {
  return e->e->get_array(name, out, count);
}
int lbfgsb_engine_set_array(lbfgsb_engine* e, const char* name, const void* in,
                            long long count) {
  return e->e->set_array(name, in, count);
}
int lbfgsb_engine_finalize(lbfgsb_engine* e) { return e->e->finalize_x(); }

int lbfgsb_engine_solve_f64(lbfgsb_engine* e, const lbfgsb_callbacks_f64* cb,
                            const lbfgsb_options_f64* opt, lbfgsb_summary_f64* summary,
                            double* iter_clock, int clock_cap) {
  std::memset(summary, 0, sizeof(*summary));
  if (!e || e->e->dtype() != 1 || !cb || !cb->funcgrad || !opt) return summary->info = -1;
  return drive<double>(e->e, cb, opt, summary, (void*)e->e->stream(), iter_clock, clock_cap);
}

int lbfgsb_engine_solve_f32(lbfgsb_engine* e, const lbfgsb_callbacks_f32* cb,
                            const lbfgsb_options_f32* opt, lbfgsb_summary_f32* summary,
                            double* iter_clock, int clock_cap) {
  std::memset(summary, 0, sizeof(*summary));
  if (!e || e->e->dtype() != 0 || !cb || !cb->funcgrad || !opt) return summary->info = -1;
  return drive<float>(e->e, cb, opt, summary, (void*)e->e->stream(), iter_clock, clock_cap);
}

void lbfgsb_engine_set_profiling(lbfgsb_engine* e, int from_iter) {
  e->e->set_profiling(from_iter, ~0u);
}
void lbfgsb_engine_set_profiling_steps(lbfgsb_engine* e, int from_iter, unsigned step_mask) {
  e->e->set_profiling(from_iter, step_mask);
}

int lbfgsb_engine_step_time(lbfgsb_engine* e, int step, double* total_ms, long long* count) {
  e->e->step_time(step, total_ms, count);
  return 0;
}

lbfgsb_problem* lbfgsb_problem_create(int dtype, int kind, long long n, long long nx,
                                      long long ny, double h, const void* b, const void* c,
                                      const void* a0, const void* a1, const void* a2) {
  if (n <= 0 || !((kind >= OBJK_GRID_QUAD && kind <= OBJK_PENTA) || kind == OBJK_DSSCFG))
    return nullptr;
  if ((kind == OBJK_GRID_QUAD || kind == OBJK_GRID_COS || kind == OBJK_DSSCFG) &&
      (nx <= 0 || ny <= 0 || nx * ny != n))
    return nullptr;
  if (kind == OBJK_ROSEN && n % 2) return nullptr;
  lbfgsb_problem* p = new lbfgsb_problem();
  p->dtype = dtype;
  p->n = n;
  p->desc.kind = kind;
  p->desc.pad = 0;
  p->desc.nx = nx;
  p->desc.ny = ny;
  p->desc.h = h;
  p->desc.b = b;
  p->desc.c = c;
  p->desc.a0 = a0;
  p->desc.a1 = a1;
  p->desc.a2 = a2;
  long long blocks = (n + kBlock - 1) / kBlock;
  p->G = (int)(blocks < kMaxBlocks ? (blocks < 1 ? 1 : blocks) : kMaxBlocks);
  if (hipMalloc((void**)&p->part, (size_t)p->G * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&p->fdev, sizeof(double)) != hipSuccess ||
      hipHostMalloc((void**)&p->fhost, sizeof(HostF), hipHostMallocCoherent) != hipSuccess) {
    delete p;
    return nullptr;
  }
  std::memset(p->fhost, 0, sizeof(HostF));
  p->fseq = 0;
  return p;
}

void lbfgsb_problem_destroy(lbfgsb_problem* p) {
  if (!p) return;
  hipFree(p->part);
  hipFree(p->fdev);
  hipHostFree(p->fhost);
  delete p;
}

int lbfgsb_problem_funcgrad_f64(void* ctx, double* x, double* f, double* g, void* stream,
                                const lbfgsb_summary_f64*) {
  return problem_funcgrad<double>((lbfgsb_problem*)ctx, x, f, g, stream);
}

int lbfgsb_problem_funcgrad_f32(void* ctx, float* x, float* f, float* g, void* stream,
                                const lbfgsb_summary_f32*) {
  return problem_funcgrad<float>((lbfgsb_problem*)ctx, x, f, g, stream);
}

int lbfgsb_problem_eval_device(lbfgsb_problem* p, const void* x, void* g, double* f_dev,
                               void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (p->dtype == 1)
    launch_objective<double>(st, p->desc, p->n, (const double*)x, (double*)g, p->part, p->G,
                             f_dev);
  else
    launch_objective<float>(st, p->desc, p->n, (const float*)x, (float*)g, p->part, p->G,
                            f_dev);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

const char* lbfgsb_amd_version(void) { return "lbfgsb_amd 0.1 (gfx950)"; }

}  // extern "C"
