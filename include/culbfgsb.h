/*
 * C++ drop-in API of the MI355X L-BFGS-B engine.
 *
 * Same names, fields and semantics as the reference's culbfgsb/culbfgsb.h
 * (LBFGSB_CUDA_OPTION / SUMMARY / STATE, lbfgsbcuda::lbfgsbminimize and
 * lbfgsbcuda::lbfgsbdefaultoption, culbfgsb.h:31-166), header-only over the C
 * ABI in lbfgsb_amd.h.  Differences a caller sees:
 *   - the callbacks take `const hipStream_t&` (the stream the solve runs on;
 *     the null stream unless the stream overload is used);
 *   - m_cublas_handle is kept as an ignored `void*` so that code which
 *     zero-initialises the state still compiles;
 *   - mode LCM_NO_ACCELERATION is not provided by this library (summary.info
 *     = -1 and a message), keep the reference's CPU path for it;
 *   - summary.info carries the reference CPU path's codes (-2 .. 5) instead
 *     of being forced to 0 at exit.
 * x, nbd, l, u are device arrays, as in the reference's LCM_CUDA mode.
 * Link with -llbfgsb_amd (lbfgsb-gpu_amd/lib).
 */
#ifndef CULBFGSB_CULBFGSB_H_
#define CULBFGSB_CULBFGSB_H_

#include <hip/hip_runtime_api.h>

#include <functional>

#include "lbfgsb_amd.h"

enum LBFGSB_CUDA_MODE { LCM_NO_ACCELERATION, LCM_CUDA };

// machine_epsilon, machine_maximum and step_scaling are kept for source
// compatibility and ignored: the update-skip test uses the CPU path's fixed
// 5e-16 in both precisions (culbfgsb/lbfgsbcpu.cpp:443, culbfgsb/ap.cpp:38).
template <typename real>
struct LBFGSB_CUDA_OPTION {
  real machine_epsilon;
  real machine_maximum;
  real step_scaling;
  real eps_g;
  real eps_f;
  real eps_x;
  int hessian_approximate_dimension;
  int max_iteration;
  LBFGSB_CUDA_MODE mode;
};

template <typename real>
struct LBFGSB_CUDA_SUMMARY {
  real residual_g;
  real residual_f;
  real residual_x;
  int num_iteration;
  int info;
};

template <typename real>
struct LBFGSB_CUDA_STATE {
  void* m_cublas_handle;  // ignored (no BLAS library is used)
  std::function<int(real*, real&, real*, const hipStream_t&,
                    const LBFGSB_CUDA_SUMMARY<real>&)>
      m_funcgrad_callback;
  std::function<int(real*, real&, real*, const hipStream_t&,
                    const LBFGSB_CUDA_SUMMARY<real>&)>
      m_after_iteration_callback;
  std::function<int(real*, real&, real*, const hipStream_t&,
                    const LBFGSB_CUDA_SUMMARY<real>&)>
      m_customized_stopping_callback;
};

namespace lbfgsbcuda {
namespace detail {

template <typename real>
struct CTypes;
template <>
struct CTypes<double> {
  typedef lbfgsb_options_f64 Options;
  typedef lbfgsb_summary_f64 Summary;
  typedef lbfgsb_callbacks_f64 Callbacks;
  static int minimize(int n, const Callbacks* cb, const Options* o, double* x, const int* nbd,
                      const double* l, const double* u, Summary* s, void* st) {
    return lbfgsb_minimize_f64(n, cb, o, x, nbd, l, u, s, st);
  }
  static void defaults(Options* o) { lbfgsb_default_options_f64(o); }
};
template <>
struct CTypes<float> {
  typedef lbfgsb_options_f32 Options;
  typedef lbfgsb_summary_f32 Summary;
  typedef lbfgsb_callbacks_f32 Callbacks;
  static int minimize(int n, const Callbacks* cb, const Options* o, float* x, const int* nbd,
                      const float* l, const float* u, Summary* s, void* st) {
    return lbfgsb_minimize_f32(n, cb, o, x, nbd, l, u, s, st);
  }
  static void defaults(Options* o) { lbfgsb_default_options_f32(o); }
};

template <typename real>
struct Trampoline {
  typedef std::function<int(real*, real&, real*, const hipStream_t&,
                            const LBFGSB_CUDA_SUMMARY<real>&)>
      Fn;
  const Fn* fn;
  static int call(void* ctx, real* x, real* f, real* g, void* stream,
                  const typename CTypes<real>::Summary* s) {
    const Trampoline* t = static_cast<const Trampoline*>(ctx);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    return (*t->fn)(x, *f, g, st, *reinterpret_cast<const LBFGSB_CUDA_SUMMARY<real>*>(s));
  }
};

}  // namespace detail

// lbfgsbminimize on an explicit stream (extension of the reference API).
template <typename real>
inline void lbfgsbminimize(const int& n, const LBFGSB_CUDA_STATE<real>& state,
                           const LBFGSB_CUDA_OPTION<real>& option, real* x, const int* nbd,
                           const real* l, const real* u, LBFGSB_CUDA_SUMMARY<real>& summary,
                           hipStream_t stream) {
  typedef detail::CTypes<real> C;
  static_assert(sizeof(LBFGSB_CUDA_SUMMARY<real>) == sizeof(typename C::Summary),
                "summary layout");
  typename C::Options o;
  o.machine_epsilon = option.machine_epsilon;
  o.machine_maximum = option.machine_maximum;
  o.step_scaling = option.step_scaling;
  o.eps_g = option.eps_g;
  o.eps_f = option.eps_f;
  o.eps_x = option.eps_x;
  o.hessian_approximate_dimension = option.hessian_approximate_dimension;
  o.max_iteration = option.max_iteration;
  o.mode = option.mode == LCM_CUDA ? LBFGSB_AMD_MODE_GPU : LBFGSB_AMD_MODE_CPU;
  detail::Trampoline<real> tf{&state.m_funcgrad_callback};
  detail::Trampoline<real> ta{&state.m_after_iteration_callback};
  detail::Trampoline<real> ts{&state.m_customized_stopping_callback};
  typename C::Callbacks cb;
  cb.funcgrad = state.m_funcgrad_callback ? &detail::Trampoline<real>::call : nullptr;
  cb.funcgrad_ctx = &tf;
  cb.after_iteration = state.m_after_iteration_callback ? &detail::Trampoline<real>::call : nullptr;
  cb.after_iteration_ctx = &ta;
  cb.customized_stopping =
      state.m_customized_stopping_callback ? &detail::Trampoline<real>::call : nullptr;
  cb.customized_stopping_ctx = &ts;
  C::minimize(n, &cb, &o, x, nbd, l, u, reinterpret_cast<typename C::Summary*>(&summary),
              static_cast<void*>(stream));
}

// lbfgsbminimize (culbfgsb/culbfgsb.h:157-163), on the null stream.
template <typename real>
inline void lbfgsbminimize(const int& n, const LBFGSB_CUDA_STATE<real>& state,
                           const LBFGSB_CUDA_OPTION<real>& option, real* x, const int* nbd,
                           const real* l, const real* u, LBFGSB_CUDA_SUMMARY<real>& summary) {
  lbfgsbminimize<real>(n, state, option, x, nbd, l, u, summary, (hipStream_t)0);
}

// lbfgsbdefaultoption (culbfgsb/lbfgsb.cpp:44-66).
template <typename real>
inline void lbfgsbdefaultoption(LBFGSB_CUDA_OPTION<real>& option) {
  typename detail::CTypes<real>::Options o;
  detail::CTypes<real>::defaults(&o);
  option.machine_epsilon = o.machine_epsilon;
  option.machine_maximum = o.machine_maximum;
  option.step_scaling = o.step_scaling;
  option.eps_g = o.eps_g;
  option.eps_f = o.eps_f;
  option.eps_x = o.eps_x;
  option.hessian_approximate_dimension = o.hessian_approximate_dimension;
  option.max_iteration = o.max_iteration;
  option.mode = LCM_CUDA;
}

}  // namespace lbfgsbcuda

#endif  // CULBFGSB_CULBFGSB_H_
