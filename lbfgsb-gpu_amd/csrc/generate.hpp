// Device generator of the benchmark problems (generate.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace lbfgsb {

// A generated problem: objective kind (OBJK_*), size, grid shape (grid
// kinds: nx * ny == n) and seed (the per-problem seed of problems.py).
struct GenSpec {
  int kind;
  int pad;
  long long n, nx, ny;
  unsigned long long seed;
};

// 0 if the spec names a problem the generator builds, -1 otherwise.
int gen_check(const GenSpec& sp);

// Write the problem's x0, nbd, l, u and the objective data of its kind
// (grid quadratic: b; grid cos: b, c; Rosenbrock: none; pentadiagonal: b,
// a0, a1, a2) into device arrays of n elements.  Returns 0, -1 on a bad spec
// or a missing array, -2 on a launch error.
template <typename T>
int launch_generate(hipStream_t s, const GenSpec& sp, T* x0, int* nbd, T* l, T* u, T* b, T* c,
                    T* a0, T* a1, T* a2);

// dst[i] = src[i] for n elements, converting between the bulk types (dtype
// 0 float, 1 double): a batch slot started from another slot's iterate.
int launch_convert_copy(hipStream_t s, const void* src, int src_dtype, void* dst, int dst_dtype,
                        long long n);

}  // namespace lbfgsb
