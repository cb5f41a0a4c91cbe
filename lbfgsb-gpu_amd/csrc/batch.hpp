// Batched form: many independent bound-constrained problems solved by one
// persistent kernel (batch.hip).  Host and device share these structs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"
#include "objectives.hpp"
#include "state.hpp"

namespace lbfgsb {

// One problem of a batch: its slice of the concatenated n-vectors.
struct BatchProb {
  long long off;   // first element in the concatenated n-vectors
  long long n;
  long long woff;  // first element of its W block (wsize(n) elements)
  ObjDesc obj;     // objective with pointers at this problem's slice
};

// Concatenated arrays of a batch (device pointers, problem p at probs[p].off).
template <typename T>
struct BatchBufs {
  T* x;   // caller's iterates (in/out)
  T* xb;  // trial points
  T* g;
  T* gb;
  const T* l;
  const T* u;
  const int* nbd;
  uint8_t* iwhere;
  T* W;
  T* d;
  T* z;
  T* r;
  T* dsub;
  Cand* ca;  // breakpoint candidates, n per problem
  Cand* cb;
};

// Per-problem outcome (written when a problem finishes or is suspended).
struct BatchResult {
  double f, residual_g, residual_f, residual_x;
  int iter, info, nfgv, done;
  // device wall clock (wall_clock64 ticks) when the problem was first taken
  // up and when it finished
  long long clk_start, clk_end;
};

// Launch the persistent kernel: `blocks` workgroups of `threads` (256 or
// 512) take problems order[0..count) from the queue counter *next
// (zeroed by the caller).  iter_budget > 0 suspends a problem after that
// many iterations in this launch (its state stays in dsa; a later launch
// resumes it); 0 runs every problem to completion.
template <typename T>
int launch_batch(hipStream_t s, int threads, int blocks, const BatchProb* probs, const int* order,
                 int count, unsigned* next, BatchBufs<T> B, DevScalars* dsa, BatchResult* res,
                 int iter_budget);

}  // namespace lbfgsb
