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
  int gang;        // workgroups that solve it together (1 = one workgroup)
  int pad;
};

// One entry of the launch queue: member `member` of problem `prob`'s gang.
// The members of a gang are consecutive in the queue.
struct BatchItem {
  int prob;
  int member;
};

// Per-problem gang control, zeroed before every launch: the arrival counter
// of the gang barrier, the breakpoint-candidate counter of the Cauchy
// rounds, and an error flag (a barrier that timed out).
struct GangCtl {
  unsigned arrive;
  int err;
  unsigned long long cand;
};

// Doubles of one member's exchange area (per barrier parity): the sums of
// the largest pass (P4_COUNT) and the max/min/argmin slots.
constexpr int kGangPart = 200;
// A gang member waits at most this long (wall-clock ticks at 100 MHz) for
// the others at a barrier before it marks the problem failed.
constexpr long long kGangWaitTicks = 30LL * 100000000LL;
// Smallest segment a gang member takes (elements).
constexpr long long kMinSegment = 4096;

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

// Launch the persistent kernels: the gang kernel (blocks_gang workgroups)
// over the gang entries items[0 .. count[0]) through next[0], then the
// one-workgroup kernel (blocks_one) over items[count[0] .. count[0] +
// count[1]) through next[1] (0 workgroups skips a kernel).  All workgroups
// of the gang kernel must be able to be resident at once (the members of a
// gang wait for each other).  iter_budget > 0 suspends a problem after that
// many iterations in this launch (its state stays in dsa; a later launch
// resumes it); 0 runs every problem to completion.  gpart holds
// 2 x kGangPart doubles per gang entry.
template <typename T>
int launch_batch(hipStream_t s, int threads, int blocks_gang, int blocks_one,
                 const BatchProb* probs, const BatchItem* items, const int* count, unsigned* next,
                 GangCtl* gang, double* gpart, BatchBufs<T> B, DevScalars* dsa,
                 BatchResult* res, int iter_budget);

// Zero the two queue counters and the gang controls of nprob problems (a
// kernel, so that the launch sequence can be captured in a graph).
void launch_batch_clear(hipStream_t s, unsigned* next, GangCtl* gang, int nprob);

// Workgroups of the batch kernels of `threads` that fit on one CU at once.
template <typename T>
int batch_blocks_per_cu(int threads);

}  // namespace lbfgsb
