// Descriptor of a built-in objective (device pointers of the bulk type).
#pragma once

#include <hip/hip_runtime.h>

#include "host_mailbox.hpp"

namespace lbfgsb {

enum ObjKind : int {
  OBJK_GRID_QUAD = 1,
  OBJK_GRID_COS = 2,
  OBJK_ROSEN = 3,
  OBJK_PENTA = 4,
  OBJK_DSSCFG = 6  // the reference example's problem; h holds lambda
};

struct ObjDesc {
  int kind;
  int pad;
  long long nx, ny;
  double h;
  const void* b;
  const void* c;
  const void* a0;
  const void* a1;
  const void* a2;
};

// g, and f into f_dev; with hf, f also into the host mailbox hf followed by
// the sequence number seq.
template <typename T>
void launch_objective(hipStream_t s, const ObjDesc& o, long long n, const T* x, T* g,
                      double* part, int G, double* f_dev, HostF* hf = nullptr,
                      unsigned seq = 0);

void launch_export_index(hipStream_t s, const uint8_t* iwhere, long long n, int track,
                         unsigned* counts, unsigned long long* totals, int* index, int* indx2);
long long export_count_blocks(long long n);

}  // namespace lbfgsb
