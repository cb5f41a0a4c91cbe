// Host engine: owns the device state of one solve and enqueues the kernel
// chain of an iteration.  The chain is speculative: every kernel checks the
// device halt flag, so a whole iteration (or everything between two function
// evaluations) is enqueued at once and the host reads one status block per
// evaluation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <string>

#include "launch.hpp"
#include "state.hpp"

namespace lbfgsb {

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int dtype() const = 0;  // 0 = fp32 bulk, 1 = fp64 bulk
  virtual int setup_begin() = 0;  // errclb + active; returns status
  virtual int setup_finish(double f) = 0;
  virtual int run(int program, int start_step, int stop_step, double f) = 0;
  // While set, the line-search steps of run() read the evaluated value from
  // this device double (written on the engine's stream by the objective;
  // rounded to fp32 for an fp32 engine, as a callback would return it)
  // instead of taking f.
  void set_value_source(const double* fdev) { fdev_src = fdev; }
  virtual DevScalars& host_scalars() = 0;
  virtual int read_scalars() = 0;
  virtual int write_scalars() = 0;
  virtual void* x_ptr(int which) = 0;  // 0 = accepted x, 1 = trial x
  virtual void* g_ptr(int which) = 0;
  virtual int get_array(const std::string& name, void* out, long long count) = 0;
  virtual int set_array(const std::string& name, const void* in, long long count) = 0;
  virtual int finalize_x() = 0;  // leaves the accepted x in the caller's buffer
  virtual hipStream_t stream() const = 0;
  virtual long long size() const = 0;
  // Per-step kernel timing with HIP events on the engine's stream.
  virtual void set_profiling(int from_iter, unsigned step_mask) = 0;  // -1 = off
  virtual void step_time(int step, double* ms, long long* count) = 0;
  hipError_t last_error = hipSuccess;
  const double* fdev_src = nullptr;
  // statistics of the host-sized breakpoint rounds: the most candidates one
  // round sorted (tiles of kSortTile, then one merge-path pass per doubling)
  long long max_sorted_round = 0;
};

EngineBase* make_engine(int dtype, long long n, int m, void* x, const int* nbd, const void* l,
                        const void* u, hipStream_t stream);

// An engine that allocates nothing: its object is built in `storage`
// (kEngineObjectBytes, 64-byte aligned) and every device buffer is carved
// from the device region [region, region + region_bytes).  nullptr when the
// region is too small.  Destroy with e->~EngineBase().
constexpr size_t kEngineObjectBytes = 16384;
EngineBase* make_engine_in(void* storage, int dtype, long long n, int m, void* x,
                           const int* nbd, const void* l, const void* u, hipStream_t stream,
                           void* region, size_t region_bytes);
// Bytes of device region an engine of this size needs.
size_t engine_region_bytes(int dtype, long long n, int m);

// Device scratch region for allocation-free solves (noalloc.hip): this
// device's address and size, looked up once per device.
int noalloc_region(void** ptr, size_t* bytes);
void launch_fill_zero(hipStream_t s, void* p, size_t bytes);

}  // namespace lbfgsb
