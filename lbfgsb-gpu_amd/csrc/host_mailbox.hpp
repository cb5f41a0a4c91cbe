// Device-to-host status delivery without a copy command or a blocking
// stream synchronisation.  A small kernel at the end of a chain writes its
// result into fine-grained (coherent) pinned host memory, then publishes a
// sequence number with a system-scope release store; the host spins on the
// sequence number with an acquire load.  Measured on MI355X: ~7 us from a
// kernel launch to the host seeing its result, against ~14 us for a copy
// plus hipStreamSynchronize.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>

namespace lbfgsb {

// Sequence word of a mailbox, on its own cache line.
struct alignas(64) MailSeq {
  unsigned seq;
  unsigned pad[15];
};

// The objective value of a built-in objective, delivered to the host.
struct alignas(64) HostF {
  double f;
  MailSeq s;
};

// Device side: publish `seq` after every store of the calling workgroup
// (each thread fences its own stores first; call from all threads).
__device__ inline void mail_publish(MailSeq* m, unsigned seq) {
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(&m->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Host side: wait until the mailbox holds `want`.  The stream is polled now
// and then so that a failed launch returns its error instead of spinning;
// after a second of spinning the wait blocks in hipStreamSynchronize, which
// the publishing kernel's completion ends.
inline hipError_t mail_wait(hipStream_t st, const MailSeq* m, unsigned want) {
  auto seen = [&]() { return __atomic_load_n(&m->seq, __ATOMIC_ACQUIRE) == want; };
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned k = 1;; k++) {
    if (seen()) return hipSuccess;
    if ((k & 1023) == 0) {
      const hipError_t q = hipStreamQuery(st);
      if (q != hipSuccess && q != hipErrorNotReady) return q;
      if (q == hipSuccess) return seen() ? hipSuccess : hipErrorUnknown;
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
        const hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        return seen() ? hipSuccess : hipErrorUnknown;
      }
    }
  }
}

}  // namespace lbfgsb
