// Host-side launch wrappers of the engine's kernels (one definition per
// .hip file; all launches are asynchronous on the given stream).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "host_mailbox.hpp"
#include "state.hpp"

namespace lbfgsb {

// bulk.hip (rev: the pass takes its chunks in descending order; see
// chunk_start in common.hpp)
template <typename T>
void launch_errclb(hipStream_t s, int G, long long n, const int* nbd, const T* l, const T* u,
                   double* part);
template <typename T>
void launch_active(hipStream_t s, int G, long long n, T* x, const T* l, const T* u,
                   const int* nbd, uint8_t* iwhere, double* part);
template <typename T>
void launch_projgr(hipStream_t s, int G, long long n, const T* x, const T* g, const T* l,
                   const T* u, const int* nbd, double* pmax);
// Block b of the pass lists the breakpoints of its elements with t below
// ds->gcp_near_tau in nearc[b * near_cap ..] (at most near_cap of them) and
// their count (listed or not) in nearn[b], for the Cauchy search's
// selections.
template <typename T>
void launch_pass_a(hipStream_t s, int G, bool fused, long long n, Pair<T> X, Pair<T> Gv,
                   const T* l, const T* u, uint8_t* iwhere, T* W, const T* d,
                   Cand* nearc, unsigned* nearn, int near_cap, DevScalars* ds, double* part,
                   double* pmax, double* pmin, long long* pmini, int rev, int spec = 0);
template <typename T>
void launch_p4(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
               const T* u, uint8_t* iwhere, const T* W, DevScalars* ds,
               double* part, int rev);
template <typename T>
void launch_p5a(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                const T* u, const uint8_t* iwhere, const T* W, T* dsub,
                DevScalars* ds, double* pmin, long long* pmini, int rev);
// x_cp of every variable from the current state (inspection only).
template <typename T>
void launch_cauchy_point(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                         const T* u, const uint8_t* iwhere,
                         const DevScalars* ds, T* xcp);
template <typename T>
void launch_p5b(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                const T* u, const uint8_t* iwhere, const T* dsub,
                DevScalars* ds, double* part, double* pmin, int rev);
template <typename T>
void launch_trial(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                  const T* u, const uint8_t* iwhere, const T* dsub, T* d,
                  const DevScalars* ds);
// The subspace point z of every variable from the current state into out
// (inspection only).
template <typename T>
void launch_subspace_point(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv,
                           const T* l, const T* u, const uint8_t* iwhere,
                           const T* dsub, const DevScalars* ds, T* out);
template <typename T>
void launch_gd(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* d,
               const DevScalars* ds, double* part, int rev);
// The line-search direction d of every variable into out (inspection only).
template <typename T>
void launch_direction(hipStream_t s, int G, long long n, Pair<T> X, const T* d,
                      const DevScalars* ds, T* out);

// small.hip
// The device state into `host` (fine-grained pinned memory), then seq into m.
void launch_publish(hipStream_t s, const DevScalars* ds, DevScalars* host, MailSeq* m,
                    unsigned seq);
// spec: sum even when the chain is halted (after the speculative pass A,
// which runs on a chain halted by an evaluation request).  With pmax (pass
// A's partials), one more block finishes the projected-gradient norm and
// the smallest breakpoint into ds->red_sbg, red_bk, red_ibk.
void launch_rowsum(hipStream_t s, const double* part, int G, int v0, int nrows, DevScalars* ds,
                   int p4_rows, int spec = 0, const double* pmax = nullptr,
                   const double* pmin = nullptr, const long long* pmini = nullptr);
void launch_resume(hipStream_t s, DevScalars* ds, int reset_bp);
void launch_setup_errclb(hipStream_t s, DevScalars* ds, const double* part, int G);
void launch_setup_active(hipStream_t s, DevScalars* ds, const double* part, int G);
void launch_setup_projgr(hipStream_t s, DevScalars* ds, const double* pmax, int G, double f);
void launch_stop_tests(hipStream_t s, DevScalars* ds, const double* part, const double* pmax,
                       const double* pmin, const long long* pmini, int G);
void launch_update_gcp(hipStream_t s, DevScalars* ds, const double* part, const double* pmin,
                       const long long* pmini, int G, int reduce);
void launch_formk_subsm(hipStream_t s, DevScalars* ds, const double* part, int G);
void launch_alpha(hipStream_t s, DevScalars* ds, const double* pmin, const long long* pmini,
                  int G);
void launch_ls_init(hipStream_t s, DevScalars* ds, const double* part, const double* pmin,
                    int G);
// The line-search kernels take the evaluated value as f, or when fdev is
// not null from the FP64 double there (rounded to fp32 when single is set).
void launch_ls_step(hipStream_t s, DevScalars* ds, const double* part, int G, double f,
                    const double* fdev, int single);
// Line-search step on the g'd of a speculative pass A (row PA_GD of the
// row-summed pass-A partials), then on acceptance the stopping tests.
void launch_ls_stop(hipStream_t s, DevScalars* ds, const double* pmax, const double* pmin,
                    const long long* pmini, int G, double f, const double* fdev, int single);

// gcp.hip
struct GcpScratch {
  double* tsum;
  double* tpre;
  int tstride;
  double* total;
  long long* tile_stop;
  double* tile_state;
  double* last_state;
};
// Pass A's breakpoint lists: G lists of up to `cap` candidates (c) and
// their counts (n).
struct NearLists {
  const Cand* c;
  const unsigned* n;
  int G;
  int cap;
};
// The candidates of the current t-range into ca: from pass A's lists when
// they cover it, else (full) from every variable's classification; without
// `full`, a range the lists do not cover halts the chain for the host to
// select.
template <typename T>
void launch_gcp_select(hipStream_t s, int G, long long n, Pair<T> X, Pair<T> Gv, const T* l,
                       const T* u, const uint8_t* iwhere, const NearLists& nl, DevScalars* ds,
                       Cand* ca, long long buf_cap, unsigned long long* bcount, bool full);
// One workgroup's round of at most kSmallRound candidates (larger ones are
// handed to the multi-block path): from ca, or (from_near) straight from
// pass A's lists, halting the chain for the host's selection when they do
// not cover the round.
template <typename T>
void launch_gcp_round_small(hipStream_t s, Cand* ca, Pair<T> Gv, const T* W, DevScalars* ds,
                            const NearLists& nl, bool from_near);
template <typename T>
// mid = 0: a host-sized round; mid = 1: the in-chain round after the
// one-workgroup round (see round_active in gcp.hip), capacity `cap`.
void launch_gcp_sortscan(hipStream_t s, Pair<T> Gv, const T* W, DevScalars* ds, Cand* ca,
                         Cand* cb, long long cap, const GcpScratch& w, int mid = 0);
// candidates the in-chain multi-block round takes without the host
constexpr long long kMidRound = 32768;

}  // namespace lbfgsb
