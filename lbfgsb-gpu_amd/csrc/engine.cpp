// Host engine of the L-BFGS-B solver (see engine.hpp).
#include "engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "objectives.hpp"

namespace lbfgsb {

#define HIPCHK(expr)                       \
  do {                                     \
    hipError_t e_ = (expr);                \
    if (e_ != hipSuccess) {                \
      last_error = e_;                     \
      return -1;                           \
    }                                      \
  } while (0)

static const int kTail[] = {K_P4, K_FORMK, K_P5A, K_ALPHA, K_P5B, K_LSINIT, K_TRIAL};

// spec: after an evaluation, pass A runs on the trial point before the
// line-search step (which then takes g'd from it) instead of after it; the
// step is usually accepted, and otherwise its work is discarded.  Not with
// an after-iteration callback, which wants the accepted point first, and
// only for whole chains (partial runs name the steps of the regular form).
static std::vector<int> program_steps(int prog, bool spec) {
  std::vector<int> s;
  switch (prog) {
    case PROG_ITER_START:
      s = {K_RESUME_RESET, K_PASSA_SA, K_UPD_REDUCE, K_GCP};
      break;
    case PROG_AFTER_EVAL:
      if (spec)
        // a trial point the line-search step requests is written by the
        // tail's k_trial (the kernels between no-op on the halted chain), so
        // an accepted step launches no k_trial of its own
        s = {K_PASSA_SPEC, K_LSSTOP, K_GCP};
      else
        s = {K_GD, K_LSSTEP, K_TRIAL, K_PASSA_FUSED, K_STOP, K_UPD, K_GCP};
      break;
    case PROG_AFTER_ACCEPT:
      s = {K_RESUME_RESET, K_PASSA_FUSED, K_STOP, K_UPD, K_GCP};
      break;
    case PROG_AFTER_CHECK:
      s = {K_RESUME, K_UPD, K_GCP};
      break;
    default:
      break;
  }
  for (int k : kTail) s.push_back(k);
  return s;
}

template <typename T>
struct Engine final : EngineBase {
  long long n;
  int m;
  int G;
  hipStream_t st;
  T* X[2];
  T* Gr[2];
  Pair<T> dX{nullptr, nullptr};  // X[0], X[1] as kernel arguments
  Pair<T> dG{nullptr, nullptr};  // Gr[0], Gr[1]
  const T* l;
  const T* u;
  const int* nbd;
  T *W = nullptr, *d = nullptr, *dsub = nullptr;
  // pass A's breakpoint lists (per block, near_cap each) and their counts
  Cand* nearc = nullptr;
  unsigned* nearn = nullptr;
  int near_cap = 0;
  // the last chain's round took its candidates from the lists: the next
  // chain's selection starts there (a miss halts it for the host)
  bool near_hint = false;
  // lists (in entries, at the last round they fed) the one-workgroup round
  // filters itself; longer ones go through the multi-block gather
  static constexpr long long kRoundGatherMax = 4096;
  // x_cp or z on request (named arrays "xcp", "z"): the iteration stores no
  // x_cp, and z only in the trial buffer until a trial overwrites it
  T* insp = nullptr;
  uint8_t* iwhere = nullptr;
  DevScalars* ds = nullptr;
  DevScalars* hs = nullptr;  // fine-grained pinned: the end of a chain writes it directly
  MailSeq* mseq = nullptr;
  unsigned mail_seq = 0;
  double *part = nullptr, *pmax = nullptr, *pmin = nullptr;
  long long* pmini = nullptr;
  Cand *ca = nullptr, *cb = nullptr;
  double *tsum = nullptr, *tpre = nullptr, *total = nullptr;
  int tstride = 0;
  double *tile_state = nullptr, *last_state = nullptr;
  long long* tile_stop = nullptr;
  int* index = nullptr;
  int* indx2 = nullptr;
  unsigned* exp_counts = nullptr;
  unsigned long long* exp_totals = nullptr;
  bool ok = false;
  // Device memory: hipMalloc per buffer (default), or every buffer carved,
  // 256-byte aligned, from one region the caller owns (region != nullptr;
  // then no allocation is made, the status is read with a plain copy and
  // hs is the host copy below).  measuring: only sizes the region.
  unsigned char* region = nullptr;
  size_t region_cap = 0, region_off = 0;
  bool measuring = false;
  DevScalars hs_store;
  // chains since the last Cauchy round one workgroup could not take
  static constexpr int kMidHorizon = 16;
  int handoff_age = kMidHorizon;
  int last_handoffs = 0;
  void note_handoffs() {
    handoff_age = hs->gcp_handoffs != last_handoffs ? 0 : handoff_age + 1;
    last_handoffs = hs->gcp_handoffs;
  }

  // ---- optional per-step timing (HIP events on st)
  struct Pending {
    int step;
    hipEvent_t a, b;
  };
  int profile_from = -1;  // record kernels once the iteration count reaches this
  unsigned profile_mask = ~0u;  // steps that are recorded (bit per step id)
  std::vector<hipEvent_t> ev_pool;
  std::vector<Pending> pending;
  double step_ms[kNumSteps] = {};
  long long step_n[kNumSteps] = {};

  Engine(long long n_, int m_, T* x, const int* nbd_, const T* l_, const T* u_, hipStream_t s,
         unsigned char* region_ = nullptr, size_t region_bytes = 0, bool measure = false)
      : n(n_), m(m_), st(s), l(l_), u(u_), nbd(nbd_), region(region_),
        region_cap(region_bytes), measuring(measure) {
    X[0] = x;
    X[1] = Gr[0] = Gr[1] = nullptr;
    long long blocks = (n + kBlock - 1) / kBlock;
    G = (int)std::max<long long>(1, std::min<long long>(blocks, kMaxBlocks));
    ok = measuring ? carve_all() == 0 : alloc() == 0;
  }

  bool owns_memory() const { return !region && !measuring; }
  // per-block candidate counts and chunk masks of a selection (2 G words):
  // tile_state, which holds nothing between rounds and has 33 n / kBlock
  // words or more
  unsigned long long* sel_count() const { return (unsigned long long*)tile_state; }
  NearLists near_lists() const { return NearLists{nearc, nearn, G, near_cap}; }

  template <typename P>
  int dmalloc(P*& p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(P);
    if (!owns_memory()) {
      const size_t at = (region_off + 255) & ~(size_t)255;
      region_off = at + bytes;
      if (measuring) {
        p = nullptr;
        return 0;
      }
      if (region_off > region_cap) return -1;
      p = (P*)(region + at);
      return 0;
    }
    HIPCHK(hipMalloc((void**)&p, bytes));
    return 0;
  }

  // Every device buffer of a solve, in one order for all three modes.
  int carve_all() {
    const size_t nn = (size_t)n;
    if (dmalloc(X[1], nn) || dmalloc(Gr[0], nn) || dmalloc(Gr[1], nn)) return -1;
    if (dmalloc(W, wsize(nn)) || dmalloc(d, nn) || dmalloc(dsub, nn) ||
        dmalloc(iwhere, nn))
      return -1;
    // room for an eighth of a block's elements in its list (64 at least)
    near_cap = (int)std::max<long long>(64, (n + 8LL * G - 1) / (8LL * G));
    if (dmalloc(nearc, (size_t)G * near_cap) || dmalloc(nearn, G)) return -1;
    if (dmalloc(ds, 1)) return -1;
    const size_t nv = std::max<int>(P4_COUNT, PA_COUNT);
    if (dmalloc(part, nv * G) || dmalloc(pmax, G) || dmalloc(pmin, G) || dmalloc(pmini, G))
      return -1;
    const size_t ntiles = nn / kBlock + 2;
    tstride = (int)ntiles;
    if (dmalloc(ca, nn) || dmalloc(cb, nn) || dmalloc(tsum, 33 * ntiles) ||
        dmalloc(tpre, 33 * ntiles) || dmalloc(total, 40) || dmalloc(tile_state, 33 * ntiles) ||
        dmalloc(last_state, 40) || dmalloc(tile_stop, ntiles))
      return -1;
    return 0;
  }

  int alloc() {
    if (carve_all()) return -1;
    const size_t wbytes = wsize((size_t)n) * sizeof(T);
    if (owns_memory()) {
      HIPCHK(hipMemsetAsync(W, 0, wbytes, st));
      HIPCHK(hipHostMalloc((void**)&hs, sizeof(DevScalars), hipHostMallocCoherent));
      HIPCHK(hipHostMalloc((void**)&mseq, sizeof(MailSeq), hipHostMallocCoherent));
      std::memset(mseq, 0, sizeof(MailSeq));
    } else {
      // the region keeps the previous call's contents: the ring must start
      // at zero (a fill kernel, no memset node for callers under capture)
      launch_fill_zero(st, W, wbytes);
      HIPCHK(hipGetLastError());
      hs = &hs_store;
    }
    std::memset(hs, 0, sizeof(DevScalars));
    dX = Pair<T>{X[0], X[1]};
    dG = Pair<T>{Gr[0], Gr[1]};
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  }

  hipEvent_t take_event() {
    if (ev_pool.empty()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      return e;
    }
    hipEvent_t e = ev_pool.back();
    ev_pool.pop_back();
    return e;
  }
  // fold finished event pairs into the per-step totals (each pair precedes
  // a completed synchronisation point on the stream)
  void drain_events() {
    for (const Pending& p : pending) {
      float ms = 0.f;
      if (hipEventSynchronize(p.b) == hipSuccess &&
          hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        step_ms[p.step] += ms;
        step_n[p.step] += 1;
      }
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }
  void set_profiling(int from_iter, unsigned step_mask) override {
    hipStreamSynchronize(st);
    drain_events();
    profile_from = from_iter;
    profile_mask = step_mask;
    for (int i = 0; i < kNumSteps; i++) {
      step_ms[i] = 0;
      step_n[i] = 0;
    }
  }
  void step_time(int step, double* ms, long long* count) override {
    hipStreamSynchronize(st);
    drain_events();
    if (step < 0 || step >= kNumSteps) {
      *ms = 0;
      *count = 0;
      return;
    }
    *ms = step_ms[step];
    *count = step_n[step];
  }

  ~Engine() override {
    if (measuring) return;
    hipStreamSynchronize(st);
    drain_events();
    for (hipEvent_t e : ev_pool) hipEventDestroy(e);
    if (!owns_memory()) return;
    T* bufs[] = {X[1], Gr[0], Gr[1], W, d, dsub, insp};
    for (T* p : bufs) hipFree(p);
    hipFree(nearc);
    hipFree(nearn);
    hipFree(iwhere);
    hipFree(ds);
    hipHostFree(hs);
    hipHostFree(mseq);
    hipFree(part);
    hipFree(pmax);
    hipFree(pmin);
    hipFree(pmini);
    hipFree(ca);
    hipFree(cb);
    hipFree(tsum);
    hipFree(tpre);
    hipFree(total);
    hipFree(tile_state);
    hipFree(last_state);
    hipFree(tile_stop);
    hipFree(index);
    hipFree(indx2);
    hipFree(exp_counts);
    hipFree(exp_totals);
  }

  int dtype() const override { return sizeof(T) == 8 ? 1 : 0; }
  hipStream_t stream() const override { return st; }
  long long size() const override { return n; }
  DevScalars& host_scalars() override { return *hs; }

  int read_scalars() override {
    HIPCHK(hipMemcpyAsync(hs, ds, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!pending.empty()) drain_events();
    return 0;
  }
  // The same at the end of a chain: a kernel writes the state into hs and
  // publishes a sequence number the host spins on (host_mailbox.hpp).
  int read_status() {
    if (!owns_memory()) return read_scalars();  // no pinned mailbox in a region
    launch_publish(st, ds, hs, mseq, ++mail_seq);
    HIPCHK(hipGetLastError());
    return wait_status();
  }
  int wait_status() {
    HIPCHK(mail_wait(st, mseq, mail_seq));
    if (!pending.empty()) drain_events();
    return 0;
  }
  // Enqueues steps[a, b) and reads the status at their end.  k_trial only
  // reads the device state, so when it is the last step the state is
  // published before it: the host wakes and enqueues what follows (the
  // objective on the trial point) while k_trial runs.
  int enqueue_and_read(const std::vector<int>& steps, size_t a, size_t b, double f) {
    const bool early = owns_memory() && b > a && steps[b - 1] == K_TRIAL;
    for (size_t k = a; k < (early ? b - 1 : b); k++) enqueue(steps[k], f);
    if (!early) {
      HIPCHK(hipGetLastError());
      return read_status();
    }
    launch_publish(st, ds, hs, mseq, ++mail_seq);
    enqueue(K_TRIAL, f);
    HIPCHK(hipGetLastError());
    return wait_status();
  }
  int write_scalars() override {
    HIPCHK(hipMemcpyAsync(ds, hs, sizeof(DevScalars), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  }

  int cur() const { return hs->phase & 1; }
  void* x_ptr(int which) override {
    if (which == 0) return X[cur()];
    return X[hs->trial_dst];
  }
  void* g_ptr(int which) override {
    if (which == 0) return Gr[cur()];
    return Gr[hs->trial_dst];
  }

  // errclb (device part) and the projection of x0 (lbfgsbcpu.cpp:258-266).
  int setup_begin() override {
    launch_errclb<T>(st, G, n, nbd, l, u, part);
    launch_setup_errclb(st, ds, part, G);
    HIPCHK(hipGetLastError());
    if (read_scalars()) return -1;
    if (hs->status == ST_DONE) return ST_DONE;  // x is left untouched on bad input
    launch_active<T>(st, G, n, X[0], l, u, nbd, iwhere, part);
    launch_setup_active(st, ds, part, G);
    HIPCHK(hipGetLastError());
    return read_scalars() ? -1 : hs->status;
  }
  // after the first evaluation at x0 (lbfgsbcpu.cpp:271-278)
  int setup_finish(double f) override {
    launch_projgr<T>(st, G, n, X[0], Gr[0], l, u, nbd, pmax);
    launch_setup_projgr(st, ds, pmax, G, f);
    HIPCHK(hipGetLastError());
    return read_scalars() ? -1 : hs->status;
  }

  void enqueue(int step, double f) {
    if (profile_from < 0 || hs->iter < profile_from || !((profile_mask >> step) & 1u))
      return launch_step(step, f);
    hipEvent_t a = take_event(), b = take_event();
    if (!a || !b) return launch_step(step, f);
    hipEventRecord(a, st);
    launch_step(step, f);
    hipEventRecord(b, st);
    pending.push_back({step, a, b});
  }

  void launch_step(int step, double f) {
    switch (step) {
      case K_RESUME_RESET:
        launch_resume(st, ds, 1);
        break;
      case K_RESUME:
        launch_resume(st, ds, 0);
        break;
      case K_PASSA_SA:
        launch_pass_a<T>(st, G, false, n, dX, dG, l, u, iwhere, W, d, nearc, nearn, near_cap,
                         ds, part, pmax, pmin, pmini, 0);
        break;
      case K_UPD_REDUCE:
        launch_rowsum(st, part, G, 0, PA_COUNT, ds, 0);
        launch_update_gcp(st, ds, part, pmin, pmini, G, 1);
        break;
      case K_PASSA_FUSED:
        launch_pass_a<T>(st, G, true, n, dX, dG, l, u, iwhere, W, d, nearc, nearn, near_cap,
                         ds, part, pmax, pmin, pmini, 0);
        break;
      case K_GCP: {
        // from pass A's lists when the last round could (a miss halts the
        // chain for the host's selection): inside the one-workgroup round
        // while the lists are short, by the multi-block gather when they
        // are long; otherwise the gather and the full selection, whichever
        // covers the round
        const bool in_round = near_hint && hs->gcp_near_listed <= kRoundGatherMax;
        if (!in_round)
          launch_gcp_select<T>(st, G, n, dX, dG, l, u, iwhere, near_lists(), ds, ca, n,
                               sel_count(), !near_hint);
        launch_gcp_round_small<T>(st, ca, dG, W, ds, near_lists(), in_round);
        // a round too large for one workgroup but within kMidRound finishes
        // here on the device; only larger ones and later ranges reach the
        // host (run_rounds).  Enqueued only while such rounds are recent:
        // when they are not, its idle launches would cost more than they save
        if (handoff_age < kMidHorizon) {
          GcpScratch w{tsum, tpre, tstride, total, tile_stop, tile_state, last_state};
          launch_gcp_sortscan<T>(st, dG, W, ds, ca, cb, std::min<long long>(n, kMidRound), w,
                                 1);
        }
        break;
      }
      case K_STOP:
        launch_rowsum(st, part, G, 0, PA_COUNT, ds, 0, 0, pmax, pmin, pmini);
        launch_stop_tests(st, ds, part, pmax, pmin, pmini, G);
        break;
      case K_PASSA_SPEC:
        launch_pass_a<T>(st, G, true, n, dX, dG, l, u, iwhere, W, d, nearc, nearn, near_cap,
                         ds, part, pmax, pmin, pmini, 0, 1);
        break;
      case K_LSSTOP:
        launch_rowsum(st, part, G, 0, PA_COUNT, ds, 0, 1, pmax, pmin, pmini);
        launch_ls_stop(st, ds, pmax, pmin, pmini, G, f, fdev_src, sizeof(T) == 4);
        break;
      case K_UPD:
        launch_update_gcp(st, ds, part, pmin, pmini, G, 0);
        break;
      case K_P4:
        launch_p4<T>(st, G, n, dX, dG, l, u, iwhere, W, ds, part, 1);
        break;
      case K_FORMK:
        launch_rowsum(st, part, G, 0, P4_COUNT, ds, 1);
        launch_formk_subsm(st, ds, part, G);
        break;
      case K_P5A:
        launch_p5a<T>(st, G, n, dX, dG, l, u, iwhere, W, dsub, ds, pmin, pmini, 0);
        break;
      case K_ALPHA:
        launch_alpha(st, ds, pmin, pmini, G);
        break;
      case K_P5B:
        launch_p5b<T>(st, G, n, dX, dG, l, u, iwhere, dsub, ds, part, pmin, 1);
        break;
      case K_LSINIT:
        launch_ls_init(st, ds, part, pmin, G);
        break;
      case K_TRIAL:
        launch_trial<T>(st, G, n, dX, dG, l, u, iwhere, dsub, d, ds);
        break;
      case K_GD:
        launch_gd<T>(st, G, n, dX, dG, d, ds, part, 1);
        break;
      case K_LSSTEP:
        launch_ls_step(st, ds, part, G, f, fdev_src, sizeof(T) == 4);
        break;
      default:
        break;
    }
  }

  // Host-sized breakpoint rounds of the Cauchy search (gcp.hip): used when a
  // round holds more candidates than one workgroup sorts, and for the rounds
  // after the first.  Entered either with the current t-range selected (the
  // one-workgroup round found too many candidates) or with a new t-range
  // still to select (gcp_need_select, set when a round passed all of its
  // candidates).
  int run_rounds() {
    GcpScratch w{tsum, tpre, tstride, total, tile_stop, tile_state, last_state};
    for (int guard = 0; guard < 128; guard++) {
      if (read_scalars()) return -1;
      if (!hs->gcp_pending) return 0;
      // the previous kernel halted the chain: clear it for the round kernels
      launch_resume(st, ds, 0);
      if (hs->gcp_need_select) {
        launch_gcp_select<T>(st, G, n, dX, dG, l, u, iwhere, near_lists(), ds, ca, n,
                             sel_count(), true);
        HIPCHK(hipGetLastError());
        continue;  // size the round with the count just selected
      }
      const long long K = (long long)hs->cand_count;
      max_sorted_round = std::max(max_sorted_round, K);
      launch_gcp_sortscan<T>(st, dG, W, ds, ca, cb, K, w);
      HIPCHK(hipGetLastError());
      if (read_scalars()) return -1;
      if (!hs->gcp_pending) return 0;
      if (hs->status != ST_GCP_ROUNDS) return -1;
    }
    return -1;
  }

  int run(int program, int start_step, int stop_step, double f) override {
    std::vector<int> steps = program_steps(
        program, !hs->has_after_iteration && start_step < 0 && stop_step == K_END);
    if (!hs->has_custom_stop && program != PROG_AFTER_CHECK && start_step != K_UPD &&
        stop_step != K_UPD) {
      // without a customised stopping test, the stopping tests already did
      // the memory update and the Cauchy controller (or ended the chain):
      // the update step would find nothing to do
      for (size_t k = 0; k + 1 < steps.size(); k++)
        if (steps[k] == K_UPD && steps[k + 1] == K_GCP && k > 0 && steps[k - 1] == K_STOP) {
          steps.erase(steps.begin() + k);
          break;
        }
    }
    size_t i = 0;
    if (start_step >= 0 && start_step != K_END) {
      while (i < steps.size() && steps[i] != start_step) i++;
      if (i == steps.size()) i = 0;
    }
    size_t j = i;
    while (j < steps.size() && steps[j] != stop_step) j++;
    if (enqueue_and_read(steps, i, j, f)) return -1;
    note_handoffs();
    if (hs->status == ST_GCP_ROUNDS) {
      hipEvent_t ea = (profile_from >= 0 && hs->iter >= profile_from &&
                       ((profile_mask >> K_END) & 1u))
                          ? take_event()
                          : nullptr;
      if (ea) hipEventRecord(ea, st);
      if (run_rounds()) return -1;
      if (ea) {
        hipEvent_t eb = take_event();
        if (eb) {
          hipEventRecord(eb, st);
          pending.push_back({K_END, ea, eb});
        } else {
          ev_pool.push_back(ea);
        }
      }
      // resume the tail from the Cauchy point pass, up to stop_step
      size_t p = 0;
      while (p < steps.size() && steps[p] != K_P4) p++;
      if (p < j && enqueue_and_read(steps, p, j, f)) return -1;
    }
    near_hint = hs->gcp_sel_near != 0;
    return hs->status;
  }

  int finalize_x() override {
    if (read_scalars()) return -1;
    if (cur() != 0) {
      HIPCHK(hipMemcpyAsync(X[0], X[1], (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    return 0;
  }

  // ---- named arrays (host copies) for inspection and state loading
  int get_array(const std::string& name, void* out, long long count) override {
    if (read_scalars()) return -1;
    const size_t nn = (size_t)n;
    auto dcopy = [&](const void* src, size_t bytes) -> int {
      HIPCHK(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      return 0;
    };
    const int c = cur();
    if (name == "x") return dcopy(X[c], nn * sizeof(T));
    if (name == "g") return dcopy(Gr[c], nn * sizeof(T));
    if (name == "xb") return dcopy(X[c ^ 1], nn * sizeof(T));
    if (name == "gb") return dcopy(Gr[c ^ 1], nn * sizeof(T));
    if (!owns_memory() && (name == "xcp" || name == "z" || name == "d" || name == "index" ||
                           name == "indx2"))
      return -3;  // inspection buffers are allocated on demand: not in a region
    if (name == "xcp" || name == "z" || name == "d") {
      // the Cauchy point, the subspace point or the line-search direction
      // as the passes compute them from the current state (d is stored only
      // once a trial moves off z)
      if (!insp && dmalloc(insp, nn)) return -1;
      if (name == "xcp")
        launch_cauchy_point<T>(st, G, n, dX, dG, l, u, iwhere, ds, insp);
      else if (name == "z")
        launch_subspace_point<T>(st, G, n, dX, dG, l, u, iwhere, dsub, ds, insp);
      else
        launch_direction<T>(st, G, n, dX, d, ds, insp);
      HIPCHK(hipGetLastError());
      return dcopy(insp, nn * sizeof(T));
    }
    if (name == "dsub") return dcopy(dsub, nn * sizeof(T));
    if (name == "iwbits") return dcopy(iwhere, nn);
    if (name == "iwhere") {
      std::vector<uint8_t> b(nn);
      HIPCHK(hipMemcpyAsync(b.data(), iwhere, nn, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      int* o = (int*)out;
      for (size_t i = 0; i < nn; i++) o[i] = iw_get(b[i]);
      return 0;
    }
    if (name == "ws" || name == "wy") {
      std::vector<T> w(wsize(nn));
      HIPCHK(hipMemcpyAsync(w.data(), W, w.size() * sizeof(T), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      const int off = name == "ws" ? kWPhys : 0;
      T* o = (T*)out;
      for (size_t i = 0; i < nn; i++)
        for (int j = 0; j < m; j++) o[i * m + j] = w[widx(i, off + ring_phys(hs, j))];
      return 0;
    }
    if (name == "index" || name == "indx2") {
      if (!index) {
        if (dmalloc(index, nn) || dmalloc(indx2, nn) ||
            dmalloc(exp_counts, 4 * (size_t)export_count_blocks(n) + 4) ||
            dmalloc(exp_totals, 4))
          return -1;
      }
      const int track = hs->iter > 0 && hs->cnstnd && !hs->shortcut;
      launch_export_index(st, iwhere, n, track, exp_counts, exp_totals, index, indx2);
      HIPCHK(hipGetLastError());
      return dcopy(name == "index" ? index : indx2, nn * sizeof(int));
    }
    double* o = (double*)out;
    auto small = [&](const double* a, int ld, int dim) {
      for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) o[i * dim + j] = a[i * ld + j];
      return 0;
    };
    if (name == "sy") return small(hs->sy, 8, m);
    if (name == "ss") return small(hs->ss, 8, m);
    if (name == "wt") return small(hs->wt, 8, m);
    if (name == "wn") return small(hs->wn, 16, 2 * m);
    if (name == "snd") return small(hs->wn1, 16, 2 * m);
    auto vec = [&](const double* a, long long cnt) {
      for (long long i = 0; i < cnt; i++) o[i] = a[i];
      return 0;
    };
    if (name == "p") return vec(hs->p, std::min<long long>(count, 16));
    if (name == "c") return vec(hs->c, std::min<long long>(count, 16));
    if (name == "mc") return vec(hs->mc, std::min<long long>(count, 16));
    if (name == "wv") return vec(hs->wv, std::min<long long>(count, 16));
    if (name == "pa") return vec(hs->pa, std::min<long long>(count, PA_COUNT));
    if (name == "dsave") return vec(hs->dsave, std::min<long long>(count, 13));
    if (name == "isave") {
      for (long long i = 0; i < std::min<long long>(count, 2); i++) o[i] = hs->isave[i];
      return 0;
    }
    return -2;
  }

  int set_array(const std::string& name, const void* in, long long count) override {
    if (read_scalars()) return -1;
    const size_t nn = (size_t)n;
    auto hcopy = [&](void* dst, size_t bytes) -> int {
      HIPCHK(hipMemcpyAsync(dst, in, bytes, hipMemcpyHostToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
      return 0;
    };
    const int c = cur();
    if (name == "x") return hcopy(X[c], nn * sizeof(T));
    if (name == "g") return hcopy(Gr[c], nn * sizeof(T));
    if (name == "xb") return hcopy(X[c ^ 1], nn * sizeof(T));
    if (name == "gb") return hcopy(Gr[c ^ 1], nn * sizeof(T));
    if (name == "iwbits") return hcopy(iwhere, nn);
    if (name == "ws" || name == "wy") {
      std::vector<T> w(wsize(nn));
      HIPCHK(hipMemcpyAsync(w.data(), W, w.size() * sizeof(T), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      const int off = name == "ws" ? kWPhys : 0;
      const T* src = (const T*)in;
      for (size_t i = 0; i < nn; i++)
        for (int j = 0; j < m; j++) w[widx(i, off + ring_phys(hs, j))] = src[i * m + j];
      HIPCHK(hipMemcpyAsync(W, w.data(), w.size() * sizeof(T), hipMemcpyHostToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
      return 0;
    }
    const double* s = (const double*)in;
    auto small = [&](double* a, int ld, int dim) {
      for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) a[i * ld + j] = s[i * dim + j];
      return write_scalars();
    };
    if (name == "sy") return small(hs->sy, 8, m);
    if (name == "ss") return small(hs->ss, 8, m);
    if (name == "wt") return small(hs->wt, 8, m);
    if (name == "wn") return small(hs->wn, 16, 2 * m);
    if (name == "snd") return small(hs->wn1, 16, 2 * m);
    auto vec = [&](double* a, long long cnt) {
      for (long long i = 0; i < cnt; i++) a[i] = s[i];
      return write_scalars();
    };
    if (name == "c") return vec(hs->c, std::min<long long>(count, 16));
    if (name == "dsave") return vec(hs->dsave, std::min<long long>(count, 13));
    if (name == "isave") {
      for (long long i = 0; i < std::min<long long>(count, 2); i++) hs->isave[i] = (int)s[i];
      return write_scalars();
    }
    return -2;
  }
};

size_t engine_region_bytes(int dtype, long long n, int m) {
  if (dtype == 1) {
    Engine<double> e(n, m, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, true);
    return e.region_off;
  }
  Engine<float> e(n, m, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, true);
  return e.region_off;
}

static_assert(sizeof(Engine<double>) <= kEngineObjectBytes, "engine object storage");
static_assert(sizeof(Engine<float>) <= kEngineObjectBytes, "engine object storage");

EngineBase* make_engine_in(void* storage, int dtype, long long n, int m, void* x,
                           const int* nbd, const void* l, const void* u, hipStream_t stream,
                           void* region, size_t region_bytes) {
  EngineBase* e;
  bool ok;
  if (dtype == 1) {
    auto* p = new (storage) Engine<double>(n, m, (double*)x, nbd, (const double*)l,
                                           (const double*)u, stream, (unsigned char*)region,
                                           region_bytes);
    ok = p->ok;
    e = p;
  } else {
    auto* p = new (storage) Engine<float>(n, m, (float*)x, nbd, (const float*)l,
                                          (const float*)u, stream, (unsigned char*)region,
                                          region_bytes);
    ok = p->ok;
    e = p;
  }
  if (!ok) {
    e->~EngineBase();
    return nullptr;
  }
  return e;
}

EngineBase* make_engine(int dtype, long long n, int m, void* x, const int* nbd, const void* l,
                        const void* u, hipStream_t stream) {
  EngineBase* e;
  if (dtype == 1)
    e = new Engine<double>(n, m, (double*)x, nbd, (const double*)l, (const double*)u, stream);
  else
    e = new Engine<float>(n, m, (float*)x, nbd, (const float*)l, (const float*)u, stream);
  bool ok = dtype == 1 ? static_cast<Engine<double>*>(e)->ok : static_cast<Engine<float>*>(e)->ok;
  if (!ok) {
    delete e;
    return nullptr;
  }
  return e;
}

}  // namespace lbfgsb
