// Serving loop over the batched form (lbfgsb_serve_stream in
// include/lbfgsb_amd.h): continuous batching of solve requests with
// admission by an HBM budget.
//
// One host thread follows the arrival clock.  Arrived requests wait in a
// first-come-first-served queue per precision; the head of a queue is
// admitted as soon as its arena has a free slot and a free range (its
// problem is generated straight into the slot), so the budget holds by
// construction: the arenas are the only device memory the server holds and
// they are sized to the budget.  Each turn of the loop launches the batch
// kernels of each arena once with an iteration budget (`quantum`), which
// advances every resident problem -- old and newly admitted alike -- and
// keeps the control state of the unfinished ones resident for the next
// launch; finished problems are answered and their slots freed at once.
// Both arenas launch on the one stream: the gang kernel relies on having
// the device's workgroup slots to itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <thread>
#include <vector>

#include "lbfgsb_amd.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Largest capacity (multiple of 64 elements) whose arena fits `bytes`.
long long capacity_for(int dtype, double bytes, int slots) {
  long long lo = 0, hi = 64;
  while (lbfgsb_batch_bytes(dtype, hi, slots) <= bytes && hi < (1LL << 40)) hi *= 2;
  // invariant: lo fits (or is 0), hi does not
  while (hi - lo > 64) {
    const long long mid = ((lo + hi) / 2 / 64) * 64;
    if (mid <= lo) break;
    if (lbfgsb_batch_bytes(dtype, mid, slots) <= bytes)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct Lane {
  lbfgsb_batch* b = nullptr;
  long long cap = 0;
  int slots = 0;
  std::vector<int> req_of;  // slot -> request index (-1 free)
  std::deque<int> waiting;
  int resident = 0;
  std::vector<double> f;
  std::vector<int> iter, info, done;
};

}  // namespace

extern "C" int lbfgsb_serve_stream(const lbfgsb_serve_config* cfg, const lbfgsb_request* reqs,
                                   int count, lbfgsb_response* out, lbfgsb_serve_stats* stats,
                                   void* stream) {
  if (!cfg || count < 0 || (count > 0 && (!reqs || !out))) return -1;
  if (!(cfg->hbm_budget_bytes > 0) || cfg->fp32_share < 0 || cfg->fp32_share > 1 ||
      cfg->max_slots <= 0 || cfg->quantum <= 0)
    return -1;
  const int threads = cfg->threads > 0 ? cfg->threads : 256;
  const int blocks = cfg->blocks > 0 ? cfg->blocks : 0;
  Lane lane[2];
  double held = 0;
  for (int d = 0; d < 2; d++) {
    const double share = d == 0 ? cfg->fp32_share : 1.0 - cfg->fp32_share;
    Lane& L = lane[d];
    L.slots = cfg->max_slots;
    L.cap = share > 0 ? capacity_for(d, share * cfg->hbm_budget_bytes, L.slots) : 0;
    if (L.cap < 64) {
      L.cap = 0;
      continue;
    }
    L.b = lbfgsb_batch_create(d, L.cap, L.slots, 8, stream);
    if (!L.b) {
      for (int e = 0; e < d; e++)
        if (lane[e].b) lbfgsb_batch_destroy(lane[e].b);
      return -2;
    }
    held += (double)lbfgsb_batch_bytes(d, L.cap, L.slots);
    if (lbfgsb_batch_configure(L.b, cfg->segment, cfg->segment > 0 ? cfg->max_gang : 1) != 0) {
      for (int e = 0; e <= d; e++)
        if (lane[e].b) lbfgsb_batch_destroy(lane[e].b);
      return -1;
    }
    lbfgsb_batch_use_graph(L.b, cfg->use_graph);
    L.req_of.assign(L.slots, -1);
    L.f.resize(L.slots);
    L.iter.resize(L.slots);
    L.info.resize(L.slots);
    L.done.resize(L.slots);
  }

  std::vector<int> ord(count);
  for (int i = 0; i < count; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });

  lbfgsb_serve_stats st{};
  st.arena_bytes = held;
  int next = 0, answered = 0, rc = 0;
  const double t0 = now_s();
  // fp32 answers waiting for their FP64 polish: the fp32 slot is held until
  // the polish is admitted (its iterate is the polish's start)
  struct Held {
    int r, slot, iter;
  };
  std::deque<Held> polishq;
  std::vector<int> polished(count, 0), iter32(count, 0);
  const bool polish = cfg->polish != 0 && lane[1].b != nullptr;
  const double polish_scale = cfg->polish_eps_scale > 0 ? cfg->polish_eps_scale : 1.0;
  auto answer = [&](int r, double f, int it, int info, int gang, double t_admit) {
    lbfgsb_response& o = out[r];
    o.f = f;
    o.iter = it;
    o.info = info;
    o.gang = gang;
    o.dtype = reqs[r].dtype;
    o.t_arrival = reqs[r].arrival;
    o.t_admit = t_admit;
    o.t_done = now_s() - t0;
    o.polished = polished[r];
    o.iter_fp32 = iter32[r];
    answered++;
  };
  std::vector<double> admit_time(count, 0.0);
  std::vector<int> gang_of(count, 1);

  while (answered < count && rc == 0) {
    const double now = now_s() - t0;
    // arrivals
    while (next < count && reqs[ord[next]].arrival <= now) {
      const int r = ord[next++];
      const lbfgsb_request& q = reqs[r];
      const int d = q.dtype == 1 ? 1 : 0;
      const long long npad = ((q.n + 63) / 64) * 64;
      if (!lane[d].b || q.n <= 0 || npad > lane[d].cap)
        answer(r, 0.0, 0, -3, 0, now);
      else
        lane[d].waiting.push_back(r);
    }
    // polishes first (their requests arrived before any waiting one), then
    // arrivals, first come first served per precision
    while (!polishq.empty()) {
      const Held h = polishq.front();
      const lbfgsb_request& q = reqs[h.r];
      Lane& P = lane[1];
      // the polish has the iterations the fp32 phase left of the request's
      // cap, and the request's tolerance scaled by polish_eps_scale
      const int left = std::max(q.max_iteration - h.iter, 1);
      const double pe = q.eps * polish_scale;
      const int slot = lbfgsb_batch_add_generated(P.b, q.kind, q.n, q.nx, q.ny, q.seed, pe, pe,
                                                  pe, left);
      if (slot == -3) break;  // the fp64 arena is full: wait for completions
      polishq.pop_front();
      if (slot < 0 || lbfgsb_batch_warm_start(P.b, slot, lane[0].b, h.slot) != 0) {
        rc = -2;
        break;
      }
      lbfgsb_batch_remove(lane[0].b, h.slot);
      lane[0].req_of[h.slot] = -1;
      lane[0].resident--;
      P.req_of[slot] = h.r;
      P.resident++;
      polished[h.r] = 1;
      iter32[h.r] = h.iter;
      gang_of[h.r] = lbfgsb_batch_gang(P.b, slot);
    }
    for (int d = 0; d < 2; d++) {
      Lane& L = lane[d];
      while (!L.waiting.empty()) {
        const int r = L.waiting.front();
        const lbfgsb_request& q = reqs[r];
        const int slot = lbfgsb_batch_add_generated(L.b, q.kind, q.n, q.nx, q.ny, q.seed, q.eps,
                                                    q.eps, q.eps, q.max_iteration);
        if (slot == -3) break;  // no slot or range free: wait for completions
        L.waiting.pop_front();
        if (slot < 0) {
          if (slot == -2) rc = -2;
          answer(r, 0.0, 0, -1, 0, now_s() - t0);
          continue;
        }
        L.req_of[slot] = r;
        L.resident++;
        admit_time[r] = now_s() - t0;
        gang_of[r] = lbfgsb_batch_gang(L.b, slot);
      }
      st.max_waiting = std::max(st.max_waiting, (int)L.waiting.size());
    }
    st.max_resident = std::max(st.max_resident, lane[0].resident + lane[1].resident);
    if (lane[0].resident + lane[1].resident == 0) {
      if (next >= count) {
        // nothing resident and nothing to come: only requests that can never
        // be admitted remain
        for (int d = 0; d < 2; d++)
          while (!lane[d].waiting.empty()) {
            answer(lane[d].waiting.front(), 0.0, 0, -3, 0, now_s() - t0);
            lane[d].waiting.pop_front();
          }
        break;
      }
      const double wait = reqs[ord[next]].arrival - (now_s() - t0);
      if (wait > 0)
        std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
      continue;
    }
    // one quantum per arena, then the answers
    for (int d = 0; d < 2 && rc == 0; d++) {
      Lane& L = lane[d];
      if (L.resident == 0) continue;
      const int left = lbfgsb_batch_run(L.b, threads, blocks, cfg->quantum);
      st.launches++;
      if (left < 0) {
        rc = left == -4 ? -1 : -2;
        break;
      }
      lbfgsb_batch_results(L.b, L.f.data(), L.iter.data(), L.info.data(), L.done.data());
      for (int s = 0; s < L.slots; s++) {
        const int r = L.req_of[s];
        if (r < 0 || !L.done[s]) continue;
        if (d == 0 && polish && L.info[s] >= 0) {
          const long long npad = ((reqs[r].n + 63) / 64) * 64;
          if (npad <= lane[1].cap) {
            bool queued = false;
            for (const Held& h : polishq) queued = queued || h.slot == s;
            if (!queued) polishq.push_back({r, s, L.iter[s]});
            continue;  // answered after the polish; the slot stays held
          }
        }
        answer(r, L.f[s], L.iter[s] + iter32[r], L.info[s], gang_of[r], admit_time[r]);
        lbfgsb_batch_remove(L.b, s);
        L.req_of[s] = -1;
        L.resident--;
      }
    }
  }
  st.seconds = now_s() - t0;
  for (int d = 0; d < 2; d++)
    if (lane[d].b) lbfgsb_batch_destroy(lane[d].b);
  if (stats) *stats = st;
  return rc;
}
