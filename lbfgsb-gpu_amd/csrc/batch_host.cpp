// Host side of the batched form: an arena of problem slots in HBM and the
// launches of the persistent per-problem kernel (batch.hip).
//
// All per-problem arrays live in slices of arena-wide arrays, allocated once
// (first fit over 64-element-aligned ranges, so a problem's W block starts at
// kWCols x its offset).  A slot holds one problem from admission to removal;
// `run` launches the kernel over every unfinished slot, largest first, either
// to completion or for an iteration budget (the serving loop resumes the
// suspended problems in its next launch).  With a segment length set, a
// problem of n elements is solved by a gang of min(max_gang, ceil(n /
// segment)) workgroups; the gang size depends on n and the configuration
// only, so a problem's result does not depend on what else is in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "lbfgsb_amd.h"
#include "batch.hpp"
#include "common.hpp"
#include "generate.hpp"

using namespace lbfgsb;

namespace {

struct Slot {
  bool used = false;
  bool done = false;
  long long off = 0, n = 0, npad = 0;
};

struct BatchBase {
  virtual ~BatchBase() {}
  virtual int add(long long n, const void* x0, const int* nbd, const void* l, const void* u,
                  int kind, long long nx, long long ny, double h, const void* b, const void* c,
                  const void* a0, const void* a1, const void* a2, double eps_g, double eps_f,
                  double eps_x, int max_iteration) = 0;
  virtual int add_generated(const GenSpec& sp, double eps_g, double eps_f, double eps_x,
                            int max_iteration) = 0;
  virtual int run(int threads, int blocks, int iter_budget) = 0;
  virtual int result(int slot, double* f, int* iter, int* info, double* residual_g,
                     int* done) = 0;
  virtual int get_x(int slot, void* out) = 0;
  virtual int remove(int slot) = 0;
  virtual long long free_elems() const = 0;
  virtual int results_all(double* f, int* iter, int* info, int* done) = 0;
  virtual int clocks(long long* start, long long* end) = 0;
  virtual int configure(long long segment, int max_gang) = 0;
  virtual void set_graph(bool on) = 0;
  virtual int gang_of(int slot) const = 0;
  // the slot's iterate in the arena (device), its size, the bulk type
  virtual void* slot_x(int slot) = 0;
  virtual long long slot_n(int slot) const = 0;
  virtual int dtype() const = 0;
  virtual hipStream_t stream() const = 0;
  bool ok = false;
};

#define BCHK(expr)                    \
  do {                                \
    if ((expr) != hipSuccess) return -2; \
  } while (0)

template <typename T>
struct Arena final : BatchBase {
  hipStream_t st;
  int m;
  long long cap;  // elements, multiple of 64
  int maxp;
  T *x = nullptr, *xb = nullptr, *g = nullptr, *gb = nullptr, *l = nullptr, *u = nullptr;
  T *W = nullptr, *d = nullptr, *dsub = nullptr;
  T *ob = nullptr, *oc = nullptr, *oa0 = nullptr, *oa1 = nullptr, *oa2 = nullptr;
  int* nbd = nullptr;
  uint8_t* iwhere = nullptr;
  Cand *ca = nullptr, *cb = nullptr;
  DevScalars* dsa = nullptr;
  BatchResult* res = nullptr;
  BatchProb* probs = nullptr;
  BatchItem* items = nullptr;
  int* count = nullptr;
  GangCtl* gang = nullptr;
  double* gpart = nullptr;
  unsigned* next = nullptr;
  long long max_items = 0;
  long long segment = 0;  // 0: one workgroup per problem
  int max_gang = 1;
  // launches replayed from a captured graph (clear kernel + both batch
  // kernels at fixed grids), re-captured when the shape changes
  bool use_graph = false;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_threads = 0, g_blocks = 0, g_budget = -1;
  std::vector<Slot> slots;
  std::vector<BatchProb> hprobs;
  std::vector<BatchResult> hres;
  std::map<long long, long long> freel;  // free ranges: offset -> length

  Arena(long long capacity, int max_problems, int m_, hipStream_t s) : st(s), m(m_) {
    cap = ((capacity + 63) / 64) * 64;
    maxp = max_problems;
    ok = alloc() == 0;
    slots.resize(maxp);
    hprobs.resize(maxp);
    hres.resize(maxp);
    std::memset(hprobs.data(), 0, sizeof(BatchProb) * hprobs.size());
    std::memset(hres.data(), 0, sizeof(BatchResult) * hres.size());
    if (cap > 0) freel[0] = cap;
  }
  template <typename P>
  int dm(P*& p, size_t count) {
    BCHK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(P)));
    return 0;
  }
  int alloc() {
    const size_t c = (size_t)cap;
    T** vec[] = {&x, &xb, &g, &gb, &l, &u, &d, &dsub, &ob, &oc, &oa0, &oa1, &oa2};
    for (T** p : vec)
      if (dm(*p, c)) return -1;
    if (dm(W, c * kWCols) || dm(nbd, c) || dm(iwhere, c) || dm(ca, c) || dm(cb, c)) return -1;
    // queue entries: one per problem plus one per extra gang member (each
    // member takes at least kMinSegment elements)
    max_items = (long long)maxp + cap / kMinSegment + 1;
    if (dm(dsa, maxp) || dm(res, maxp) || dm(probs, maxp) || dm(items, (size_t)max_items) ||
        dm(count, 2) || dm(gang, maxp) || dm(gpart, (size_t)max_items * 2 * kGangPart) ||
        dm(next, 2))
      return -1;
    BCHK(hipMemsetAsync(W, 0, c * kWCols * sizeof(T), st));
    BCHK(hipStreamSynchronize(st));
    return 0;
  }
  ~Arena() override {
    hipStreamSynchronize(st);
    if (gexec) hipGraphExecDestroy(gexec);
    if (cap_stream) hipStreamDestroy(cap_stream);
    void* ptrs[] = {x,   xb,   g,      gb, l,  u,   W,   d,     dsub, ob,
                    oc,  oa0,  oa1,    oa2, nbd, iwhere, ca, cb, dsa, res, probs, items,
                    count, gang, gpart, next};
    for (void* p : ptrs)
      if (p) hipFree(p);
  }

  long long free_elems() const override {
    long long s = 0;
    for (const auto& kv : freel) s += kv.second;
    return s;
  }
  long long take_range(long long len) {
    for (auto it = freel.begin(); it != freel.end(); ++it) {
      if (it->second >= len) {
        const long long off = it->first, rest = it->second - len;
        freel.erase(it);
        if (rest > 0) freel[off + len] = rest;
        return off;
      }
    }
    return -1;
  }
  void give_range(long long off, long long len) {
    auto it = freel.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != freel.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      freel.erase(nx);
    }
    if (it != freel.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        freel.erase(it);
      }
    }
  }

  // errclb's scalar checks (lbfgsbcpu.cpp:879-890), the option checks and
  // the objective's shape
  int check(long long n, int kind, long long nx, long long ny, double eps_g, double eps_f,
            double eps_x, int max_iteration) const {
    if (n <= 0 || m <= 0 || m > n || eps_f < 0 || eps_g < 0 || eps_x < 0 || max_iteration < 0)
      return -1;
    if (kind < OBJK_GRID_QUAD || kind > OBJK_PENTA) return -1;
    if ((kind == OBJK_GRID_QUAD || kind == OBJK_GRID_COS) && (nx <= 0 || ny <= 0 || nx * ny != n))
      return -1;
    if (n >= (1LL << 31)) return -1;  // candidate indices are 32-bit
    return 0;
  }

  // A free slot and a range of npad elements, or -3.
  int claim(long long n) {
    int slot = -1;
    for (int k = 0; k < maxp; k++)
      if (!slots[k].used) {
        slot = k;
        break;
      }
    if (slot < 0) return -3;
    const long long npad = ((n + 63) / 64) * 64;
    const long long off = take_range(npad);
    if (off < 0) return -3;
    Slot& s = slots[slot];
    s.used = true;
    s.done = false;
    s.off = off;
    s.n = n;
    s.npad = npad;
    return slot;
  }

  // Objective descriptor, zeroed memory pairs and the starting control
  // state of a filled slot; waits for the slot's uploads.
  int init_slot(int slot, int kind, long long nx, long long ny, double h, double eps_g,
                double eps_f, double eps_x, int max_iteration) {
    const Slot& sl = slots[slot];
    const long long off = sl.off, n = sl.n;
    // the memory slots of a reused range may hold another problem's pairs:
    // they are multiplied by zero coefficients only, but keep them finite
    BCHK(hipMemsetAsync(W + off * kWCols, 0, (size_t)sl.npad * kWCols * sizeof(T), st));
    BatchProb& p = hprobs[slot];
    p.off = off;
    p.n = n;
    p.woff = off * kWCols;
    p.obj.kind = kind;
    p.obj.nx = nx;
    p.obj.ny = ny;
    p.obj.h = h;
    p.obj.b = ob + off;
    p.obj.c = oc + off;
    p.obj.a0 = oa0 + off;
    p.obj.a1 = oa1 + off;
    p.obj.a2 = oa2 + off;
    p.gang = 1;
    // the state the host driver starts from (capi.cpp: drive)
    DevScalars ds;
    std::memset(&ds, 0, sizeof(ds));
    ds.n = n;
    ds.m = m;
    ds.theta = 1.0;
    ds.nfree = (int)n;
    ds.eps_g = eps_g;
    ds.eps_f = eps_f;
    ds.eps_x = eps_x;
    ds.max_iteration = max_iteration;
    ds.ibd = -1;
    BCHK(hipMemcpyAsync(dsa + slot, &ds, sizeof(ds), hipMemcpyHostToDevice, st));
    std::memset(&hres[slot], 0, sizeof(BatchResult));
    BCHK(hipStreamSynchronize(st));
    return slot;
  }

  int add(long long n, const void* x0, const int* nbd_in, const void* l_in, const void* u_in,
          int kind, long long nx, long long ny, double h, const void* b, const void* c,
          const void* a0, const void* a1, const void* a2, double eps_g, double eps_f,
          double eps_x, int max_iteration) override {
    if (check(n, kind, nx, ny, eps_g, eps_f, eps_x, max_iteration) != 0) return -1;
    const int slot = claim(n);
    if (slot < 0) return slot;
    const long long off = slots[slot].off;
    const size_t bytes = (size_t)n * sizeof(T);
    auto cp = [&](T* dst, const void* src) -> int {
      if (src && hipMemcpyAsync(dst + off, src, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return -2;
      return 0;
    };
    if (cp(x, x0) || cp(l, l_in) || cp(u, u_in) || cp(ob, b) || cp(oc, c) || cp(oa0, a0) ||
        cp(oa1, a1) || cp(oa2, a2) ||
        hipMemcpyAsync(nbd + off, nbd_in, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, st) !=
            hipSuccess) {
      remove(slot);
      return -2;
    }
    const int r = init_slot(slot, kind, nx, ny, h, eps_g, eps_f, eps_x, max_iteration);
    if (r < 0) remove(slot);
    return r;
  }

  int add_generated(const GenSpec& sp, double eps_g, double eps_f, double eps_x,
                    int max_iteration) override {
    if (gen_check(sp) != 0 ||
        check(sp.n, sp.kind, sp.nx, sp.ny, eps_g, eps_f, eps_x, max_iteration) != 0)
      return -1;
    const int slot = claim(sp.n);
    if (slot < 0) return slot;
    const long long off = slots[slot].off;
    if (launch_generate<T>(st, sp, x + off, nbd + off, l + off, u + off, ob + off, oc + off,
                           oa0 + off, oa1 + off, oa2 + off) != 0) {
      remove(slot);
      return -2;
    }
    const bool grid = sp.kind == OBJK_GRID_QUAD || sp.kind == OBJK_GRID_COS;
    const double h = grid ? 1.0 / (double)(sp.nx + 1) : 0.0;
    const int r = init_slot(slot, sp.kind, sp.nx, sp.ny, h, eps_g, eps_f, eps_x, max_iteration);
    if (r < 0) remove(slot);
    return r;
  }

  int configure(long long seg, int mg) override {
    if (seg < 0 || (seg > 0 && seg < kMinSegment) || mg < 1) return -1;
    segment = seg;
    max_gang = mg;
    return 0;
  }
  int gang_size(long long n) const {
    if (segment <= 0) return 1;
    const long long s = (n + segment - 1) / segment;
    return (int)std::max<long long>(1, std::min<long long>(s, max_gang));
  }
  void set_graph(bool on) override { use_graph = on; }
  int gang_of(int slot) const override {
    if (slot < 0 || slot >= maxp || !slots[slot].used) return -1;
    return gang_size(slots[slot].n);
  }

  int run(int threads, int blocks, int iter_budget) override {
    std::vector<int> ord;
    for (int k = 0; k < maxp; k++)
      if (slots[k].used && !slots[k].done) ord.push_back(k);
    if (ord.empty()) return 0;
    // largest problems first: the gangs start together and the tail of the
    // persistent queue is short work
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return slots[a].n > slots[b].n; });
    if (threads != 256 && threads != 512) threads = 256;
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // every workgroup of the launch is resident at once (gang members wait
    // for each other)
    const int per_cu = batch_blocks_per_cu<T>(threads);
    if (per_cu <= 0) return -2;
    const int resident = std::max(1, cus) * per_cu;
    if (blocks <= 0) blocks = std::max(1, cus) * (threads == 256 ? 2 : 1);
    blocks = std::min(blocks, resident);
    // gang entries first (members consecutive), then the one-workgroup ones
    std::vector<BatchItem> q, solo;
    for (int k : ord) {
      const int S = gang_size(slots[k].n);
      if (S > blocks) return -4;  // the gang cannot be resident at once
      hprobs[k].gang = S;
      if (S > 1)
        for (int j = 0; j < S; j++) q.push_back(BatchItem{k, j});
      else
        solo.push_back(BatchItem{k, 0});
    }
    const int cnt[2] = {(int)q.size(), (int)solo.size()};
    q.insert(q.end(), solo.begin(), solo.end());
    if ((long long)q.size() > max_items) return -2;
    BCHK(hipMemcpyAsync(probs, hprobs.data(), sizeof(BatchProb) * maxp, hipMemcpyHostToDevice,
                        st));
    BCHK(hipMemcpyAsync(items, q.data(), sizeof(BatchItem) * q.size(), hipMemcpyHostToDevice,
                        st));
    BCHK(hipMemcpyAsync(count, cnt, sizeof(cnt), hipMemcpyHostToDevice, st));
    BatchBufs<T> B{x, xb, g, gb, l, u, nbd, iwhere, W, d, dsub, ca, cb};
    if (use_graph) {
      // the queue sizes are read from `count` on the device: one graph
      // serves every launch of this shape (workgroups beyond a queue's
      // length find it empty and exit)
      if (!gexec || g_threads != threads || g_blocks != blocks || g_budget != iter_budget) {
        if (gexec) {
          BCHK(hipGraphExecDestroy(gexec));
          gexec = nullptr;
        }
        if (!cap_stream) BCHK(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
        hipGraph_t gr = nullptr;
        BCHK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeRelaxed));
        launch_batch_clear(cap_stream, next, gang, maxp);
        const int lr = launch_batch<T>(cap_stream, threads, blocks, blocks, probs, items, count,
                                       next, gang, gpart, B, dsa, res, iter_budget);
        BCHK(hipStreamEndCapture(cap_stream, &gr));
        if (lr != 0) {
          hipGraphDestroy(gr);
          return -1;
        }
        const hipError_t ie = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (ie != hipSuccess) {
          gexec = nullptr;
          return -2;
        }
        g_threads = threads;
        g_blocks = blocks;
        g_budget = iter_budget;
      }
      BCHK(hipGraphLaunch(gexec, st));
    } else {
      launch_batch_clear(st, next, gang, maxp);
      if (launch_batch<T>(st, threads, std::min(blocks, cnt[0]), std::min(blocks, cnt[1]), probs,
                          items, count, next, gang, gpart, B, dsa, res, iter_budget) != 0)
        return -1;
    }
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(hres.data(), res, sizeof(BatchResult) * maxp, hipMemcpyDeviceToHost,
                        st));
    BCHK(hipStreamSynchronize(st));
    int left = 0;
    for (int k : ord) {
      slots[k].done = hres[k].done != 0;
      if (!slots[k].done) left++;
    }
    return left;
  }

  int result(int slot, double* f, int* iter, int* info, double* residual_g, int* done) override {
    if (slot < 0 || slot >= maxp || !slots[slot].used) return -1;
    const BatchResult& r = hres[slot];
    if (f) *f = r.f;
    if (iter) *iter = r.iter;
    if (info) *info = r.info;
    if (residual_g) *residual_g = r.residual_g;
    if (done) *done = r.done;
    return 0;
  }
  int results_all(double* f, int* iter, int* info, int* done) override {
    for (int k = 0; k < maxp; k++) {
      const BatchResult& r = hres[k];
      const bool u = slots[k].used;
      if (f) f[k] = u ? r.f : 0.0;
      if (iter) iter[k] = u ? r.iter : 0;
      if (info) info[k] = u ? r.info : 0;
      if (done) done[k] = u ? r.done : 0;
    }
    return 0;
  }
  int clocks(long long* start, long long* end) override {
    for (int k = 0; k < maxp; k++) {
      const bool u = slots[k].used;
      if (start) start[k] = u ? hres[k].clk_start : 0;
      if (end) end[k] = u && hres[k].done ? hres[k].clk_end : 0;
    }
    return 0;
  }
  void* slot_x(int slot) override {
    if (slot < 0 || slot >= maxp || !slots[slot].used) return nullptr;
    return x + slots[slot].off;
  }
  long long slot_n(int slot) const override {
    if (slot < 0 || slot >= maxp || !slots[slot].used) return -1;
    return slots[slot].n;
  }
  int dtype() const override { return sizeof(T) == 8 ? 1 : 0; }
  hipStream_t stream() const override { return st; }
  int get_x(int slot, void* out) override {
    if (slot < 0 || slot >= maxp || !slots[slot].used) return -1;
    BCHK(hipMemcpyAsync(out, x + slots[slot].off, (size_t)slots[slot].n * sizeof(T),
                        hipMemcpyDeviceToDevice, st));
    BCHK(hipStreamSynchronize(st));
    return 0;
  }
  int remove(int slot) override {
    if (slot < 0 || slot >= maxp || !slots[slot].used) return -1;
    give_range(slots[slot].off, slots[slot].npad);
    slots[slot] = Slot();
    return 0;
  }
};

}  // namespace

struct lbfgsb_batch {
  BatchBase* b;
};

namespace {

// Device bytes an arena of `capacity` elements and max_problems slots
// allocates (the allocations of Arena::alloc).
long long arena_bytes(int dtype, long long capacity, int max_problems) {
  const long long s = dtype == 1 ? 8 : 4;
  const long long cap = ((capacity + 63) / 64) * 64;
  const long long items = (long long)max_problems + cap / kMinSegment + 1;
  long long b = cap * (13 * s + kWCols * s + (long long)sizeof(int) + 1 + 2 * (long long)sizeof(Cand));
  b += (long long)max_problems *
       (long long)(sizeof(DevScalars) + sizeof(BatchResult) + sizeof(BatchProb) + sizeof(GangCtl));
  b += items * ((long long)sizeof(BatchItem) + 2LL * kGangPart * (long long)sizeof(double));
  b += 2 * (long long)sizeof(int) + 2 * (long long)sizeof(unsigned);
  return b;
}

}  // namespace

extern "C" {

lbfgsb_batch* lbfgsb_batch_create(int dtype, long long capacity, int max_problems, int m,
                                  void* stream) {
  if (capacity <= 0 || max_problems <= 0 || m <= 0) return nullptr;
  const int mm = m < 8 ? m : 8;
  BatchBase* b = nullptr;
  if (dtype == 0)
    b = new Arena<float>(capacity, max_problems, mm, (hipStream_t)stream);
  else
    b = new Arena<double>(capacity, max_problems, mm, (hipStream_t)stream);
  if (!b->ok) {
    delete b;
    return nullptr;
  }
  return new lbfgsb_batch{b};
}

void lbfgsb_batch_destroy(lbfgsb_batch* b) {
  if (!b) return;
  delete b->b;
  delete b;
}

int lbfgsb_batch_add(lbfgsb_batch* b, long long n, const void* x0, const int* nbd, const void* l,
                     const void* u, int kind, long long nx, long long ny, double h, const void* bv,
                     const void* c, const void* a0, const void* a1, const void* a2, double eps_g,
                     double eps_f, double eps_x, int max_iteration) {
  if (!b) return -1;
  return b->b->add(n, x0, nbd, l, u, kind, nx, ny, h, bv, c, a0, a1, a2, eps_g, eps_f, eps_x,
                   max_iteration);
}

int lbfgsb_batch_run(lbfgsb_batch* b, int threads, int blocks, int iter_budget) {
  if (!b) return -1;
  return b->b->run(threads, blocks, iter_budget);
}

int lbfgsb_batch_result(lbfgsb_batch* b, int slot, double* f, int* iter, int* info,
                        double* residual_g, int* done) {
  if (!b) return -1;
  return b->b->result(slot, f, iter, info, residual_g, done);
}

int lbfgsb_batch_results(lbfgsb_batch* b, double* f, int* iter, int* info, int* done) {
  if (!b) return -1;
  return b->b->results_all(f, iter, info, done);
}

int lbfgsb_batch_clocks(lbfgsb_batch* b, long long* start, long long* end,
                        double* ticks_per_second) {
  if (!b) return -1;
  if (ticks_per_second) {
    int dev = 0, khz = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    *ticks_per_second = 1e3 * (double)khz;
  }
  return b->b->clocks(start, end);
}

int lbfgsb_batch_get_x(lbfgsb_batch* b, int slot, void* out) {
  if (!b) return -1;
  return b->b->get_x(slot, out);
}

int lbfgsb_batch_warm_start(lbfgsb_batch* dst, int dst_slot, lbfgsb_batch* src,
                            int src_slot) {
  if (!dst || !src) return -1;
  void* xd = dst->b->slot_x(dst_slot);
  const void* xs = src->b->slot_x(src_slot);
  const long long n = dst->b->slot_n(dst_slot);
  if (!xd || !xs || n != src->b->slot_n(src_slot)) return -1;
  // the source's iterate is complete on its stream before the copy reads it
  if (hipStreamSynchronize(src->b->stream()) != hipSuccess) return -2;
  hipStream_t st = dst->b->stream();
  if (launch_convert_copy(st, xs, src->b->dtype(), xd, dst->b->dtype(), n) != 0) return -2;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -2;
}

int lbfgsb_batch_remove(lbfgsb_batch* b, int slot) {
  if (!b) return -1;
  return b->b->remove(slot);
}

long long lbfgsb_batch_free_elems(lbfgsb_batch* b) { return b ? b->b->free_elems() : -1; }

long long lbfgsb_batch_bytes(int dtype, long long capacity, int max_problems) {
  if (capacity <= 0 || max_problems <= 0) return -1;
  return arena_bytes(dtype, capacity, max_problems);
}

int lbfgsb_batch_configure(lbfgsb_batch* b, long long segment, int max_gang) {
  if (!b) return -1;
  return b->b->configure(segment, max_gang);
}

int lbfgsb_batch_gang(lbfgsb_batch* b, int slot) {
  if (!b) return -1;
  return b->b->gang_of(slot);
}

int lbfgsb_batch_use_graph(lbfgsb_batch* b, int on) {
  if (!b) return -1;
  b->b->set_graph(on != 0);
  return 0;
}

int lbfgsb_batch_add_generated(lbfgsb_batch* b, int kind, long long n, long long nx, long long ny,
                               unsigned long long seed, double eps_g, double eps_f, double eps_x,
                               int max_iteration) {
  if (!b) return -1;
  GenSpec sp{kind, 0, n, nx, ny, seed};
  return b->b->add_generated(sp, eps_g, eps_f, eps_x, max_iteration);
}

int lbfgsb_generate(int dtype, int kind, long long n, long long nx, long long ny,
                    unsigned long long seed, void* x0, int* nbd, void* l, void* u, void* bv,
                    void* c, void* a0, void* a1, void* a2, void* stream) {
  GenSpec sp{kind, 0, n, nx, ny, seed};
  hipStream_t st = (hipStream_t)stream;
  int r;
  if (dtype == 1)
    r = launch_generate<double>(st, sp, (double*)x0, nbd, (double*)l, (double*)u, (double*)bv,
                                (double*)c, (double*)a0, (double*)a1, (double*)a2);
  else
    r = launch_generate<float>(st, sp, (float*)x0, nbd, (float*)l, (float*)u, (float*)bv,
                               (float*)c, (float*)a0, (float*)a1, (float*)a2);
  if (r != 0) return r;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -2;
}

}  // extern "C"
