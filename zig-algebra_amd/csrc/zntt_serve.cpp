// Serving loop for coset low-degree extensions: requests queue as they are
// submitted, are admitted against an HBM budget and an active-request limit,
// and run by continuous batching.  Every scheduler step holds the next pass
// of every admitted request, in a few persistent launches grouped by kernel
// (HipDevice::classify); a request admitted while others are on the device
// joins the next step.  The steps run in order on the server's own stream,
// with up to kDepth of them enqueued ahead, so the device never waits for
// the host between steps.
//
// The scheduling core (queue, admission, budget, steps, retirement,
// cancellation) talks to the device through the Device interface.  HipDevice
// does the real work; SimDevice does none, so the same core can be driven
// step by step from a host without a GPU to test admission and failure
// handling.
//
// No reference counterpart: the reference computes an LDE only inside its
// STARK example, by Horner's rule (examples/stark_prover.zig:117-129).  A
// request's result equals zntt_lde on the same coefficients: zero-padding to
// 2^(log_n + log_blowup), the coset scale shift^i and the forward transform
// with the standard primitive root, fused into the transform's passes (the
// first pass reads the 2^log_n coefficients and treats the rest as zero).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "ctx_internal.hpp"

using namespace zntt_detail;

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kDepth = 2;  // steps enqueued ahead of the device
constexpr uint64_t kDefaultBudget = 64ull << 30;
constexpr int kDefaultActive = 256;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Request {
  uint64_t ticket = 0;
  zntt_lde_request req = {};
  Elem shift = {};
  int L = 0;
  int npass = 0, next_pass = 0;
  zntt::PassPlan plan = {};
  size_t out_bytes = 0, scratch_bytes = 0, in_bytes = 0, bytes = 0;
  void *out = nullptr, *scratch = nullptr, *in = nullptr;
  bool own_out = true;  // out is server memory (not the caller's)
  TwiddleEntry *tw = nullptr;
  CosetEntry *coset = nullptr;
  int state = ZNTT_REQ_QUEUED;
  int status = ZNTT_OK;
  bool cancelled = false;
  bool resident = false;  // holds device memory
  Clock::time_point t_submit, t_admit, t_done;
};

struct StepItem {
  Request *r;
  int pass;
};

// What the scheduling core needs from a device.
struct Device {
  virtual ~Device() = default;
  // Allocates the request's buffers, resolves its tables, starts generating
  // its input.  OUT_OF_MEMORY means "retry when memory has been freed".
  virtual int prepare(Request *r) = 0;
  // Frees the request's buffers (the device is done with them).  The result
  // buffer is kept when `keep_out`.
  virtual void release(Request *r, bool keep_out) = 0;
  virtual void free_result(Request *r) = 0;
  // Enqueues one step; returns its id (>= 0) or -status.
  virtual long submit_step(const std::vector<StepItem> &items) = 0;
  // True when step `id` has completed (block: wait for it).  `busy_ms` gets
  // its device time.
  virtual int step_done(long id, bool block, bool *done, double *busy_ms) = 0;
};

}  // namespace

struct zntt_server {
  zntt_ctx *ctx = nullptr;
  zntt_server_config cfg = {};
  std::unique_ptr<Device> dev;
  std::mutex mu;
  std::condition_variable cv_work;  // scheduler: new submissions / stop
  std::condition_variable cv_done;  // waiters: request state changes
  std::deque<Request *> queue;
  std::vector<Request *> active;  // admitted, not yet retired
  struct Step {
    long id;
    std::vector<Request *> finishing;  // requests whose last pass is in it
  };
  std::deque<Step> inflight;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
  uint64_t next_ticket = 1;
  uint64_t in_use = 0;
  zntt_server_stats stats = {};
  int broken = ZNTT_OK;
  bool stop = false;
  std::thread thread;
  int tile_log = 11, kmax = 9;
  // Result reads of more than one ring slot: a page-locked ring and a stream
  // of their own, so a read neither takes the context's lock nor waits behind
  // the scheduler's steps (made at the first such read).
  std::mutex read_mu;
  zntt_detail::HostRing *read_ring = nullptr;
  hipStream_t read_stream = nullptr;
  bool read_ring_failed = false;
};

namespace {

// ---------------------------------------------------------------------------
// Scheduling core.  Caller holds s->mu.
// ---------------------------------------------------------------------------

void finish(zntt_server *s, Request *r, int state, int status) {
  r->state = state;
  r->status = status;
  r->t_done = Clock::now();
  if (state == ZNTT_REQ_DONE) {
    s->stats.completed++;
    s->stats.elements_done += (uint64_t)1 << r->L;
  } else if (state == ZNTT_REQ_FAILED) {
    s->stats.failed++;
  } else if (state == ZNTT_REQ_CANCELLED) {
    s->stats.cancelled++;
  }
  s->cv_done.notify_all();
}

void retire_request(zntt_server *s, Request *r) {
  const bool keep = r->req.keep_result && !r->cancelled;
  s->dev->release(r, keep);
  const uint64_t freed = keep ? r->bytes - r->out_bytes : r->bytes;
  s->in_use -= freed;
  if (!keep) r->resident = false;
  s->active.erase(std::find(s->active.begin(), s->active.end(), r));
  finish(s, r, r->cancelled ? ZNTT_REQ_CANCELLED : ZNTT_REQ_DONE, ZNTT_OK);
}

void fail_all(zntt_server *s, int status) {
  s->broken = status;
  for (Request *r : s->queue) finish(s, r, ZNTT_REQ_FAILED, status);
  s->queue.clear();
  for (Request *r : s->active) finish(s, r, ZNTT_REQ_FAILED, status);
  s->active.clear();
  s->inflight.clear();
}

// One scheduler iteration; returns true if anything changed.
bool iterate(zntt_server *s) {
  if (s->broken) return false;
  bool progress = false;
  // 1. Retire completed steps, oldest first.
  while (!s->inflight.empty()) {
    bool done = false;
    double busy = 0;
    int st = s->dev->step_done(s->inflight.front().id, false, &done, &busy);
    if (st) {
      fail_all(s, st);
      return true;
    }
    if (!done) break;
    s->stats.busy_ms += busy;
    for (Request *r : s->inflight.front().finishing) retire_request(s, r);
    s->inflight.pop_front();
    progress = true;
  }
  // 2. Admit from the head of the queue (first come, first served).
  while (!s->queue.empty() && (int)s->active.size() < s->cfg.max_active) {
    Request *r = s->queue.front();
    if (r->bytes > s->cfg.hbm_budget) {
      s->queue.pop_front();
      finish(s, r, ZNTT_REQ_FAILED, ZNTT_OUT_OF_MEMORY);
      progress = true;
      continue;
    }
    if (s->in_use + r->bytes > s->cfg.hbm_budget) break;  // wait for retirements
    const int st = s->dev->prepare(r);
    if (st == ZNTT_OUT_OF_MEMORY && !s->active.empty()) break;  // memory frees up later
    s->queue.pop_front();
    progress = true;
    if (st) {
      finish(s, r, ZNTT_REQ_FAILED, st);
      continue;
    }
    r->resident = true;
    s->in_use += r->bytes;
    s->stats.peak_bytes = std::max(s->stats.peak_bytes, s->in_use);
    s->stats.admitted++;
    r->state = ZNTT_REQ_RUNNING;
    r->t_admit = Clock::now();
    s->active.push_back(r);
    s->cv_done.notify_all();
  }
  // 3. Launch the next step: one pass of every admitted request with passes left.
  if ((int)s->inflight.size() < kDepth) {
    std::vector<StepItem> items;
    zntt_server::Step step;
    for (Request *r : s->active) {
      if (r->next_pass >= r->npass) continue;
      items.push_back(StepItem{r, r->next_pass});
      if (r->next_pass == r->npass - 1) step.finishing.push_back(r);
    }
    if (!items.empty()) {
      const long id = s->dev->submit_step(items);
      if (id < 0) {
        fail_all(s, (int)-id);
        return true;
      }
      for (auto &it : items) it.r->next_pass++;
      step.id = id;
      s->inflight.push_back(std::move(step));
      s->stats.steps++;
      progress = true;
    }
  }
  return progress;
}

void scheduler_main(zntt_server *s) {
  if (s->ctx) hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    const bool progress = iterate(s);
    if (s->stop && s->queue.empty() && s->active.empty()) break;
    if (progress) continue;
    if (!s->inflight.empty() && !s->broken) {
      if ((int)s->inflight.size() >= kDepth) {
        // No step can be launched before the oldest one completes.
        const long id = s->inflight.front().id;
        lk.unlock();
        bool done = false;
        double busy = 0;
        const int st = s->dev->step_done(id, true, &done, &busy);
        lk.lock();
        if (st) fail_all(s, st);
      } else {
        // A step slot is free: wake for a submission or poll the device.
        s->cv_work.wait_for(lk, std::chrono::microseconds(100));
      }
    } else {
      s->cv_work.wait(lk);
    }
  }
}

// ---------------------------------------------------------------------------
// Memory: a first-fit arena with coalescing free blocks.
// ---------------------------------------------------------------------------
struct Arena {
  uint8_t *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> length
  void reset(uint8_t *b, size_t n) {
    base = b;
    size = n;
    free_blocks.clear();
    if (n) free_blocks[0] = n;
  }
  void *alloc(size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1), 256);
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_blocks.erase(it);
      if (len > bytes) free_blocks[off + bytes] = len - bytes;
      return base + off;
    }
    return nullptr;
  }
  void release(void *p, size_t bytes) {
    if (!p) return;
    bytes = align_up(std::max<size_t>(bytes, 1), 256);
    size_t off = (uint8_t *)p - base;
    auto next = free_blocks.lower_bound(off);
    if (next != free_blocks.end() && next->first == off + bytes) {
      bytes += next->second;
      next = free_blocks.erase(next);
    }
    if (next != free_blocks.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        bytes += prev->second;
        free_blocks.erase(prev);
      }
    }
    free_blocks[off] = bytes;
  }
};

// Allocation through the caller's allocator or the server's arena.
struct Memory {
  const zntt_allocator *user = nullptr;
  Arena arena;
  int alloc(size_t bytes, void **out) {
    *out = nullptr;
    if (bytes == 0) return ZNTT_OK;
    if (user) {
      const int st = user->alloc(user->user, bytes, out);
      if (st == ZNTT_OK && !*out) return ZNTT_OUT_OF_MEMORY;
      return st;
    }
    *out = arena.alloc(bytes);
    return *out ? ZNTT_OK : ZNTT_OUT_OF_MEMORY;
  }
  void free(void *p, size_t bytes) {
    if (!p) return;
    if (user)
      user->free(user->user, p, bytes);
    else
      arena.release(p, bytes);
  }
  // Frees the request's buffers; returns 0 or the first allocation's status.
  int alloc_request(Request *r) {
    int st = ZNTT_OK;
    if (r->own_out)
      st = alloc(r->out_bytes, &r->out);
    else
      r->out = r->req.out;
    if (st == ZNTT_OK) st = alloc(r->scratch_bytes, &r->scratch);
    if (st == ZNTT_OK) st = alloc(r->in_bytes, &r->in);
    if (st) free_request(r, false);
    return st;
  }
  void free_request(Request *r, bool keep_out) {
    if (!keep_out) {
      if (r->own_out) free(r->out, r->out_bytes);
      r->out = nullptr;
    }
    free(r->scratch, r->scratch_bytes);
    free(r->in, r->in_bytes);
    r->scratch = r->in = nullptr;
  }
};

// ---------------------------------------------------------------------------
// Simulated device: bookkeeping only.
// ---------------------------------------------------------------------------
struct SimDevice : Device {
  Memory mem;
  uintptr_t next_fake = 0x100000;
  long next_id = 0;
  long completed_upto = -1;  // steps with id <= this are complete
  std::vector<long> submitted;
  explicit SimDevice(const zntt_allocator *a) {
    mem.user = a;
  }
  int prepare(Request *r) override {
    if (mem.user) return mem.alloc_request(r);
    // Fake, never dereferenced addresses.
    r->out = r->own_out ? (void *)next_fake : r->req.out;
    next_fake += align_up(r->out_bytes, 256);
    r->scratch = r->scratch_bytes ? (void *)next_fake : nullptr;
    next_fake += align_up(r->scratch_bytes, 256);
    r->in = r->in_bytes ? (void *)next_fake : nullptr;
    next_fake += align_up(r->in_bytes, 256);
    return ZNTT_OK;
  }
  void release(Request *r, bool keep_out) override {
    if (mem.user) {
      mem.free_request(r, keep_out);
    } else {
      if (!keep_out) r->out = nullptr;
      r->scratch = r->in = nullptr;
    }
  }
  void free_result(Request *r) override {
    if (mem.user && r->own_out) mem.free(r->out, r->out_bytes);
    r->out = nullptr;
  }
  long submit_step(const std::vector<StepItem> &) override {
    submitted.push_back(next_id);
    return next_id++;
  }
  int step_done(long id, bool, bool *done, double *busy_ms) override {
    *done = id <= completed_upto;
    *busy_ms = 0;
    return ZNTT_OK;
  }
};

// ---------------------------------------------------------------------------
// HIP device.
// ---------------------------------------------------------------------------
struct HipDevice : Device {
  zntt_ctx *ctx;
  int field;
  int tile_log, grid = 0;
  hipStream_t stream = nullptr;
  Memory mem;
  void *arena_mem = nullptr;
  // Per in-flight step: pinned staging, device copy, events.
  struct Slot {
    void *host = nullptr;
    void *dev = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    long id = -1;
  };
  Slot slots[kDepth + 1];
  size_t slot_bytes = 0;
  long next_id = 0;
  // Branch streams of a step's launches beyond the first, with the step's
  // fork event and one join event per branch stream.
  static constexpr int kBranches = 8;
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;

  HipDevice(zntt_ctx *c, int f, int tl) : ctx(c), field(f), tile_log(tl) {}

  int init(const zntt_server_config &cfg) {
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_status(e);
    e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (int k = 0; k < kBranches && e == hipSuccess; k++) {
      hipStream_t st = nullptr;
      hipEvent_t ev = nullptr;
      e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      if (e == hipSuccess) aux.push_back(st);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) joins.push_back(ev);
    }
    if (e != hipSuccess) return hip_status(e);
    e = zntt::configure_batch_lds();
    if (e != hipSuccess) return hip_status(e);
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    grid = std::max(1, cus * 2);
    if (cfg.allocator) {
      mem.user = cfg.allocator;
    } else {
      e = hipMalloc(&arena_mem, cfg.hbm_budget);
      if (e != hipSuccess) {
        hipGetLastError();
        arena_mem = nullptr;
        return hip_status(e);
      }
      mem.arena.reset((uint8_t *)arena_mem, cfg.hbm_budget);
    }
    // [counter | BatchXform x max_active | BatchEntry x max_active]
    slot_bytes = 256 + align_up(sizeof(zntt::BatchXform) * cfg.max_active, 256) +
                 sizeof(zntt::BatchEntry) * cfg.max_active;
    for (auto &sl : slots) {
      e = hipHostMalloc(&sl.host, slot_bytes, hipHostMallocDefault);
      if (e == hipSuccess) e = hipMalloc(&sl.dev, slot_bytes);
      if (e == hipSuccess) e = hipEventCreate(&sl.start);
      if (e == hipSuccess) e = hipEventCreate(&sl.end);
      if (e != hipSuccess) return hip_status(e);
    }
    return ZNTT_OK;
  }

  ~HipDevice() override {
    if (stream) hipStreamSynchronize(stream);
    for (auto &sl : slots) {
      if (sl.host) hipHostFree(sl.host);
      if (sl.dev) hipFree(sl.dev);
      if (sl.start) hipEventDestroy(sl.start);
      if (sl.end) hipEventDestroy(sl.end);
    }
    for (auto st : aux) hipStreamSynchronize(st);
    if (arena_mem) hipFree(arena_mem);
    for (auto st : aux) hipStreamDestroy(st);
    for (auto ev : joins) hipEventDestroy(ev);
    if (fork) hipEventDestroy(fork);
    if (stream) hipStreamDestroy(stream);
  }

  int prepare(Request *r) override {
    int st = mem.alloc_request(r);
    if (st) return st;
    {
      CallGuard g(ctx);
      st = power_table(ctx, field, f_primitive_root(field, r->L), r->L, ctx->stream, &r->tw);
      if (st == ZNTT_OK) {
        const uint32_t *cos = nullptr;
        st = coset_table(ctx, field, r->shift, (size_t)1 << r->req.log_n, ctx->stream, &cos);
        if (st == ZNTT_OK)
          for (auto &ce : ctx->cosets)
            if (ce.table == cos) r->coset = &ce;
      }
      if (st == ZNTT_OK) {
        r->tw->pins++;
        r->coset->pins++;
      } else {
        r->tw = nullptr;
        r->coset = nullptr;
      }
    }
    if (st == ZNTT_OK && r->in) {
      hipError_t e = zntt::launch_fill_counter(field, (uint32_t *)r->in, (size_t)1 << r->req.log_n, r->req.seed,
                                               r->req.stream_id, stream);
      if (e != hipSuccess) st = hip_status(e);
    }
    if (st) {
      // A table build that failed for lack of memory is not retried: the
      // context's cache has its own budget, separate from the server's.
      if (st == ZNTT_OUT_OF_MEMORY && !r->tw) st = ZNTT_DEVICE_ERROR;
      unpin(r);
      mem.free_request(r, false);
    }
    return st;
  }

  void unpin(Request *r) {
    if (!r->tw && !r->coset) return;
    CallGuard g(ctx);
    if (r->tw) r->tw->pins--;
    if (r->coset) r->coset->pins--;
    r->tw = nullptr;
    r->coset = nullptr;
  }

  void release(Request *r, bool keep_out) override {
    unpin(r);
    mem.free_request(r, keep_out);
  }

  void free_result(Request *r) override {
    if (r->own_out) mem.free(r->out, r->out_bytes);
    r->out = nullptr;
  }

  // A step's launches.  Each item (one pass of one request) goes to one of
  // three kernels: a request that fits one register-round tile (2^9 up to
  // the field's largest) to the whole-transform kernel, one launch per size;
  // a pass with a full tile to the register-round level kernel, one launch
  // per (pass length, parity of its first stage), so the passes of all large
  // requests that share those run in one launch; the rest (small requests)
  // to the generic level kernel.  A step's launches are independent: they
  // run as parallel branches of the server stream.
  enum Kind { kGeneric, kWhole, kLevelR8 };
  struct Group {
    Kind kind;
    int a, b;  // kWhole: log size; kLevelR8: pass length, parity
    std::vector<const StepItem *> items;
  };
  static constexpr int kMaxGroups = 32;  // counter pairs in a step's 256-byte header

  // Whether a launch holds the first pass of a request whose coefficients
  // are fewer than its evaluations (zero padding): the level kernel then
  // skips the padding's products.
  static bool pads_first_pass(const Group &g) {
    for (const StepItem *it : g.items)
      if (it->pass == 0 && (int)it->r->req.log_n < it->r->L) return true;
    return false;
  }

  static int s0_of(const Request *r, int p) {
    int s0 = 1;
    for (int q = 0; q < p; q++) s0 += r->plan.K[q];
    return s0;
  }

  std::vector<Group> classify(const std::vector<StepItem> &items) const {
    std::vector<Group> g;
    auto find = [&](Kind k, int a, int b) -> Group & {
      for (auto &x : g)
        if (x.kind == k && x.a == a && x.b == b) return x;
      g.push_back(Group{k, a, b, {}});
      return g.back();
    };
    for (const auto &it : items) {
      const Request *r = it.r;
      const int K = r->plan.K[it.pass], tl = K + r->plan.log_g[it.pass];
      if (r->npass == 1 && r->L >= zntt::kWholeMinLog && r->L <= zntt::whole_max_log(field))
        find(kWhole, r->L, 0).items.push_back(&it);
      else if (tl == tile_log && zntt::level_r8_ok(field, K, tl))
        find(kLevelR8, K, s0_of(r, it.pass) & 1).items.push_back(&it);
      else
        find(kGeneric, 0, 0).items.push_back(&it);
    }
    // Largest launches first, so that the smaller ones fill their tails.
    auto work = [](const Group &x) {
      uint64_t w = 0;
      for (const StepItem *it : x.items) w += (uint64_t)1 << it->r->L;
      return w;
    };
    std::stable_sort(g.begin(), g.end(), [&](const Group &u, const Group &v) { return work(u) > work(v); });
    return g;
  }

  long submit_step(const std::vector<StepItem> &items) override {
    Slot *sl = nullptr;
    for (auto &x : slots)
      if (x.id < 0) {
        sl = &x;
        break;
      }
    if (!sl) return -ZNTT_DEVICE_ERROR;  // more steps in flight than slots: a scheduler bug
    const std::vector<Group> groups = classify(items);
    if ((int)groups.size() > kMaxGroups) return -ZNTT_DEVICE_ERROR;
    const size_t n = items.size();
    uint8_t *h = (uint8_t *)sl->host;
    memset(h, 0, 256);  // the launches' counters (next item, finished workgroups) start at zero
    auto *hx = (zntt::BatchXform *)(h + 256);
    auto *he = (zntt::BatchEntry *)(h + 256 + align_up(sizeof(zntt::BatchXform) * n, 256));
    struct L {
      size_t ent_off, nent;
      uint32_t nitems, uniform;
    };
    std::vector<L> ls;
    size_t i = 0;
    for (const Group &g : groups) {
      const size_t off = i;
      uint32_t first = 0, uniform = 0;
      for (const StepItem *item : g.items) {
        const Request *r = item->r;
        zntt::BatchXform &x = hx[i];
        memset(&x, 0, sizeof x);
        x.src = (const uint32_t *)(r->in ? r->in : r->req.coeffs);
        x.src_len = (uint64_t)1 << r->req.log_n;
        x.data = (uint32_t *)r->out;
        x.scratch = (uint32_t *)r->scratch;
        x.tw = r->tw->table;
        x.coset = r->coset->table;
        x.L = r->L;
        x.npass = r->npass;
        int s0 = 1;
        for (int p = 0; p < r->npass; p++) {
          x.K[p] = r->plan.K[p];
          x.lg[p] = r->plan.log_g[p];
          x.s0[p] = s0;
          x.tiles[p] = (uint32_t)(((size_t)1 << r->L) >> (r->plan.K[p] + r->plan.log_g[p]));
          s0 += r->plan.K[p];
        }
        const uint32_t tiles = g.kind == kWhole ? 1u : x.tiles[item->pass];
        he[i] = zntt::BatchEntry{(uint32_t)i, (uint32_t)item->pass, first, 0};
        if (i == off) uniform = tiles;
        if (tiles != uniform) uniform = 0;
        first += tiles;
        i++;
      }
      ls.push_back(L{off, i - off, first, uniform});
    }
    const size_t bytes = (uint8_t *)(he + n) - h;
    uint8_t *d = (uint8_t *)sl->dev;
    const auto *dx = (const zntt::BatchXform *)(d + 256);
    const auto *de = (const zntt::BatchEntry *)(d + ((uint8_t *)he - h));
    hipError_t e = hipEventRecord(sl->start, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream);
    if (groups.size() > 1 && e == hipSuccess) e = hipEventRecord(fork, stream);
    for (size_t k = 0; k < groups.size() && e == hipSuccess; k++) {
      const Group &g = groups[k];
      const L &l = ls[k];
      hipStream_t st = k == 0 ? stream : aux[(k - 1) % aux.size()];
      if (k > 0) e = hipStreamWaitEvent(st, fork, 0);
      uint32_t *ctr = (uint32_t *)d + 2 * k;
      if (e != hipSuccess) break;
      if (g.kind == kGeneric)
        e = zntt::launch_batch_level(field, dx, de + l.ent_off, (uint32_t)l.nent, l.nitems, l.uniform, ctr, grid,
                                     tile_log, st);
      else if (g.kind == kLevelR8)
        e = zntt::launch_batch_level_r8(field, g.a, g.b != 0, pads_first_pass(g), tile_log, dx, de + l.ent_off,
                                        (uint32_t)l.nent, l.nitems, l.uniform, ctr, grid, st);
      else if (field == ZNTT_FIELD_BN254_FR)
        e = zntt::launch_whole_fr(g.a, dx, de + l.ent_off, (uint32_t)l.nent, ctr, nullptr, st);
      else if (field == ZNTT_FIELD_GOLDILOCKS)
        e = zntt::launch_whole_gl(g.a, dx, de + l.ent_off, (uint32_t)l.nent, ctr, nullptr, st);
      else
        e = zntt::launch_whole_f31(field, g.a, dx, de + l.ent_off, (uint32_t)l.nent, ctr, nullptr, st);
      if (e == hipSuccess && k > 0) e = hipEventRecord(joins[(k - 1) % aux.size()], st);
      if (e == hipSuccess && k > 0) e = hipStreamWaitEvent(stream, joins[(k - 1) % aux.size()], 0);
    }
    if (e == hipSuccess) e = hipEventRecord(sl->end, stream);
    if (e != hipSuccess) return -hip_status(e);
    sl->id = next_id++;
    return sl->id;
  }

  int step_done(long id, bool block, bool *done, double *busy_ms) override {
    *done = false;
    *busy_ms = 0;
    Slot *sl = nullptr;
    for (auto &x : slots)
      if (x.id == id) sl = &x;
    if (!sl) return ZNTT_DEVICE_ERROR;
    hipError_t e = block ? hipEventSynchronize(sl->end) : hipEventQuery(sl->end);
    if (e == hipErrorNotReady) return ZNTT_OK;
    if (e != hipSuccess) return hip_status(e);
    if (!block) {  // the non-blocking query retires the step and frees its slot
      float ms = 0;
      hipEventElapsedTime(&ms, sl->start, sl->end);
      *busy_ms = ms;
      sl->id = -1;
    }
    *done = true;
    return ZNTT_OK;
  }
};

// Smallest generator of the multiplicative group: the coset shift prepared
// at creation (5 BN254 Fr, 7 Goldilocks, 31 BabyBear, 3 KoalaBear).
uint64_t default_generator(int field) {
  switch (field) {
    case ZNTT_FIELD_BN254_FR: return 5;
    case ZNTT_FIELD_GOLDILOCKS: return 7;
    case ZNTT_FIELD_BABYBEAR: return 31;
    default: return 3;
  }
}

int check_request(const zntt_server *s, const zntt_lde_request *r) {
  if (!r || !r->shift) return ZNTT_INVALID_ARGUMENT;
  if (r->log_n >= 64 || r->log_blowup >= 64 || r->log_n + r->log_blowup >= 64) return ZNTT_LOG_TOO_LARGE;
  if (r->len != ((size_t)1 << r->log_n)) return ZNTT_LENGTH_MISMATCH;
  const size_t L = r->log_n + r->log_blowup;
  if (L > (size_t)two_adicity(s->cfg.field)) return ZNTT_ORDER_TOO_LARGE;
  if (L > ZNTT_MAX_LOG_N || L < 3) return ZNTT_UNSUPPORTED;
  return ZNTT_OK;
}

}  // namespace

extern "C" {

int zntt_server_create(zntt_ctx *ctx, const zntt_server_config *cfg, zntt_server **out) {
  if (!cfg || !out) return ZNTT_INVALID_ARGUMENT;
  *out = nullptr;
  if (!field_ok(cfg->field)) return ZNTT_INVALID_ARGUMENT;
  if (!cfg->simulate && !ctx) return ZNTT_INVALID_ARGUMENT;
  if (cfg->max_active < 0 || cfg->max_log_n < 0 || cfg->max_log_n > ZNTT_MAX_LOG_N) return ZNTT_INVALID_ARGUMENT;
  if (cfg->allocator && (!cfg->allocator->alloc || !cfg->allocator->free)) return ZNTT_INVALID_ARGUMENT;
  auto s = std::make_unique<zntt_server>();
  s->ctx = ctx;
  s->cfg = *cfg;
  if (!s->cfg.hbm_budget) s->cfg.hbm_budget = kDefaultBudget;
  if (!s->cfg.max_active) s->cfg.max_active = kDefaultActive;
  if (cfg->field != ZNTT_FIELD_BN254_FR) {
    s->tile_log = 12;
    s->kmax = 10;
  }
  if (cfg->simulate) {
    s->dev = std::make_unique<SimDevice>(cfg->allocator);
  } else {
    if (hipSetDevice(ctx->device) != hipSuccess) return ZNTT_DEVICE_ERROR;
    auto d = std::make_unique<HipDevice>(ctx, cfg->field, s->tile_log);
    int st = d->init(s->cfg);
    if (st) return st;
    if (cfg->max_log_n >= 3) {
      // Resident tables: the powers table of the largest size serves every
      // smaller standard root; the generator's coset table every request
      // that uses it.
      CallGuard g(ctx);
      TwiddleEntry *t = nullptr;
      st = power_table(ctx, cfg->field, f_primitive_root(cfg->field, cfg->max_log_n), cfg->max_log_n, ctx->stream,
                       &t);
      if (st) return st;
      const Elem gen = f_from_u64(cfg->field, default_generator(cfg->field));
      const uint32_t *cos = nullptr;
      st = coset_table(ctx, cfg->field, gen, (size_t)1 << (cfg->max_log_n - 3), ctx->stream, &cos);
      if (st) return st;
    }
    s->dev = std::move(d);
  }
  zntt_server *raw = s.release();
  if (ctx) ctx_retain(ctx);  // released by zntt_server_destroy
  if (!cfg->simulate) raw->thread = std::thread(scheduler_main, raw);
  *out = raw;
  return ZNTT_OK;
}

int zntt_server_submit(zntt_server *s, const zntt_lde_request *r, uint64_t *ticket) {
  if (!s || !ticket) return ZNTT_INVALID_ARGUMENT;
  *ticket = 0;
  int st = check_request(s, r);
  if (st) return st;
  auto q = std::make_unique<Request>();
  q->req = *r;
  q->shift = load_elem(s->cfg.field, r->shift);
  q->L = (int)(r->log_n + r->log_blowup);
  q->plan = zntt::plan_passes_shape(q->L, s->tile_log, s->kmax);
  q->npass = q->plan.passes;
  if (q->npass > zntt::kBatchMaxPasses) return ZNTT_UNSUPPORTED;
  const size_t n = (size_t)1 << q->L;
  // Sizes are kept 256-byte aligned, the arena's granule, so that admission
  // and retirement account exactly what the allocations hold.
  q->own_out = r->out == nullptr;
  q->out_bytes = q->own_out ? align_up(n * user_bytes(s->cfg.field), 256) : 0;
  q->scratch_bytes = q->npass > 1 ? align_up(n * dev_bytes(s->cfg.field), 256) : 0;
  q->in_bytes = r->coeffs ? 0 : align_up(((size_t)1 << r->log_n) * user_bytes(s->cfg.field), 256);
  q->bytes = q->out_bytes + q->scratch_bytes + q->in_bytes;
  q->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->broken) return s->broken;
  if (s->stop) return ZNTT_INVALID_ARGUMENT;
  q->ticket = s->next_ticket++;
  *ticket = q->ticket;
  s->queue.push_back(q.get());
  s->all[q->ticket] = std::move(q);
  s->stats.submitted++;
  s->cv_work.notify_one();
  return ZNTT_OK;
}

int zntt_server_query(zntt_server *s, uint64_t ticket, zntt_request_info *out) {
  if (!s || !out) return ZNTT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return ZNTT_INVALID_ARGUMENT;
  const Request *r = it->second.get();
  memset(out, 0, sizeof *out);
  out->state = r->state;
  out->status = r->status;
  out->log_n = r->req.log_n;
  out->log_blowup = r->req.log_blowup;
  if (r->state != ZNTT_REQ_QUEUED && r->state != ZNTT_REQ_FAILED && r->state != ZNTT_REQ_CANCELLED)
    out->queued_ms = ms_between(r->t_submit, r->t_admit);
  if (r->state >= ZNTT_REQ_DONE) out->latency_ms = ms_between(r->t_submit, r->t_done);
  if (r->state == ZNTT_REQ_DONE && (r->req.keep_result || !r->own_out)) out->data = r->own_out ? r->out : r->req.out;
  return ZNTT_OK;
}

int zntt_server_read(zntt_server *s, uint64_t ticket, void *host, size_t len) {
  if (!s || !host || s->cfg.simulate) return ZNTT_INVALID_ARGUMENT;
  const void *src = nullptr;
  size_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return ZNTT_INVALID_ARGUMENT;
    const Request *r = it->second.get();
    if (r->state != ZNTT_REQ_DONE) return ZNTT_INVALID_ARGUMENT;
    src = r->own_out ? r->out : r->req.out;
    if (!src) return ZNTT_INVALID_ARGUMENT;  // result not kept
    if (len != ((size_t)1 << r->L)) return ZNTT_LENGTH_MISMATCH;
    bytes = len * user_bytes(s->cfg.field);
  }
  // The request is done: its last step completed before it was marked so.
  hipSetDevice(s->ctx->device);
  if (ring_wanted(bytes)) {
    std::lock_guard<std::mutex> lk(s->read_mu);
    if (!s->read_ring && !s->read_ring_failed) {
      if (hipStreamCreateWithFlags(&s->read_stream, hipStreamNonBlocking) == hipSuccess)
        s->read_ring = ring_create();
      s->read_ring_failed = s->read_ring == nullptr;
    }
    if (s->read_ring) return ring_copy_to_host(s->read_ring, host, src, bytes, s->read_stream);
  }
  return hip_status(hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost));
}

int zntt_server_wait(zntt_server *s, uint64_t ticket, double timeout_ms) {
  if (!s) return ZNTT_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return ZNTT_INVALID_ARGUMENT;
  const Request *r = it->second.get();
  auto finished = [&] { return r->state >= ZNTT_REQ_DONE; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, finished);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
    return ZNTT_DEVICE_ERROR;
  }
  return ZNTT_OK;
}

int zntt_server_cancel(zntt_server *s, uint64_t ticket) {
  if (!s) return ZNTT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return ZNTT_INVALID_ARGUMENT;
  Request *r = it->second.get();
  if (r->state == ZNTT_REQ_QUEUED) {
    s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r));
    r->cancelled = true;
    finish(s, r, ZNTT_REQ_CANCELLED, ZNTT_OK);
  } else if (r->state == ZNTT_REQ_RUNNING) {
    r->cancelled = true;  // retirement frees everything and marks it cancelled
  }
  return ZNTT_OK;
}

int zntt_server_release(zntt_server *s, uint64_t ticket) {
  if (!s) return ZNTT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return ZNTT_INVALID_ARGUMENT;
  Request *r = it->second.get();
  if (r->state < ZNTT_REQ_DONE) return ZNTT_INVALID_ARGUMENT;  // still queued or running
  if (r->resident) {
    s->dev->free_result(r);
    s->in_use -= r->out_bytes;
    r->resident = false;
    s->cv_work.notify_one();  // memory for the queue
  }
  s->all.erase(it);
  return ZNTT_OK;
}

int zntt_server_drain(zntt_server *s) {
  if (!s) return ZNTT_INVALID_ARGUMENT;
  if (s->cfg.simulate) return ZNTT_INVALID_ARGUMENT;  // nothing runs the steps
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->broken || (s->queue.empty() && s->active.empty()); });
  return s->broken;
}

int zntt_server_stats_get(zntt_server *s, zntt_server_stats *out) {
  if (!s || !out) return ZNTT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->bytes_in_use = s->in_use;
  out->queued = s->queue.size();
  out->active = s->active.size();
  return ZNTT_OK;
}

int zntt_server_sim_pump(zntt_server *s, int *in_flight) {
  if (!s || !s->cfg.simulate) return ZNTT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  iterate(s);
  if (in_flight) *in_flight = (int)s->inflight.size();
  return s->broken;
}

int zntt_server_sim_complete(zntt_server *s, int n) {
  if (!s || !s->cfg.simulate || n < 0) return ZNTT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto *d = static_cast<SimDevice *>(s->dev.get());
  int k = 0;
  for (const auto &st : s->inflight) {
    if (k++ >= n) break;
    d->completed_upto = std::max(d->completed_upto, st.id);
  }
  return ZNTT_OK;
}

void zntt_server_destroy(zntt_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_one();
  }
  if (s->thread.joinable()) s->thread.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &kv : s->all) {
      Request *r = kv.second.get();
      // Kept results, and requests a simulation or a device failure left
      // admitted, still hold memory.
      if (r->resident) s->dev->release(r, false);
    }
    s->all.clear();
  }
  s->dev.reset();
  ring_destroy(s->read_ring);
  if (s->read_stream) hipStreamDestroy(s->read_stream);
  zntt_ctx *ctx = s->ctx;
  delete s;
  if (ctx) ctx_release(ctx);
}

int zntt_fill_counter(zntt_ctx *c, int field, uint64_t seed, uint64_t stream_id, void *data, size_t n,
                      void *stream) {
  if (!c || !field_ok(field) || (!data && n)) return ZNTT_INVALID_ARGUMENT;
  if (hipSetDevice(c->device) != hipSuccess) return ZNTT_DEVICE_ERROR;
  return hip_status(zntt::launch_fill_counter(field, (uint32_t *)data, n, seed, stream_id, (hipStream_t)stream));
}

}  // extern "C"
