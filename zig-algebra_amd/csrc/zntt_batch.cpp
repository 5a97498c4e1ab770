// Batched transforms: many transforms of uneven size and mixed direction run
// as one persistent launch per pass level, captured once in a HIP graph and
// replayed, with the tile shape, pass length, grid occupancy and work order
// chosen by a small autotuner.  No reference counterpart (zig-ntt transforms one slice per
// call); each transform's result equals zntt_execute_device on it alone.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "ctx_internal.hpp"

using namespace zntt_detail;

namespace {

struct Shape {
  int tile_log;    // elements per workgroup tile = 2^tile_log
  int kmax;        // longest pass, in stages
  int wgs_per_cu;  // resident workgroups of the persistent grid per CU
  int order;       // transform order inside a level: 0 as given, 1 largest first
  int r8;          // 1: passes the register-round level kernel can take run in it
  bool operator<(const Shape &o) const {
    return std::tie(tile_log, kmax, wgs_per_cu, order, r8) <
           std::tie(o.tile_log, o.kmax, o.wgs_per_cu, o.order, o.r8);
  }
};

Shape default_shape(int field) {
  return field == ZNTT_FIELD_BN254_FR ? Shape{11, 9, 2, 0, 1} : Shape{12, 10, 3, 0, 1};
}

// The shapes the autotuner may pick: tiles of the field's largest size or
// half of it, passes two or three stages shorter than the tile, one to four
// resident workgroups per CU as LDS (160 KiB per CU) allows.
bool shape_ok(int field, const Shape &sh) {
  const int base_tl = field == ZNTT_FIELD_BN254_FR ? 11 : 12;
  if (sh.tile_log != base_tl - 1 && sh.tile_log != base_tl) return false;
  if (sh.kmax != sh.tile_log - 2 && sh.kmax != sh.tile_log - 3) return false;
  if (sh.wgs_per_cu < 1 || sh.wgs_per_cu > 4) return false;
  if (((size_t)zntt::dev_words(field) * 4 << sh.tile_log) * sh.wgs_per_cu > 160 * 1024) return false;
  return (sh.order == 0 || sh.order == 1) && (sh.r8 == 0 || sh.r8 == 1);
}

}  // namespace

struct zntt_batch {
  zntt_ctx *ctx = nullptr;
  int field = 0;
  std::vector<zntt::BatchXform> xf;       // host copy (data / table pointers filled)
  std::vector<TwiddleEntry *> pinned_tw;
  std::vector<CosetEntry *> pinned_cos;
  void *scratch = nullptr;                 // all multi-pass scratch
  size_t scratch_bytes = 0;
  void *dev = nullptr;                     // xf + entries + level counters
  size_t dev_bytes = 0;
  // The launches of one execution, level by level.  Launches of one level
  // are independent and run as parallel branches (of the captured graph).
  enum Kind { kGeneric, kWhole, kLevelR8 };
  struct Launch {
    Kind kind;
    int level;
    int whole_log;  // kWhole: single-pass transforms of 2^whole_log elements
    int K, odd;     // kLevelR8: pass length and parity of the first stage
    uint32_t ent_off, nent, nitems, uniform;
    const uint32_t *tw_all;  // whole launches: the one table all transforms use, else null
  };
  std::vector<Launch> launches;
  std::vector<hipStream_t> aux;  // branch streams (launches of one level beyond the first)
  std::vector<hipEvent_t> ev;    // fork event, then one join event per branch stream
  zntt::BatchXform *d_xf = nullptr;
  zntt::BatchEntry *d_ent = nullptr;
  uint32_t *d_counters = nullptr;
  Shape shape = {};
  int grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double tuned_ms = 0;
};

namespace {

void drop_graph(zntt_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// (Re)builds the pass plans, the work list and the device copies for `shape`.
int build_plan(zntt_batch *b, const Shape &shape) {
  zntt_ctx *c = b->ctx;
  drop_graph(b);
  if (b->dev) {
    hipStreamSynchronize(c->stream);
    dev_free(c, b->dev, b->dev_bytes);
    b->dev = nullptr;
  }
  int levels = 0;
  for (size_t t = 0; t < b->xf.size(); t++) {
    zntt::BatchXform &x = b->xf[t];
    const zntt::PassPlan plan = zntt::plan_passes_shape(x.L, shape.tile_log, shape.kmax);
    if (plan.passes > zntt::kBatchMaxPasses) return ZNTT_UNSUPPORTED;
    x.npass = plan.passes;
    levels = std::max(levels, plan.passes);
    int s0 = 1;
    for (int p = 0; p < plan.passes; p++) {
      x.K[p] = plan.K[p];
      x.lg[p] = plan.log_g[p];
      x.s0[p] = s0;
      x.tiles[p] = (uint32_t)(((size_t)1 << x.L) >> (plan.K[p] + plan.log_g[p]));
      s0 += plan.K[p];
    }
  }
  // Entries level by level; inside a level the transforms in the batch's
  // order (order 0) or largest first (order 1, so the longest-running tiles
  // start early and the tail is short).
  std::vector<uint32_t> idx(b->xf.size());
  for (size_t t = 0; t < idx.size(); t++) idx[t] = (uint32_t)t;
  if (shape.order == 1)
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t u, uint32_t v) { return b->xf[u].L > b->xf[v].L; });
  std::vector<zntt::BatchEntry> ent;
  b->launches.clear();
  // Transforms that fit one register-round tile go to the whole-transform
  // kernel, one launch per size.  Passes with a full tile go to the
  // register-round level kernel, one launch per (pass length, parity of the
  // first stage); the rest to the generic level kernel.
  auto whole = [&](const zntt::BatchXform &x) {
    return x.npass == 1 && x.L >= zntt::kWholeMinLog && x.L <= zntt::whole_max_log(b->field);
  };
  auto r8 = [&](const zntt::BatchXform &x, int p) {
    return shape.r8 && x.K[p] + x.lg[p] == shape.tile_log && zntt::level_r8_ok(b->field, x.K[p], shape.tile_log);
  };
  for (int p = 0; p < levels; p++) {
    std::map<std::pair<int, int>, std::vector<uint32_t>> groups;  // (K, odd) -> transforms
    std::vector<uint32_t> generic;
    for (uint32_t t : idx) {
      const zntt::BatchXform &x = b->xf[t];
      if (p >= x.npass || whole(x)) continue;
      if (r8(x, p))
        groups[{x.K[p], x.s0[p] & 1}].push_back(t);
      else
        generic.push_back(t);
    }
    // One launch over the given transforms' pass p.
    auto add = [&](zntt_batch::Kind kind, int K, int odd, const std::vector<uint32_t> &ts) {
      const uint32_t off = (uint32_t)ent.size();
      uint32_t first = 0, uniform = 0;
      bool equal = true;
      for (uint32_t t : ts) {
        const zntt::BatchXform &x = b->xf[t];
        ent.push_back(zntt::BatchEntry{t, (uint32_t)p, first, 0});
        first += x.tiles[p];
        if (uniform == 0) uniform = x.tiles[p];
        equal = equal && uniform == x.tiles[p];
      }
      b->launches.push_back({kind, p, 0, K, odd, off, (uint32_t)ent.size() - off, first, equal ? uniform : 0, nullptr});
    };
    // Largest launches first, so that smaller ones fill the tails.
    std::vector<std::pair<uint64_t, std::pair<int, int>>> by_size;
    for (auto &g : groups) {
      uint64_t items = 0;
      for (uint32_t t : g.second) items += b->xf[t].tiles[p];
      by_size.push_back({items, g.first});
    }
    std::sort(by_size.begin(), by_size.end(), [](const auto &u, const auto &v) { return u.first > v.first; });
    for (auto &g : by_size) add(zntt_batch::kLevelR8, g.second.first, g.second.second, groups[g.second]);
    if (!generic.empty()) add(zntt_batch::kGeneric, 0, 0, generic);
    if (p == 0) {
      for (int L = zntt::whole_max_log(b->field); L >= zntt::kWholeMinLog; L--) {
        const uint32_t woff = (uint32_t)ent.size();
        uint32_t k = 0;
        const uint32_t *tw = nullptr;
        bool same = true;
        for (uint32_t t : idx)
          if (whole(b->xf[t]) && b->xf[t].L == L) {
            ent.push_back(zntt::BatchEntry{t, 0, k++, 0});
            if (!tw) tw = b->xf[t].tw;
            same = same && tw == b->xf[t].tw;
          }
        if (k) b->launches.push_back({zntt_batch::kWhole, 0, L, 0, 0, woff, k, k, 1, same ? tw : nullptr});
      }
    }
  }
  // Branch streams and events for the widest level.
  size_t width = 0;
  for (int p = 0; p < levels; p++) {
    size_t w = 0;
    for (auto &l : b->launches) w += l.level == p ? 1 : 0;
    width = std::max(width, w);
  }
  while (b->aux.size() + 1 < width) {
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_status(e);
    b->aux.push_back(st);
  }
  while (b->ev.size() < b->aux.size() + 1) {
    hipEvent_t e0 = nullptr;
    hipError_t e = hipEventCreateWithFlags(&e0, hipEventDisableTiming);
    if (e != hipSuccess) return hip_status(e);
    b->ev.push_back(e0);
  }
  const size_t xf_bytes = b->xf.size() * sizeof(zntt::BatchXform);
  const size_t it_bytes = ent.size() * sizeof(zntt::BatchEntry);
  const size_t st_bytes = 2 * b->launches.size() * sizeof(uint32_t);  // next item, finished workgroups
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  b->dev_bytes = up(xf_bytes) + up(it_bytes) + up(st_bytes);
  int st = dev_alloc(c, b->dev_bytes, &b->dev);
  if (st) return st;
  uint8_t *base = (uint8_t *)b->dev;
  b->d_xf = (zntt::BatchXform *)base;
  b->d_ent = (zntt::BatchEntry *)(base + up(xf_bytes));
  b->d_counters = (uint32_t *)(base + up(xf_bytes) + up(it_bytes));
  hipError_t e = hipMemcpyAsync(b->d_xf, b->xf.data(), xf_bytes, hipMemcpyHostToDevice, c->stream);
  // Counters start at zero; every launch returns its own to zero, so replays
  // need no clearing node in the graph.
  if (e == hipSuccess) e = hipMemsetAsync(b->d_counters, 0, st_bytes, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_ent, ent.data(), it_bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_status(e);
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->grid = std::max(1, cus * shape.wgs_per_cu);
  b->shape = shape;
  return ZNTT_OK;
}

hipError_t launch_one(zntt_batch *b, size_t i, hipStream_t s) {
  const zntt_batch::Launch &l = b->launches[i];
  uint32_t *ctr = b->d_counters + 2 * i;
  const zntt::BatchEntry *ent = b->d_ent + l.ent_off;
  switch (l.kind) {
    case zntt_batch::kGeneric:
      return zntt::launch_batch_level(b->field, b->d_xf, ent, l.nent, l.nitems, l.uniform, ctr, b->grid,
                                      b->shape.tile_log, s);
    case zntt_batch::kLevelR8:
      return zntt::launch_batch_level_r8(b->field, l.K, l.odd != 0, false, b->shape.tile_log, b->d_xf, ent,
                                         l.nent, l.nitems, l.uniform, ctr, b->grid, s);
    default:
      if (b->field == ZNTT_FIELD_BN254_FR) return zntt::launch_whole_fr(l.whole_log, b->d_xf, ent, l.nent, ctr, l.tw_all, s);
      if (b->field == ZNTT_FIELD_GOLDILOCKS) return zntt::launch_whole_gl(l.whole_log, b->d_xf, ent, l.nent, ctr, l.tw_all, s);
      return zntt::launch_whole_f31(b->field, l.whole_log, b->d_xf, ent, l.nent, ctr, l.tw_all, s);
  }
}

// The persistent launches level by level (each leaves its counters zero).
// A level's first launch goes on `s`, the others on branch streams that fork
// from `s` and join it again before the next level.
hipError_t enqueue(zntt_batch *b, hipStream_t s) {
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < b->launches.size() && e == hipSuccess;) {
    size_t j = i;
    while (j < b->launches.size() && b->launches[j].level == b->launches[i].level) j++;
    if (j - i > 1) e = hipEventRecord(b->ev[0], s);
    for (size_t k = i; k < j && e == hipSuccess; k++) {
      hipStream_t st = k == i ? s : b->aux[k - i - 1];
      if (k > i) e = hipStreamWaitEvent(st, b->ev[0], 0);
      if (e == hipSuccess) e = launch_one(b, k, st);
      if (e == hipSuccess && k > i) e = hipEventRecord(b->ev[k - i], st);
    }
    for (size_t k = i + 1; k < j && e == hipSuccess; k++) e = hipStreamWaitEvent(s, b->ev[k - i], 0);
    i = j;
  }
  return e;
}

void release(zntt_batch *b) {
  zntt_ctx *c = b->ctx;
  hipStreamSynchronize(c->stream);
  hipDeviceSynchronize();
  drop_graph(b);
  if (b->dev) dev_free(c, b->dev, b->dev_bytes);
  if (b->scratch) dev_free(c, b->scratch, b->scratch_bytes);
  b->dev = b->scratch = nullptr;
  for (auto *t : b->pinned_tw) t->pins--;
  for (auto *e : b->pinned_cos) e->pins--;
  b->pinned_tw.clear();
  b->pinned_cos.clear();
  for (auto st : b->aux) hipStreamDestroy(st);
  for (auto e : b->ev) hipEventDestroy(e);
  b->aux.clear();
  b->ev.clear();
}

// Autotune decisions keyed by the batch's signature (field, sizes, directions),
// shared by every context of the process.
std::map<std::vector<int>, Shape> g_tuned;
std::mutex g_tuned_mu;

std::vector<int> signature(int field, const std::vector<zntt::BatchXform> &xf) {
  std::vector<int> k{field};
  for (auto &x : xf) k.push_back(x.L * 4 + (x.has_scale ? 1 : 0) + (x.coset ? 2 : 0));
  std::sort(k.begin() + 1, k.end());
  return k;
}

}  // namespace

extern "C" {

int zntt_batch_create(zntt_ctx *c, const zntt_transform *ts, size_t count, zntt_batch **out) {
  if (!c || !out || (!ts && count)) return ZNTT_INVALID_ARGUMENT;
  *out = nullptr;
  if (count == 0 || count > (1u << 24)) return ZNTT_INVALID_ARGUMENT;
  const int field = ts[0].field;
  if (!field_ok(field)) return ZNTT_INVALID_ARGUMENT;
  for (size_t i = 0; i < count; i++) {
    const zntt_transform &t = ts[i];
    if (t.field != field || !t.root || !t.data) return ZNTT_INVALID_ARGUMENT;
    if (t.direction < ZNTT_FORWARD || t.direction > ZNTT_COSET_FORWARD) return ZNTT_INVALID_ARGUMENT;
    if (t.direction == ZNTT_COSET_FORWARD && !t.shift) return ZNTT_INVALID_ARGUMENT;
    if (t.log_n >= 64) return ZNTT_LOG_TOO_LARGE;
    if (t.log_n < 3 || t.log_n > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;  // a tile holds >= 8 elements
  }
  CallGuard g(c);
  hipError_t he = zntt::configure_batch_lds();
  if (he != hipSuccess) return hip_status(he);
  auto *b = new zntt_batch();
  b->ctx = c;
  b->field = field;
  b->xf.resize(count);
  // Largest transforms first, so that nested roots (the standard ones) are
  // served by the first, largest table.
  std::vector<size_t> order(count);
  for (size_t i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t x) { return ts[a].log_n > ts[x].log_n; });
  size_t scratch_elems = 0;
  int st = ZNTT_OK;
  for (size_t oi = 0; oi < count && st == ZNTT_OK; oi++) {
    const size_t i = order[oi];
    const zntt_transform &t = ts[i];
    zntt::BatchXform &x = b->xf[i];
    memset(&x, 0, sizeof x);
    x.L = (int)t.log_n;
    x.data = (uint32_t *)t.data;
    x.src = x.data;
    x.src_len = (uint64_t)1 << x.L;
    Elem root = load_elem(field, t.root);
    if (t.direction == ZNTT_INVERSE) {
      root = f_inv(field, root);
      const auto sc = f_to_dev(field, f_inv(field, f_from_u64(field, (uint64_t)1 << t.log_n)));
      memcpy(x.scale, sc.data(), sizeof x.scale);
      x.has_scale = 1;
    }
    TwiddleEntry *te = nullptr;
    st = power_table(c, field, root, x.L, c->stream, &te);
    if (st) break;
    te->pins++;
    b->pinned_tw.push_back(te);
    x.tw = te->table;
    if (t.direction == ZNTT_COSET_FORWARD) {
      const uint32_t *cos = nullptr;
      st = coset_table(c, field, load_elem(field, t.shift), (size_t)1 << t.log_n, c->stream, &cos);
      if (st) break;
      for (auto &ce : c->cosets)
        if (ce.table == cos) {
          ce.pins++;
          b->pinned_cos.push_back(&ce);
          break;
        }
      x.coset = cos;
    }
    scratch_elems += (size_t)1 << t.log_n;
  }
  if (st == ZNTT_OK) {
    // Every transform gets its own scratch (any may need several passes
    // under some tuned shape).
    b->scratch_bytes = scratch_elems * dev_bytes(field);
    st = dev_alloc(c, b->scratch_bytes, &b->scratch);
  }
  if (st == ZNTT_OK) {
    size_t off = 0;
    for (auto &x : b->xf) {
      x.scratch = (uint32_t *)((uint8_t *)b->scratch + off * dev_bytes(field));
      off += (size_t)1 << x.L;
    }
    Shape sh = default_shape(field);
    {
      std::lock_guard<std::mutex> lk(g_tuned_mu);
      auto it = g_tuned.find(signature(field, b->xf));
      if (it != g_tuned.end()) sh = it->second;
    }
    st = build_plan(b, sh);
  }
  if (st) {
    release(b);
    delete b;
    return st;
  }
  ctx_retain(c);  // the batch reads the context's stream and pinned tables until destroyed
  *out = b;
  return ZNTT_OK;
}

int zntt_batch_execute(zntt_batch *b, void *stream) {
  if (!b) return ZNTT_INVALID_ARGUMENT;
  zntt_ctx *c = b->ctx;
  CallGuard g(c);
  hipStream_t s = (hipStream_t)stream;
  if (!b->exec) {
    // Capture on the context's own stream (the null stream cannot be
    // captured); the instantiated graph then launches on any stream.
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t le = enqueue(b, c->stream);
      e = hipStreamEndCapture(c->stream, &b->graph);
      if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGetLastError();
      drop_graph(b);
      return hip_status(e);
    }
  }
  return hip_status(hipGraphLaunch(b->exec, s));
}

int zntt_batch_status(zntt_batch *b, void *stream) {
  if (!b) return ZNTT_INVALID_ARGUMENT;
  return hip_status(hipStreamSynchronize((hipStream_t)stream));
}

int zntt_batch_autotune(zntt_batch *b, int reps, double *best_ms) {
  if (!b || reps < 1) return ZNTT_INVALID_ARGUMENT;
  zntt_ctx *c = b->ctx;
  std::vector<Shape> cands;
  const int base_tl = b->field == ZNTT_FIELD_BN254_FR ? 11 : 12;
  for (int r8 : {1, 0})
    for (int tl : {base_tl - 1, base_tl})
      for (int kd : {2, 3})
        for (int w : {1, 2, 3, 4})
          for (int order : {0, 1}) {
            const Shape sh{tl, tl - kd, w, order, r8};
            if (shape_ok(b->field, sh)) cands.push_back(sh);
          }
  // Time on shadow copies of the data so the caller's buffers are untouched.
  std::vector<uint32_t *> real(b->xf.size());
  size_t total = 0;
  for (auto &x : b->xf) total += ((size_t)1 << x.L) * user_bytes(b->field);
  void *shadow = nullptr;
  int st;
  {
    CallGuard g(c);
    st = dev_alloc(c, total, &shadow);
    if (st) return st;
    hipMemsetAsync(shadow, 0, total, c->stream);
    size_t off = 0;
    for (size_t i = 0; i < b->xf.size(); i++) {
      real[i] = b->xf[i].data;
      b->xf[i].data = (uint32_t *)((uint8_t *)shadow + off);
      b->xf[i].src = b->xf[i].data;
      off += ((size_t)1 << b->xf[i].L) * user_bytes(b->field);
    }
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  Shape best = b->shape;
  double best_t = 1e30;
  for (const Shape &sh : cands) {
    CallGuard g(c);
    st = build_plan(b, sh);
    if (st) break;
    hipError_t e = enqueue(b, c->stream);  // warm-up
    hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = enqueue(b, c->stream);
    hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e != hipSuccess) {
      st = hip_status(e);
      break;
    }
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms / reps < best_t) {
      best_t = ms / reps;
      best = sh;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  CallGuard g(c);
  for (size_t i = 0; i < b->xf.size(); i++) b->xf[i].data = (uint32_t *)(b->xf[i].src = real[i]);
  hipStreamSynchronize(c->stream);
  dev_free(c, shadow, total);
  if (st) {
    build_plan(b, b->shape);
    return st;
  }
  st = build_plan(b, best);
  if (st) return st;
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[signature(b->field, b->xf)] = best;
  }
  b->tuned_ms = best_t;
  if (best_ms) *best_ms = best_t;
  return ZNTT_OK;
}

int zntt_batch_shape(const zntt_batch *b, int *tile_log, int *kmax, int *wgs_per_cu, int *order,
                     int *level_kernel) {
  if (!b) return ZNTT_INVALID_ARGUMENT;
  if (tile_log) *tile_log = b->shape.tile_log;
  if (kmax) *kmax = b->shape.kmax;
  if (wgs_per_cu) *wgs_per_cu = b->shape.wgs_per_cu;
  if (order) *order = b->shape.order;
  if (level_kernel) *level_kernel = b->shape.r8;
  return ZNTT_OK;
}

int zntt_batch_set_shape(zntt_batch *b, int tile_log, int kmax, int wgs_per_cu, int order, int level_kernel) {
  if (!b) return ZNTT_INVALID_ARGUMENT;
  const Shape sh{tile_log, kmax, wgs_per_cu, order, level_kernel};
  if (!shape_ok(b->field, sh)) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(b->ctx);
  const int st = build_plan(b, sh);
  if (st) build_plan(b, b->shape);  // keep the batch runnable in its previous shape
  return st;
}

void zntt_batch_destroy(zntt_batch *b) {
  if (!b) return;
  zntt_ctx *c = b->ctx;
  {
    CallGuard g(c);
    release(b);
  }
  delete b;
  ctx_release(c);
}

}  // extern "C"
