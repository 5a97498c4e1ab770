// C ABI of the MI355X NTT library: context, device memory under an HBM budget,
// the device-resident twiddle cache, and the zig-ntt entry points
// (libs/ntt/src/root.zig:33-212) with the reference's validation order.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <list>
#include <memory>
#include <mutex>
#include <vector>

#include "host_field.hpp"
#include "kernels.hpp"
#include "zntt_gpu.h"
#include "ctx_internal.hpp"

using zntt::host::U256;
using namespace zntt_detail;

namespace zntt_detail {


bool field_ok(int field) { return field >= ZNTT_FIELD_BN254_FR && field <= ZNTT_FIELD_KOALABEAR; }
const zntt::host::F31 *f31(int field) {
  return field == ZNTT_FIELD_BABYBEAR ? &zntt::host::kBabyBear
                                      : (field == ZNTT_FIELD_KOALABEAR ? &zntt::host::kKoalaBear : nullptr);
}
size_t user_bytes(int field) { return field == ZNTT_FIELD_BN254_FR ? 32 : 8; }
size_t dev_bytes(int field) { return 4 * (size_t)zntt::dev_words(field); }

Elem load_elem(int field, const void *p) {
  Elem e = {};
  memcpy(e.w, p, user_bytes(field));
  return e;
}

// --- host field dispatch on storage-layout elements -------------------------
U256 to_u256(const Elem &e) {
  U256 x;
  memcpy(x.l, e.w, 32);
  return x;
}
Elem from_u256(const U256 &x) {
  Elem e = {};
  memcpy(e.w, x.l, 32);
  return e;
}
uint64_t to_u64(const Elem &e) { return (uint64_t)e.w[0] | ((uint64_t)e.w[1] << 32); }
Elem from_u64e(uint64_t v) {
  Elem e = {};
  e.w[0] = (uint32_t)v;
  e.w[1] = (uint32_t)(v >> 32);
  return e;
}

Elem f_mul(int field, const Elem &a, const Elem &b) {
  if (field == ZNTT_FIELD_BN254_FR) return from_u256(zntt::host::Fr::mont(to_u256(a), to_u256(b)));
  if (const auto *f = f31(field)) return from_u64e(f->mul(to_u64(a), to_u64(b)));
  return from_u64e(zntt::host::Gl::mul(to_u64(a), to_u64(b)));
}
Elem f_inv(int field, const Elem &a) {
  if (field == ZNTT_FIELD_BN254_FR) return from_u256(zntt::host::Fr::inv(to_u256(a)));
  if (const auto *f = f31(field)) return from_u64e(f->inv(to_u64(a)));
  return from_u64e(zntt::host::Gl::inv(to_u64(a)));
}
Elem f_from_u64(int field, uint64_t v) {
  if (field == ZNTT_FIELD_BN254_FR) return from_u256(zntt::host::Fr::from_u64(v));
  if (const auto *f = f31(field)) return from_u64e(v % f->p);
  return from_u64e(v % zntt::host::Gl::P);
}
Elem f_one(int field) { return f_from_u64(field, 1); }
Elem f_primitive_root(int field, int log_size) {
  if (field == ZNTT_FIELD_BN254_FR) return from_u256(zntt::host::Fr::primitive_root(log_size));
  if (const auto *f = f31(field)) return from_u64e(f->primitive_root(log_size));
  return from_u64e(zntt::host::Gl::primitive_root(log_size));
}
int two_adicity(int field) {
  if (field == ZNTT_FIELD_BN254_FR) return 28;
  if (const auto *f = f31(field)) return f->two_adicity;
  return 32;
}
// Device constant form (Fr: 29-bit limbs of c * 2^261; Goldilocks: the value;
// 31-bit fields: c * 2^32 mod p).
std::array<uint32_t, 9> f_to_dev(int field, const Elem &a) {
  std::array<uint32_t, 9> d = {};
  if (field == ZNTT_FIELD_BN254_FR) {
    zntt::host::Fr::to_dev(to_u256(a), d.data());
  } else if (const auto *f = f31(field)) {
    d[0] = f->to_mont(to_u64(a));
  } else {
    d[0] = a.w[0];
    d[1] = a.w[1];
  }
  return d;
}

}  // namespace zntt_detail

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

namespace zntt_detail {


int hip_status(hipError_t e) {
  if (e == hipSuccess) return ZNTT_OK;
  if (e == hipErrorOutOfMemory) return ZNTT_OUT_OF_MEMORY;
  return ZNTT_DEVICE_ERROR;
}

void dev_free(zntt_ctx *c, void *p, size_t bytes) {
  if (!p) return;
  hipFree(p);
  c->in_use -= bytes;
}

// Frees the least recently used cache entry; false if nothing is evictable.
bool evict_one(zntt_ctx *c) {
  auto tb = c->tables.end();
  for (auto it = c->tables.begin(); it != c->tables.end(); ++it)
    if (it->key.empty() && it->pins == 0 && it->last_use < c->protect_from &&
        (tb == c->tables.end() || it->last_use < tb->last_use))
      tb = it;
  auto cb = c->cosets.end();
  for (auto it = c->cosets.begin(); it != c->cosets.end(); ++it)
    if (it->pins == 0 && it->last_use < c->protect_from && (cb == c->cosets.end() || it->last_use < cb->last_use))
      cb = it;
  const bool have_t = tb != c->tables.end(), have_c = cb != c->cosets.end();
  if (!have_t && !have_c) return false;
  hipDeviceSynchronize();  // in-flight work may still read the entry
  if (have_t && (!have_c || tb->last_use <= cb->last_use)) {
    dev_free(c, tb->table, tb->bytes);
    dev_free(c, tb->inv_table, tb->inv_bytes);
    c->tables.erase(tb);
  } else {
    dev_free(c, cb->table, cb->bytes);
    c->cosets.erase(cb);
  }
  return true;
}

int dev_alloc(zntt_ctx *c, size_t bytes, void **out) {
  *out = nullptr;
  if (bytes == 0) return ZNTT_OK;
  while (c->in_use + bytes > c->budget)
    if (!evict_one(c)) return ZNTT_OUT_OF_MEMORY;
  for (;;) {
    hipError_t e = hipMalloc(out, bytes);
    if (e == hipSuccess) break;
    hipGetLastError();
    if (e != hipErrorOutOfMemory || !evict_one(c)) return hip_status(e);
  }
  c->in_use += bytes;
  return ZNTT_OK;
}

// Grows a reusable buffer to at least `bytes`.  On a context with a
// reservation the new buffer is allocated before the old one is freed, and a
// failed growth keeps the old buffer, so the reserved sizes stay covered.
int ensure(zntt_ctx *c, void **buf, size_t *have, size_t bytes) {
  if (*have >= bytes) return ZNTT_OK;
  if (*buf) {
    hipStreamSynchronize(c->stream);
    hipDeviceSynchronize();
    if (c->resv_log_n >= 0) {
      void *grown = nullptr;
      int st = dev_alloc(c, bytes, &grown);
      if (st) return st;
      dev_free(c, *buf, *have);
      *buf = grown;
      *have = bytes;
      return ZNTT_OK;
    }
    dev_free(c, *buf, *have);
    *buf = nullptr;
    *have = 0;
  }
  int st = dev_alloc(c, bytes, buf);
  if (st == ZNTT_OK) *have = bytes;
  return st;
}

int ensure_pinned(zntt_ctx *c, size_t bytes) {
  if (c->pinned_bytes >= bytes) return ZNTT_OK;
  if (c->pinned) {
    hipStreamSynchronize(c->stream);
    hipDeviceSynchronize();
    hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_bytes = 0;
  }
  hipError_t e = hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    hipGetLastError();
    c->pinned = nullptr;
    return hip_status(e);
  }
  c->pinned_bytes = bytes;
  return ZNTT_OK;
}

// Frees every cache entry no batch or server has pinned (every entry with
// `all`), and the context's reusable buffers.  Pinned entries stay: live
// children read them.  A reservation stays too (its buffers and slots), unless
// `all`.  Caller holds the context lock.
void release_cache(zntt_ctx *c, bool all = false) {
  hipStreamSynchronize(c->stream);
  hipDeviceSynchronize();
  for (auto it = c->tables.begin(); it != c->tables.end();) {
    if (it->pins > 0 && !all) {
      ++it;
      continue;
    }
    dev_free(c, it->table, it->bytes);
    dev_free(c, it->inv_table, it->inv_bytes);
    it = c->tables.erase(it);
  }
  for (auto it = c->cosets.begin(); it != c->cosets.end();) {
    if (it->pins > 0 && !all) {
      ++it;
      continue;
    }
    dev_free(c, it->table, it->bytes);
    it = c->cosets.erase(it);
  }
  if (all || c->resv_log_n < 0) {
    dev_free(c, c->scratch, c->scratch_bytes);
    dev_free(c, c->staging, c->staging_bytes);
    dev_free(c, c->aux, c->aux_bytes);
    c->scratch = c->staging = c->aux = nullptr;
    c->scratch_bytes = c->staging_bytes = c->aux_bytes = 0;
  }
  if (all) {
    for (ResvSlot &sl : c->resv) {
      dev_free(c, sl.table, sl.bytes);
      sl = ResvSlot();
    }
    dev_free(c, c->resv_consts, c->resv_consts_bytes);
    c->resv_consts = nullptr;
    c->resv_consts_bytes = 0;
    c->resv_log_n = -1;
  }
  dev_free(c, c->m31_table, c->m31_table_bytes);
  c->m31_table = nullptr;
  c->m31_table_bytes = 0;
  c->m31_log_n = -1;
  if (c->pinned) hipHostFree(c->pinned);
  c->pinned = nullptr;
  c->pinned_bytes = 0;
  if (all || c->resv_log_n < 0) ring_free(c);
}

void ctx_retain(zntt_ctx *c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

// Drops one reference; the last one frees everything (no child is left, so
// nothing is pinned any more) and the context itself.
void ctx_release(zntt_ctx *c) {
  if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  {
    CallGuard g(c);
    release_cache(c, true);
  }
  hipEventDestroy(c->shared_done);
  hipStreamDestroy(c->stream);
  delete c;
}

std::vector<Elem> root_chain(int field, const Elem &root, int L) {
  std::vector<Elem> ch(L + 1);
  ch[0] = root;
  for (int k = 1; k <= L; k++) ch[k] = f_mul(field, ch[k - 1], ch[k - 1]);
  return ch;
}

// Uploads `words` u32 to a fresh device buffer and waits for the copy.
int upload_small(zntt_ctx *c, const std::vector<uint32_t> &words, hipStream_t s, void **out) {
  int st = dev_alloc(c, words.size() * 4, out);
  if (st) return st;
  hipError_t e = hipMemcpyAsync(*out, words.data(), words.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return hip_status(e);
}

// The constants a table build reads, in the dev layout: the root chain's
// first max(L, 1) entries, then one.
std::vector<uint32_t> chain_consts(int field, const std::vector<Elem> &chain, int L) {
  std::vector<uint32_t> consts;
  const int W = zntt::dev_words(field);
  for (int k = 0; k < std::max(L, 1); k++) {
    auto d = f_to_dev(field, chain[k]);
    consts.insert(consts.end(), d.begin(), d.begin() + W);
  }
  auto one = f_to_dev(field, f_one(field));
  consts.insert(consts.end(), one.begin(), one.begin() + W);
  return consts;
}

// Builds the table of powers of a root (2^L entries, see power_table) at
// `table` from its chain constants, uploaded through `cbuf` (consts.size()
// words of device memory), and waits for it.
hipError_t build_powers(int field, int L, const std::vector<uint32_t> &consts, uint32_t *table, uint32_t *cbuf,
                        hipStream_t s) {
  hipError_t he = hipMemcpyAsync(cbuf, consts.data(), consts.size() * 4, hipMemcpyHostToDevice, s);
  const uint32_t *one_dev = cbuf + (size_t)std::max(L, 1) * zntt::dev_words(field);
  if (he == hipSuccess) he = zntt::launch_build_table(field, L, table, cbuf, one_dev, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);  // `consts` is the caller's
  return he;
}

TwiddleEntry *find_power_table(zntt_ctx *c, int field, const Elem &root, int L);

// Device table for the DIT DAG of ntt(root) at size 2^L.  A table built for
// (root', L') with L' >= L and root'^(2^(L'-L)) == root serves the query: stage
// s of both tables is {root'^(j 2^(L'-s))}, so the smaller table is a prefix.
int power_table(zntt_ctx *c, int field, const Elem &root, int L, hipStream_t s, TwiddleEntry **out) {
  for (auto &e : c->tables) {
    if (e.field != field || !e.key.empty() || e.L < L) continue;
    if (e.chain[e.L - L] == root) {
      e.last_use = ++c->tick;
      *out = &e;
      return ZNTT_OK;
    }
  }
  TwiddleEntry e;
  e.field = field;
  e.L = L;
  e.chain = root_chain(field, root, L);
  e.bytes = ((size_t)1 << L) * dev_bytes(field);
  int st = dev_alloc(c, e.bytes, (void **)&e.table);
  if (st) return st;
  const std::vector<uint32_t> consts = chain_consts(field, e.chain, L);
  void *cbuf = nullptr;
  st = dev_alloc(c, consts.size() * 4, &cbuf);
  if (st) {
    dev_free(c, e.table, e.bytes);
    return st;
  }
  hipError_t he = build_powers(field, L, consts, e.table, (uint32_t *)cbuf, s);
  dev_free(c, cbuf, consts.size() * 4);
  if (he != hipSuccess) {
    dev_free(c, e.table, e.bytes);
    return hip_status(he);
  }
  e.last_use = ++c->tick;
  c->tables.push_back(std::move(e));
  *out = &c->tables.back();
  return ZNTT_OK;
}

// The reservation's slot to rebuild: an empty one, else the least recently
// used.
ResvSlot &lru_slot(zntt_ctx *c) {
  ResvSlot *best = &c->resv[0];
  for (ResvSlot &sl : c->resv) {
    if (sl.field < 0) return sl;
    if (sl.last_use < best->last_use) best = &sl;
  }
  return *best;
}

// The table of powers a zig-ntt entry point (ntt, intt) runs over: a cached
// table that serves it, else, on a context whose reservation covers 2^L, one
// built into a reservation slot (no allocation), else a new cache entry.
int call_power_table(zntt_ctx *c, int field, const Elem &root, int L, hipStream_t s, const uint32_t **out) {
  if (TwiddleEntry *e = find_power_table(c, field, root, L)) {
    *out = e->table;
    return ZNTT_OK;
  }
  if (L > c->resv_log_n) {
    TwiddleEntry *e = nullptr;
    int st = power_table(c, field, root, L, s, &e);
    if (st == ZNTT_OK) *out = e->table;
    return st;
  }
  for (ResvSlot &sl : c->resv)
    if (sl.field == field && sl.inv_of == 0 && sl.L >= L && sl.chain[sl.L - L] == root) {
      sl.last_use = ++c->tick;
      *out = sl.table;
      return ZNTT_OK;
    }
  ResvSlot &sl = lru_slot(c);
  sl.field = -1;  // holds nothing until the build succeeds
  std::vector<Elem> chain = root_chain(field, root, L);
  hipError_t he = build_powers(field, L, chain_consts(field, chain, L), sl.table, c->resv_consts, s);
  if (he != hipSuccess) return hip_status(he);
  sl.field = field;
  sl.L = L;
  sl.chain = std::move(chain);
  sl.inv_of = 0;
  sl.last_use = ++c->tick;
  *out = sl.table;
  return ZNTT_OK;
}

// shift^i for i < count (dev layout), cached with prefix reuse.
int coset_table(zntt_ctx *c, int field, const Elem &shift, size_t count, hipStream_t s, const uint32_t **out) {
  for (auto &e : c->cosets) {
    if (e.field == field && e.count >= count && e.shift == shift) {
      e.last_use = ++c->tick;
      *out = e.table;
      return ZNTT_OK;
    }
  }
  CosetEntry e;
  e.field = field;
  e.shift = shift;
  e.count = count;
  e.bytes = count * dev_bytes(field);
  int st = dev_alloc(c, e.bytes, (void **)&e.table);
  if (st) return st;
  const int W = zntt::dev_words(field);
  std::vector<uint32_t> consts;
  Elem p = shift;
  int npows = 0;
  for (; npows < 64 && ((size_t)1 << npows) <= count; npows++) {
    auto d = f_to_dev(field, p);
    consts.insert(consts.end(), d.begin(), d.begin() + W);
    p = f_mul(field, p, p);
  }
  if (npows == 0) {
    auto d = f_to_dev(field, shift);
    consts.insert(consts.end(), d.begin(), d.begin() + W);
    npows = 1;
  }
  auto one = f_to_dev(field, f_one(field));
  consts.insert(consts.end(), one.begin(), one.begin() + W);
  void *cbuf = nullptr;
  st = upload_small(c, consts, s, &cbuf);
  if (st) {
    dev_free(c, e.table, e.bytes);
    return st;
  }
  const uint32_t *pows = (const uint32_t *)cbuf;
  hipError_t he = zntt::launch_powers(field, count, e.table, pows, npows, pows + (size_t)npows * W, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  dev_free(c, cbuf, consts.size() * 4);
  if (he != hipSuccess) {
    dev_free(c, e.table, e.bytes);
    return hip_status(he);
  }
  e.last_use = ++c->tick;
  c->cosets.push_back(std::move(e));
  *out = c->cosets.back().table;
  return ZNTT_OK;
}

// Lookups of the two caches that never build or allocate (the no-allocation
// entry point): the entry power_table / coset_table would return, or null.
TwiddleEntry *find_power_table(zntt_ctx *c, int field, const Elem &root, int L) {
  for (auto &e : c->tables)
    if (e.field == field && e.key.empty() && e.L >= L && e.chain[e.L - L] == root) {
      e.last_use = ++c->tick;
      return &e;
    }
  return nullptr;
}
const uint32_t *find_coset_table(zntt_ctx *c, int field, const Elem &shift, size_t count) {
  for (auto &e : c->cosets)
    if (e.field == field && e.count >= count && e.shift == shift) {
      e.last_use = ++c->tick;
      return e.table;
    }
  return nullptr;
}

int check_transform_args(zntt_ctx *c, int field, const void *data, size_t len, size_t log_n) {
  if (!c || !field_ok(field)) return ZNTT_INVALID_ARGUMENT;
  if (log_n >= 64) return ZNTT_LOG_TOO_LARGE;         // root.zig:33-36
  if (len != ((size_t)1 << log_n)) return ZNTT_LENGTH_MISMATCH;  // root.zig:68
  if (!data) return ZNTT_INVALID_ARGUMENT;
  if (log_n > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  return ZNTT_OK;
}

// Runs one transform over device data on stream s.  `tw` covers >= 2^L - 1
// entries; `coset` (nullable) covers 2^L; `scale` (nullable) is a dev constant.
int run_device(zntt_ctx *c, int field, int L, void *d_data, const uint32_t *tw, const uint32_t *coset,
               const std::array<uint32_t, 9> *scale, hipStream_t s, bool tw_one) {
  if (L == 0) {
    if (!scale && !coset) return ZNTT_OK;
  }
  const zntt::PassPlan plan = zntt::plan_passes(field, L);
  if (plan.passes > 1) {
    int st = ensure(c, &c->scratch, &c->scratch_bytes, ((size_t)1 << L) * dev_bytes(field));
    if (st) return st;
  }
  hipError_t e = zntt::launch_transform(field, L, (uint32_t *)d_data, (uint32_t *)c->scratch, tw, coset,
                                        scale ? scale->data() : nullptr, tw_one, s);
  return hip_status(e);
}

// Host-buffer wrapper: stage data, run, copy back, wait.
template <class Fn>
int with_staging(zntt_ctx *c, int field, void *data, size_t n, Fn &&fn) {
  const size_t bytes = n * user_bytes(field);
  int st = ensure(c, &c->staging, &c->staging_bytes, bytes);
  if (st) return st;
  st = copy_to_device(c, c->staging, data, bytes, c->stream);
  if (!st) st = fn(c->staging);
  if (st) {
    hipStreamSynchronize(c->stream);
    return st;
  }
  return copy_to_host(c, data, c->staging, bytes, c->stream);
}

// Fingerprint of a host twiddle table (stage s holds 2^s elements): a hash
// of sampled entries -- entries 0, 1, the middle and the last of every
// stage, and 64 evenly spaced entries of each of the three largest stages.
// A registered table's host buffers freed and reused for other content (a
// table of another root, data) change it with overwhelming probability.
uint64_t table_fingerprint(int field, size_t log_n, const void *const *stages) {
  const size_t ub = user_bytes(field);
  uint64_t h = 0x9E3779B97F4A7C15ull ^ log_n;
  auto mix = [&](const void *p) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < ub; i += 8) {
      uint64_t w;
      memcpy(&w, b + i, 8);
      h = (h ^ w) * 0xBF58476D1CE4E5B9ull;
      h ^= h >> 31;
    }
  };
  for (size_t s = 0; s < log_n; s++) {
    const size_t len = (size_t)1 << s;
    const uint8_t *st = (const uint8_t *)stages[s];
    mix(st);
    mix(st + (len - 1) * ub);
    if (len > 2) {
      mix(st + ub);
      mix(st + (len / 2) * ub);
    }
    if (s + 3 >= log_n && len > 64)
      for (size_t k = 0; k < 64; k++) mix(st + (k * (len / 64) + k % (len / 64)) * ub);
  }
  return h;
}

// The registered table for these host stages, if their content still
// matches what was registered.  A registration whose buffers now hold other
// content is stale (freed without zntt_free_twiddles and reused): it is
// dropped, and the table is then used as a foreign one.
TwiddleEntry *find_registered(zntt_ctx *c, int field, size_t log_n, const void *const *stages) {
  for (auto it = c->tables.begin(); it != c->tables.end(); ++it) {
    TwiddleEntry &e = *it;
    if (e.field != field || e.key.empty() || (size_t)e.L != log_n) continue;
    bool same = true;
    for (size_t s = 0; s < log_n && same; s++) same = e.key[s] == stages[s];
    if (!same) continue;
    if (e.fingerprint != table_fingerprint(field, log_n, stages)) {
      if (e.pins == 0) {
        hipDeviceSynchronize();  // in-flight work may still read it
        dev_free(c, e.table, e.bytes);
        dev_free(c, e.inv_table, e.inv_bytes);
        c->tables.erase(it);
      } else {
        e.key.clear();  // no longer reachable by its host pointers
      }
      return nullptr;
    }
    e.last_use = ++c->tick;
    return &e;
  }
  return nullptr;
}

// Uploads a caller's table (stage s holds 2^s user-layout elements) into the
// dev layout at c->aux, followed by the trailing one.
int upload_foreign(zntt_ctx *c, int field, size_t log_n, const void *const *stages, hipStream_t s,
                   uint32_t **out) {
  const size_t n = (size_t)1 << log_n;
  const size_t ub = user_bytes(field), db = dev_bytes(field);
  // aux holds: dev table (n entries) + user staging (n entries) + inverse table (n entries)
  int st = ensure(c, &c->aux, &c->aux_bytes, n * (2 * db + ub));
  if (st) return st;
  uint32_t *dev = (uint32_t *)c->aux;
  uint8_t *user = (uint8_t *)c->aux + n * db;
  // Stage by stage from the caller's buffers into the flat user-layout
  // staging (stage k at 2^k - 1), then the trailing one.
  const Elem one = f_one(field);
  for (size_t k = 0; k < log_n && !st; k++)
    st = copy_to_device(c, user + (((size_t)1 << k) - 1) * ub, stages[k], ((size_t)1 << k) * ub, s);
  if (!st) st = copy_to_device(c, user + (n - 1) * ub, one.w, ub, s);
  hipError_t e = st ? hipSuccess : zntt::launch_table_from_user(field, n, dev, (const uint32_t *)user, s);
  const hipError_t e2 = hipStreamSynchronize(s);  // `one` goes out of scope
  if (st) return st;
  *out = dev;
  return hip_status(e != hipSuccess ? e : e2);
}

// Checks a *_with_twiddles call's arguments in the reference's order
// (root.zig:148-161, 180-193): LogTooLarge, LengthMismatch, the stage
// count, then every stage length -- all before any data is touched.
int check_twiddled(zntt_ctx *c, int field, const void *data, size_t len, size_t log_n, const void *const *stages,
                   const size_t *stage_lens, size_t num_stages) {
  int st = check_transform_args(c, field, data, len, log_n);
  if (st) return st;
  if (num_stages != log_n) return ZNTT_INVALID_TWIDDLES;  // root.zig:152
  if (log_n > 0 && (!stages || !stage_lens)) return ZNTT_INVALID_ARGUMENT;
  for (size_t s = 0; s < log_n; s++)  // root.zig:161, checked before data is touched
    if (stage_lens[s] != ((size_t)1 << s)) return ZNTT_INVALID_TWIDDLES;
  for (size_t s = 0; s < log_n; s++)
    if (!stages[s]) return ZNTT_INVALID_ARGUMENT;
  return ZNTT_OK;
}

// The device table a *_with_twiddles call runs over: the registered copy of
// `stages` (its derived inverse table for `inverse`), or the caller's table
// uploaded as given.  `tw_one`: every stage of the result starts with 1.
// Caller holds the context lock; log_n >= 1.
int resolve_twiddles(zntt_ctx *c, int field, size_t log_n, const void *const *stages, bool inverse, hipStream_t s,
                     const uint32_t **tw, bool *tw_one) {
  const int L = (int)log_n;
  // Registered tables are powers of a root, and derived inverse tables start
  // every stage with 1; a foreign forward table is used as given, so stages
  // may skip the multiplication by entry 0 only if each one holds 1.
  *tw_one = true;
  TwiddleEntry *reg = find_registered(c, field, log_n, stages);
  if (reg) {
    *tw = reg->table;
    if (inverse && !reg->inv_table && L <= c->resv_log_n) {
      // A reserved call: the derived inverse goes to a reservation slot.
      for (ResvSlot &sl : c->resv)
        if (sl.field == field && sl.inv_of == reg->reg_id) {
          sl.last_use = ++c->tick;
          *tw = sl.table;
          return ZNTT_OK;
        }
      ResvSlot &sl = lru_slot(c);
      sl.field = -1;
      hipError_t e = zntt::launch_inverse_table(field, L, sl.table, reg->table, s);
      if (e != hipSuccess) return hip_status(e);
      sl.field = field;
      sl.L = L;
      sl.chain.clear();
      sl.inv_of = reg->reg_id;
      sl.last_use = ++c->tick;
      *tw = sl.table;
      return ZNTT_OK;
    }
    if (inverse) {
      if (!reg->inv_table) {
        reg->inv_bytes = reg->bytes;
        int st = dev_alloc(c, reg->inv_bytes, (void **)&reg->inv_table);
        if (st) return st;
        hipError_t e = zntt::launch_inverse_table(field, L, reg->inv_table, reg->table, s);
        if (e != hipSuccess) return hip_status(e);
      }
      *tw = reg->inv_table;
    }
    return ZNTT_OK;
  }
  uint32_t *dev = nullptr;
  int st = upload_foreign(c, field, log_n, stages, s, &dev);
  if (st) return st;
  *tw = dev;
  const Elem one = f_one(field);
  for (size_t k = 0; k < log_n && !inverse; k++)
    if (!(load_elem(field, stages[k]) == one)) *tw_one = false;
  if (inverse) {
    const size_t n = (size_t)1 << log_n;
    uint32_t *inv = (uint32_t *)((uint8_t *)c->aux + n * (dev_bytes(field) + user_bytes(field)));
    hipError_t e = zntt::launch_inverse_table(field, L, inv, dev, s);
    if (e != hipSuccess) return hip_status(e);
    *tw = inv;
  }
  return ZNTT_OK;
}

int twiddled(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *const *stages,
             const size_t *stage_lens, size_t num_stages, bool inverse) {
  int st = check_twiddled(c, field, data, len, log_n, stages, stage_lens, num_stages);
  if (st) return st;
  CallGuard g(c, c->stream);
  if (log_n == 0) return ZNTT_OK;
  const uint32_t *tw = nullptr;
  bool tw_one = true;
  st = resolve_twiddles(c, field, log_n, stages, inverse, c->stream, &tw, &tw_one);
  if (st) return st;
  std::array<uint32_t, 9> scale;
  if (inverse) scale = f_to_dev(field, f_inv(field, f_from_u64(field, len)));  // root.zig:208-211
  return with_staging(c, field, data, len, [&](void *d) {
    return run_device(c, field, (int)log_n, d, tw, nullptr, inverse ? &scale : nullptr, c->stream, tw_one);
  });
}

}  // namespace zntt_detail

// ---------------------------------------------------------------------------
// Exported entry points
// ---------------------------------------------------------------------------
extern "C" {

int zntt_ctx_create(zntt_ctx **out, int device, uint64_t hbm_budget_bytes) {
  if (!out) return ZNTT_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    hipGetLastError();
    return ZNTT_NO_DEVICE;
  }
  if (device < 0 || device >= count) return ZNTT_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return ZNTT_DEVICE_ERROR;
  auto *c = new zntt_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (64ull << 30);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ZNTT_DEVICE_ERROR;
  }
  if (hipEventCreateWithFlags(&c->shared_done, hipEventDisableTiming) != hipSuccess) {
    hipStreamDestroy(c->stream);
    delete c;
    return ZNTT_DEVICE_ERROR;
  }
  // The no-allocation entry point's scratch on this device (looked up once
  // per device, with the device current).
  if (zntt::noalloc_scratch_init(device) != hipSuccess) {
    hipGetLastError();
    hipEventDestroy(c->shared_done);
    hipStreamDestroy(c->stream);
    delete c;
    return ZNTT_DEVICE_ERROR;
  }
  *out = c;
  return ZNTT_OK;
}

int zntt_ctx_reserve(zntt_ctx *c, size_t max_log_n) {
  if (!c) return ZNTT_INVALID_ARGUMENT;
  if (max_log_n > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  CallGuard g(c);
  if ((int)max_log_n <= c->resv_log_n) return ZNTT_OK;
  // The largest element of any field, in the caller's layout and on the device.
  size_t ub = 0, db = 0;
  for (int f = 0; f < 16; f++)
    if (field_ok(f)) {
      ub = std::max(ub, user_bytes(f));
      db = std::max(db, dev_bytes(f));
    }
  const size_t n = (size_t)1 << max_log_n;
  // Buffers of the calls at 2^max_log_n: staged data, pass scratch, and a
  // foreign table's upload (dev table, user staging, derived inverse).
  int st = ensure(c, &c->staging, &c->staging_bytes, n * ub);
  if (!st) st = ensure(c, &c->scratch, &c->scratch_bytes, n * db);
  if (!st) st = ensure(c, &c->aux, &c->aux_bytes, n * (2 * db + ub));
  const size_t consts = (size_t)(ZNTT_MAX_LOG_N + 2) * 9 * 4;
  if (!st && !c->resv_consts) {
    st = dev_alloc(c, consts, (void **)&c->resv_consts);
    if (!st) c->resv_consts_bytes = consts;
  }
  // The page-locked ring of large host-buffer copies (optional: without it
  // the calls copy from pageable memory, more slowly).
  if (!st && ring_wanted(n * ub)) (void)ring_ensure(c);
  for (ResvSlot &sl : c->resv) {
    if (st || sl.bytes >= n * db) continue;
    uint32_t *t = nullptr;
    st = dev_alloc(c, n * db, (void **)&t);
    if (st) break;
    hipStreamSynchronize(c->stream);
    dev_free(c, sl.table, sl.bytes);
    sl = ResvSlot();
    sl.table = t;
    sl.bytes = n * db;
  }
  if (st) return st;
  c->resv_log_n = (int)max_log_n;
  return ZNTT_OK;
}

int zntt_ctx_trim(zntt_ctx *c) {
  if (!c) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c);
  release_cache(c);
  return ZNTT_OK;
}

void zntt_ctx_destroy(zntt_ctx *c) {
  if (!c) return;
  {
    CallGuard g(c);
    release_cache(c);
  }
  ctx_release(c);
}

uint64_t zntt_ctx_bytes_in_use(const zntt_ctx *c) { return c ? c->in_use : 0; }

void *zntt_ctx_stream(zntt_ctx *c) { return c ? (void *)c->stream : nullptr; }

int zntt_ctx_synchronize(zntt_ctx *c) {
  if (!c) return ZNTT_INVALID_ARGUMENT;
  return hip_status(hipStreamSynchronize(c->stream));
}

const char *zntt_status_string(int status) {
  switch (status) {
    case ZNTT_OK: return "ok";
    case ZNTT_LOG_TOO_LARGE: return "LogTooLarge";
    case ZNTT_LENGTH_MISMATCH: return "LengthMismatch";
    case ZNTT_INVALID_LENGTH: return "InvalidLength";
    case ZNTT_INVALID_TWIDDLES: return "InvalidTwiddles";
    case ZNTT_OUT_OF_MEMORY: return "OutOfMemory";
    case ZNTT_DEVICE_ERROR: return "DeviceError";
    case ZNTT_NO_DEVICE: return "NoDevice";
    case ZNTT_INVALID_ARGUMENT: return "InvalidArgument";
    case ZNTT_UNSUPPORTED: return "Unsupported";
    case ZNTT_ORDER_TOO_LARGE: return "OrderTooLarge";
    case ZNTT_DEGREE_TOO_LARGE: return "DegreeTooLarge";
    case ZNTT_INVALID_PARAMETERS: return "InvalidParameters";
    case ZNTT_NOT_PREPARED: return "NotPrepared";
    default: return "unknown";
  }
}

size_t zntt_element_size(int field) { return field_ok(field) ? user_bytes(field) : 0; }

int zntt_primitive_root(int field, size_t log_size, void *out) {
  if (!field_ok(field) || !out) return ZNTT_INVALID_ARGUMENT;
  if (log_size > (size_t)two_adicity(field)) return ZNTT_ORDER_TOO_LARGE;
  const Elem r = f_primitive_root(field, (int)log_size);
  memcpy(out, r.w, user_bytes(field));
  return ZNTT_OK;
}

int zntt_bit_reverse(zntt_ctx *c, int field, void *data, size_t len) {
  if (!c || !field_ok(field)) return ZNTT_INVALID_ARGUMENT;
  if (len == 0 || (len & (len - 1)) != 0) return ZNTT_INVALID_LENGTH;  // root.zig:43
  if (!data) return ZNTT_INVALID_ARGUMENT;
  const int L = __builtin_ctzll(len);
  if (L > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  CallGuard g(c, c->stream);
  return with_staging(c, field, data, len, [&](void *d) {
    return hip_status(zntt::launch_bit_reverse(field, L, (uint32_t *)d, c->stream));
  });
}

int zntt_ntt(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *root) {
  int st = check_transform_args(c, field, data, len, log_n);
  if (st) return st;
  if (!root) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c, c->stream);
  if (log_n == 0) return ZNTT_OK;
  const uint32_t *tw = nullptr;
  st = call_power_table(c, field, load_elem(field, root), (int)log_n, c->stream, &tw);
  if (st) return st;
  return with_staging(c, field, data, len, [&](void *d) {
    return run_device(c, field, (int)log_n, d, tw, nullptr, nullptr, c->stream);
  });
}

int zntt_intt(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *root) {
  int st = check_transform_args(c, field, data, len, log_n);
  if (st) return st;
  if (!root) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c, c->stream);
  if (log_n == 0) return ZNTT_OK;  // n^-1 = 1
  const Elem root_inv = f_inv(field, load_elem(field, root));  // root.zig:100
  const uint32_t *tw = nullptr;
  st = call_power_table(c, field, root_inv, (int)log_n, c->stream, &tw);
  if (st) return st;
  const std::array<uint32_t, 9> scale = f_to_dev(field, f_inv(field, f_from_u64(field, len)));  // root.zig:103
  return with_staging(c, field, data, len, [&](void *d) {
    return run_device(c, field, (int)log_n, d, tw, nullptr, &scale, c->stream);
  });
}

int zntt_precompute_twiddles(zntt_ctx *c, int field, size_t log_n, const void *root, void *const *stages,
                             size_t num_stages) {
  if (!c || !field_ok(field)) return ZNTT_INVALID_ARGUMENT;
  if (log_n >= 64) return ZNTT_LOG_TOO_LARGE;  // root.zig:114
  if (num_stages != log_n || !root || (log_n > 0 && !stages)) return ZNTT_INVALID_ARGUMENT;
  for (size_t s = 0; s < log_n; s++)
    if (!stages[s]) return ZNTT_INVALID_ARGUMENT;
  if (log_n > ZNTT_MAX_LOG_N) return ZNTT_OUT_OF_MEMORY;
  CallGuard g(c, c->stream);
  if (log_n == 0) return ZNTT_OK;
  const int L = (int)log_n;
  // Drop a stale registration of the same host buffers.
  if (TwiddleEntry *old = find_registered(c, field, log_n, (const void *const *)stages)) {
    dev_free(c, old->table, old->bytes);
    dev_free(c, old->inv_table, old->inv_bytes);
    c->tables.remove_if([&](const TwiddleEntry &e) { return &e == old; });
  }
  TwiddleEntry *src = nullptr;
  int st = power_table(c, field, load_elem(field, root), L, c->stream, &src);
  if (st) return st;
  // A registered entry owns its own copy of the table (the cache may evict
  // power tables; registered ones live until zntt_free_twiddles).
  TwiddleEntry e;
  e.field = field;
  e.L = L;
  e.bytes = ((size_t)1 << L) * dev_bytes(field);
  st = dev_alloc(c, e.bytes, (void **)&e.table);
  if (st) return st;
  const size_t ub = user_bytes(field);
  const size_t n = (size_t)1 << L;
  st = ensure(c, &c->staging, &c->staging_bytes, n * ub);
  if (st) {
    dev_free(c, e.table, e.bytes);
    return st;
  }
  // src may cover a larger size: copy the prefix of 2^L - 1 entries, then one.
  hipError_t he = hipMemcpyAsync(e.table, src->table, (n - 1) * dev_bytes(field), hipMemcpyDeviceToDevice, c->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(e.table + (n - 1) * zntt::dev_words(field),
                        src->table + (((size_t)1 << src->L) - 1) * zntt::dev_words(field), dev_bytes(field),
                        hipMemcpyDeviceToDevice, c->stream);
  if (he == hipSuccess) he = zntt::launch_table_to_user(field, n - 1, (uint32_t *)c->staging, e.table, c->stream);
  // Stage s of the flat user-layout table (at 2^s - 1) straight into the
  // caller's stage buffer.
  int cst = hip_status(he);
  for (size_t s = 0; s < log_n && !cst; s++)
    cst = copy_to_host(c, stages[s], (uint8_t *)c->staging + (((size_t)1 << s) - 1) * ub, ((size_t)1 << s) * ub,
                       c->stream);
  if (cst) {
    hipStreamSynchronize(c->stream);
    dev_free(c, e.table, e.bytes);
    return cst;
  }
  e.key.assign(stages, stages + log_n);
  e.fingerprint = table_fingerprint(field, log_n, (const void *const *)stages);
  e.reg_id = c->next_reg_id++;
  e.last_use = ++c->tick;
  c->tables.push_back(std::move(e));
  return ZNTT_OK;
}

int zntt_free_twiddles(zntt_ctx *c, const void *const *stages, size_t num_stages) {
  if (!c) return ZNTT_INVALID_ARGUMENT;
  if (num_stages == 0) return ZNTT_OK;
  if (!stages) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c);
  for (auto it = c->tables.begin(); it != c->tables.end(); ++it) {
    if (it->key.size() != num_stages) continue;
    if (!std::equal(it->key.begin(), it->key.end(), stages)) continue;
    hipStreamSynchronize(c->stream);
    dev_free(c, it->table, it->bytes);
    dev_free(c, it->inv_table, it->inv_bytes);
    c->tables.erase(it);
    return ZNTT_OK;
  }
  return ZNTT_OK;  // not registered (e.g. built elsewhere): nothing on the device
}

int zntt_ntt_with_twiddles(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *const *stages,
                           const size_t *stage_lens, size_t num_stages) {
  return twiddled(c, field, data, len, log_n, stages, stage_lens, num_stages, false);
}

int zntt_intt_with_twiddles(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *const *stages,
                            const size_t *stage_lens, size_t num_stages) {
  return twiddled(c, field, data, len, log_n, stages, stage_lens, num_stages, true);
}

// nttWithTwiddles / inttWithTwiddles on device-resident data, enqueued on
// `stream`: the same checks, tables and DAG as the host-buffer calls.
static int twiddled_device(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *const *stages,
                           const size_t *stage_lens, size_t num_stages, bool inverse, void *stream) {
  int st = check_twiddled(c, field, data, len, log_n, stages, stage_lens, num_stages);
  if (st) return st;
  const hipStream_t s = (hipStream_t)stream;
  CallGuard g(c, s);
  if (log_n == 0) return ZNTT_OK;
  const uint32_t *tw = nullptr;
  bool tw_one = true;
  st = resolve_twiddles(c, field, log_n, stages, inverse, s, &tw, &tw_one);
  if (st) return st;
  std::array<uint32_t, 9> scale;
  if (inverse) scale = f_to_dev(field, f_inv(field, f_from_u64(field, len)));  // root.zig:208-211
  return run_device(c, field, (int)log_n, data, tw, nullptr, inverse ? &scale : nullptr, s, tw_one);
}

int zntt_ntt_with_twiddles_device(zntt_ctx *c, int field, void *data, size_t len, size_t log_n,
                                  const void *const *stages, const size_t *stage_lens, size_t num_stages,
                                  void *stream) {
  return twiddled_device(c, field, data, len, log_n, stages, stage_lens, num_stages, false, stream);
}

int zntt_intt_with_twiddles_device(zntt_ctx *c, int field, void *data, size_t len, size_t log_n,
                                   const void *const *stages, const size_t *stage_lens, size_t num_stages,
                                   void *stream) {
  return twiddled_device(c, field, data, len, log_n, stages, stage_lens, num_stages, true, stream);
}

int zntt_coset_ntt(zntt_ctx *c, int field, void *data, size_t len, size_t log_n, const void *root,
                   const void *shift) {
  int st = check_transform_args(c, field, data, len, log_n);
  if (st) return st;
  if (!root || !shift) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c, c->stream);
  TwiddleEntry *t = nullptr;
  if (log_n > 0) {
    st = power_table(c, field, load_elem(field, root), (int)log_n, c->stream, &t);
    if (st) return st;
  }
  const uint32_t *cos = nullptr;
  st = coset_table(c, field, load_elem(field, shift), len, c->stream, &cos);
  if (st) return st;
  if (log_n == 0) return ZNTT_OK;  // data[0] *= shift^0
  return with_staging(c, field, data, len, [&](void *d) {
    return run_device(c, field, (int)log_n, d, t->table, cos, nullptr, c->stream);
  });
}

}  // extern "C"

namespace {

// Degree + 1 of a coefficient array: its length without trailing zeros.
size_t effective_len(int field, const void *p, size_t n) {
  const size_t ub = user_bytes(field);
  const uint8_t *b = (const uint8_t *)p;
  while (n > 0) {
    const uint8_t *e = b + (n - 1) * ub;
    bool zero = true;
    for (size_t i = 0; i < ub && zero; i++) zero = e[i] == 0;  // zero is all-zero bytes in both layouts
    if (!zero) break;
    n--;
  }
  return n;
}

// Product of a (na coefficients) and b (nb) into out (na + nb - 1), all on
// the device, on stream s: pad to 2^L, forward both, pointwise product,
// inverse with the n^-1 scale (times 2^5 for BN254 Fr's data x data
// Montgomery product).  Caller holds the context lock.
int poly_mul_device(zntt_ctx *c, int field, const void *a, size_t na, const void *b, size_t nb, void *out,
                    hipStream_t s) {
  const size_t m = na + nb - 1;
  int L = 1;  // two constants are multiplied as a 2-point cyclic product
  while (((size_t)1 << L) < m) L++;
  if (L > two_adicity(field)) return ZNTT_ORDER_TOO_LARGE;
  if (L > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  const size_t n = (size_t)1 << L, ub = user_bytes(field);
  int st = ensure(c, &c->staging, &c->staging_bytes, 2 * n * ub);
  if (st) return st;
  uint8_t *A = (uint8_t *)c->staging, *B = A + n * ub;
  hipError_t e = hipMemcpyAsync(A, a, na * ub, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(A + na * ub, 0, (n - na) * ub, s);
  if (e == hipSuccess) e = hipMemcpyAsync(B, b, nb * ub, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(B + nb * ub, 0, (n - nb) * ub, s);
  if (e != hipSuccess) return hip_status(e);
  const Elem root = f_primitive_root(field, L);
  std::array<uint32_t, 9> scale;
  Elem sc = f_inv(field, f_from_u64(field, n));
  // The pointwise Montgomery product of two data values carries a factor
  // 2^-5 for BN254 Fr (data in 2^256 form, R' = 2^261) and 2^-32 for the
  // 31-bit fields (plain data, R = 2^32); Goldilocks multiplies exactly.
  if (field == ZNTT_FIELD_BN254_FR) sc = f_mul(field, sc, f_from_u64(field, 32));
  if (field == ZNTT_FIELD_BABYBEAR || field == ZNTT_FIELD_KOALABEAR)
    sc = f_mul(field, sc, f_from_u64(field, (uint64_t)1 << 32));
  scale = f_to_dev(field, sc);
  {
    TwiddleEntry *fw = nullptr, *inv = nullptr;
    st = power_table(c, field, root, L, s, &fw);
    if (st) return st;
    fw->pins++;
    st = power_table(c, field, f_inv(field, root), L, s, &inv);
    fw->pins--;
    if (st) return st;
    st = run_device(c, field, L, A, fw->table, nullptr, nullptr, s);
    if (st == ZNTT_OK) st = run_device(c, field, L, B, fw->table, nullptr, nullptr, s);
    if (st) return st;
    e = zntt::launch_pointwise_mul(field, (uint32_t *)A, (const uint32_t *)B, n, s);
    if (e != hipSuccess) return hip_status(e);
    st = run_device(c, field, L, A, inv->table, nullptr, &scale, s);
    if (st) return st;
  }
  return hip_status(hipMemcpyAsync(out, A, m * ub, hipMemcpyDeviceToDevice, s));
}

}  // namespace

extern "C" {

int zntt_lde(zntt_ctx *c, int field, const void *coeffs, size_t log_k, size_t log_blowup, const void *shift,
             void *out) {
  if (!c || !field_ok(field) || !coeffs || !shift || !out) return ZNTT_INVALID_ARGUMENT;
  if (log_k >= 64 || log_blowup >= 64 || log_k + log_blowup >= 64) return ZNTT_LOG_TOO_LARGE;
  const size_t L = log_k + log_blowup;
  if (L > (size_t)two_adicity(field)) return ZNTT_ORDER_TOO_LARGE;
  if (L > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  const size_t k = (size_t)1 << log_k, n = (size_t)1 << L, ub = user_bytes(field);
  CallGuard g(c, c->stream);
  TwiddleEntry *t = nullptr;
  const Elem root = f_primitive_root(field, (int)L);
  int st = ZNTT_OK;
  if (L > 0) {
    st = power_table(c, field, root, (int)L, c->stream, &t);
    if (st) return st;
  }
  const uint32_t *cos = nullptr;
  st = coset_table(c, field, load_elem(field, shift), n, c->stream, &cos);
  if (st) return st;
  st = ensure(c, &c->staging, &c->staging_bytes, n * ub);
  if (st) return st;
  hipError_t e = hipMemsetAsync((uint8_t *)c->staging + k * ub, 0, (n - k) * ub, c->stream);
  if (e != hipSuccess) return hip_status(e);
  st = copy_to_device(c, c->staging, coeffs, k * ub, c->stream);
  if (st) return st;
  if (L == 0) {
    // f(shift) with a single coefficient is the coefficient itself.
    memcpy(out, coeffs, ub);
    return hip_status(hipStreamSynchronize(c->stream));
  }
  st = run_device(c, field, (int)L, c->staging, t->table, cos, nullptr, c->stream);
  if (st) return st;
  return copy_to_host(c, out, c->staging, n * ub, c->stream);
}

int zntt_poly_mul(zntt_ctx *c, int field, const void *a, size_t a_len, const void *b, size_t b_len, void *out,
                  size_t out_cap, size_t *out_len) {
  if (!c || !field_ok(field) || !out_len || (!a && a_len) || (!b && b_len)) return ZNTT_INVALID_ARGUMENT;
  *out_len = 0;
  const size_t na = effective_len(field, a, a_len), nb = effective_len(field, b, b_len);
  if (na == 0 || nb == 0) return ZNTT_OK;  // the zero polynomial, poly.zig:198
  const size_t m = na + nb - 1;
  if (m > out_cap) return ZNTT_DEGREE_TOO_LARGE;  // poly.zig:199-200
  if (!out) return ZNTT_INVALID_ARGUMENT;
  const size_t ub = user_bytes(field);
  CallGuard g(c, c->stream);
  void *dev = nullptr;
  int st = ensure(c, &c->aux, &c->aux_bytes, (na + nb + m) * ub);
  if (st) return st;
  dev = c->aux;
  uint8_t *da = (uint8_t *)dev, *db = da + na * ub, *dout = db + nb * ub;
  st = copy_to_device(c, da, a, na * ub, c->stream);
  if (!st) st = copy_to_device(c, db, b, nb * ub, c->stream);
  if (!st) st = poly_mul_device(c, field, da, na, db, nb, dout, c->stream);
  if (!st) st = copy_to_host(c, out, dout, m * ub, c->stream);
  if (st) {
    hipStreamSynchronize(c->stream);
    return st;
  }
  *out_len = m;
  return ZNTT_OK;
}

int zntt_poly_mul_device(zntt_ctx *c, int field, const void *a, size_t a_len, const void *b, size_t b_len,
                         void *out, void *stream) {
  if (!c || !field_ok(field) || !a || !b || !out || a_len == 0 || b_len == 0) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c, (hipStream_t)stream);
  return poly_mul_device(c, field, a, a_len, b, b_len, out, (hipStream_t)stream);
}

int zntt_execute_device(zntt_ctx *c, const zntt_transform *t, void *stream) {
  if (!c || !t || !field_ok(t->field) || !t->root) return ZNTT_INVALID_ARGUMENT;
  if (t->log_n >= 64) return ZNTT_LOG_TOO_LARGE;
  if (t->log_n > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  if (!t->data) return ZNTT_INVALID_ARGUMENT;
  if (t->direction == ZNTT_COSET_FORWARD && !t->shift) return ZNTT_INVALID_ARGUMENT;
  if (t->direction < ZNTT_FORWARD || t->direction > ZNTT_COSET_FORWARD) return ZNTT_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;  // HIP's convention: NULL is the null stream
  CallGuard g(c, s);
  const int field = t->field, L = (int)t->log_n;
  const size_t n = (size_t)1 << L;
  Elem root = load_elem(field, t->root);
  std::array<uint32_t, 9> scale;
  const std::array<uint32_t, 9> *scale_p = nullptr;
  if (t->direction == ZNTT_INVERSE) {
    root = f_inv(field, root);
    scale = f_to_dev(field, f_inv(field, f_from_u64(field, n)));
    scale_p = &scale;
  }
  TwiddleEntry *tw = nullptr;
  int st = ZNTT_OK;
  if (L > 0) {
    st = power_table(c, field, root, L, s, &tw);
    if (st) return st;
  }
  const uint32_t *cos = nullptr;
  if (t->direction == ZNTT_COSET_FORWARD) {
    st = coset_table(c, field, load_elem(field, t->shift), n, s, &cos);
    if (st) return st;
  }
  if (L == 0) return ZNTT_OK;
  return run_device(c, field, L, t->data, tw->table, cos, scale_p, s);
}

int zntt_execute_device_noalloc(zntt_ctx *c, const zntt_transform *t, void *stream) {
  if (!c || !t || !field_ok(t->field) || !t->root) return ZNTT_INVALID_ARGUMENT;
  if (t->log_n >= 64) return ZNTT_LOG_TOO_LARGE;
  if (t->log_n > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  if (!t->data) return ZNTT_INVALID_ARGUMENT;
  if (t->direction == ZNTT_COSET_FORWARD && !t->shift) return ZNTT_INVALID_ARGUMENT;
  if (t->direction < ZNTT_FORWARD || t->direction > ZNTT_COSET_FORWARD) return ZNTT_INVALID_ARGUMENT;
  const int field = t->field, L = (int)t->log_n;
  const size_t n = (size_t)1 << L;
  // Scratch need of a multi-pass transform against the static buffer.
  const bool multi = L > 0 && zntt::plan_passes(field, L).passes > 1;
  if (multi && n * dev_bytes(field) > (size_t)ZNTT_NOALLOC_SCRATCH_BYTES) return ZNTT_OUT_OF_MEMORY;
  uint32_t *scratch = (uint32_t *)zntt::noalloc_scratch_ptr(c->device);
  if (!scratch) return ZNTT_DEVICE_ERROR;
  const hipStream_t s = (hipStream_t)stream;
  CallGuard g(c);  // the cache's lock; the static scratch is this call's alone only on one stream
  Elem root = load_elem(field, t->root);
  std::array<uint32_t, 9> scale;
  const uint32_t *scale_p = nullptr;
  if (t->direction == ZNTT_INVERSE) {
    root = f_inv(field, root);
    scale = f_to_dev(field, f_inv(field, f_from_u64(field, n)));
    scale_p = scale.data();
  }
  const TwiddleEntry *tw = nullptr;
  if (L > 0 && !(tw = find_power_table(c, field, root, L))) return ZNTT_NOT_PREPARED;
  const uint32_t *cos = nullptr;
  if (t->direction == ZNTT_COSET_FORWARD && !(cos = find_coset_table(c, field, load_elem(field, t->shift), n)))
    return ZNTT_NOT_PREPARED;
  if (L == 0) return ZNTT_OK;
  return hip_status(zntt::launch_transform(field, L, (uint32_t *)t->data, scratch, tw->table, cos, scale_p, true, s));
}

int zntt_debug_field_op(zntt_ctx *c, int field, int op, const void *a, const void *b, void *out, size_t n) {
  if (!c || !field_ok(field) || !a || !b || !out || op < 0 || op > 2) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c, c->stream);
  const size_t bytes = n * user_bytes(field);
  void *buf = nullptr;
  int st = dev_alloc(c, 3 * bytes, &buf);
  if (st) return st;
  uint8_t *da = (uint8_t *)buf, *db = da + bytes, *dout = db + bytes;
  hipError_t e = hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = zntt::launch_field_op(field, op, n, (const uint32_t *)da, (const uint32_t *)db, (uint32_t *)dout, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dev_free(c, buf, 3 * bytes);
  return hip_status(e);
}

int zntt_fill_uniform(int field, uint64_t seed, void *out, size_t n) {
  if (!field_ok(field) || (!out && n)) return ZNTT_INVALID_ARGUMENT;
  // SplitMix64 expands the seed into the xoshiro256++ state.
  uint64_t sm = seed, st[4];
  for (int i = 0; i < 4; i++) {
    uint64_t z = (sm += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    st[i] = z ^ (z >> 31);
  }
  auto next = [&st]() {
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    const uint64_t r = rotl(st[0] + st[3], 23) + st[0];
    const uint64_t t = st[1] << 17;
    st[2] ^= st[0];
    st[3] ^= st[1];
    st[1] ^= st[2];
    st[0] ^= st[3];
    st[2] ^= t;
    st[3] = rotl(st[3], 45);
    return r;
  };
  if (field == ZNTT_FIELD_BN254_FR) {
    U256 *o = (U256 *)out;
    for (size_t i = 0; i < n;) {
      U256 v;
      for (int k = 0; k < 4; k++) v.l[k] = next();
      if (zntt::host::Fr::lt(v, zntt::host::Fr::P)) o[i++] = zntt::host::Fr::from_canonical(v);
    }
  } else {
    // ceil(bits / 8) little-endian bytes of each draw, rejected at or above
    // p (field.zig:224-232): 8 bytes for Goldilocks, 4 for the 31-bit fields.
    const auto *f = f31(field);
    const uint64_t p = f ? f->p : zntt::host::Gl::P;
    const uint64_t mask = f ? 0xFFFFFFFFull : ~0ull;
    uint64_t *o = (uint64_t *)out;
    for (size_t i = 0; i < n;) {
      const uint64_t v = next() & mask;
      if (v < p) o[i++] = v;
    }
  }
  return ZNTT_OK;
}

int zntt_probe_peaks(zntt_ctx *c, double *mad_per_s, double *copy_bytes_per_s) {
  if (!c || !mad_per_s || !copy_bytes_per_s) return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c, c->stream);
  return hip_status(zntt::probe_peaks(mad_per_s, copy_bytes_per_s, c->stream));
}

}  // extern "C"
