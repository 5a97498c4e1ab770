// Context internals shared by the C ABI translation units (zntt_api.cpp,
// zntt_batch.cpp): the context, the device twiddle cache, budgeted device
// memory and host-side field helpers on storage-layout elements.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <array>
#include <atomic>
#include <list>
#include <mutex>
#include <vector>

#include "host_field.hpp"
#include "kernels.hpp"
#include "zntt_gpu.h"

namespace zntt_detail {
// One element in the caller's storage layout (Fr 8 words, Goldilocks 2).
struct Elem {
  uint32_t w[8];
  bool operator==(const Elem &o) const { return memcmp(w, o.w, sizeof w) == 0; }
};
struct HostRing;
}  // namespace zntt_detail

struct TwiddleEntry {
  int field = 0;
  int L = 0;
  std::vector<zntt_detail::Elem> chain;           // chain[k] = root^(2^k), k = 0..L (powers tables)
  uint32_t *table = nullptr;         // dev layout, 2^L entries (stages + trailing one)
  size_t bytes = 0;
  uint32_t *inv_table = nullptr;     // inverse-with-twiddles table, built on demand
  size_t inv_bytes = 0;
  std::vector<const void *> key;     // host stage pointers of a registered table
  uint64_t fingerprint = 0;          // of sampled entries of a registered table's host stages
  uint64_t reg_id = 0;               // a registered table's number, unique in its context
  uint64_t last_use = 0;
  int pins = 0;                      // live batch plans reading this table
};

// A table slot of a reservation (zntt_ctx_reserve): device storage for one
// table of up to 2^resv_log_n entries of any field.  The reserved calls build
// the tables they need into the least recently used slot instead of
// allocating; nothing else reads the slots, so no batch or server pins them.
struct ResvSlot {
  uint32_t *table = nullptr;
  size_t bytes = 0;
  int field = -1;                                // -1: holds no table
  int L = 0;
  std::vector<zntt_detail::Elem> chain;          // a table of powers: chain[k] = root^(2^k)
  uint64_t inv_of = 0;                           // nonzero: inverse of the registered table reg_id
  uint64_t last_use = 0;
};

struct CosetEntry {
  int field = 0;
  zntt_detail::Elem shift;
  size_t count = 0;
  uint32_t *table = nullptr;  // dev layout, shift^i for i < count
  size_t bytes = 0;
  uint64_t last_use = 0;
  int pins = 0;
};

struct zntt_ctx {
  // References: one for the caller (dropped by zntt_ctx_destroy) and one per
  // live batch or server made on the context.  The context, its stream and
  // its cache are freed when the last reference goes, so children may
  // outlive the caller's handle.
  std::atomic<int> refs{1};
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  uint64_t tick = 0;
  uint64_t protect_from = 0;  // entries used since this tick belong to the running call
  std::mutex mu;
  std::list<TwiddleEntry> tables;
  std::list<CosetEntry> cosets;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  void *staging = nullptr;
  size_t staging_bytes = 0;
  void *aux = nullptr;  // foreign tables and other per-call temporaries
  size_t aux_bytes = 0;
  void *pinned = nullptr;  // page-locked host staging (small read-backs and uploads)
  size_t pinned_bytes = 0;
  zntt_detail::HostRing *ring = nullptr;  // large host-buffer copies (host_copy.cpp)
  bool ring_failed = false;               // its setup failed: plain copies until the next trim
  // The eight-lane M31 transform's twiddle table for (m31_root, m31_log_n);
  // kept so that repeated calls with one root skip the rebuild.
  void *m31_table = nullptr;
  size_t m31_table_bytes = 0;
  uint64_t m31_root = 0;
  int m31_log_n = -1;
  // Reservation (zntt_ctx_reserve): the seven zig-ntt entry points at sizes
  // up to 2^resv_log_n run in the reusable buffers above, pre-sized, and in
  // these slots, and allocate nothing.  -1: no reservation.
  int resv_log_n = -1;
  ResvSlot resv[2];
  uint32_t *resv_consts = nullptr;  // a table build's root chain upload
  size_t resv_consts_bytes = 0;
  uint64_t next_reg_id = 1;
  // Ordering of the shared buffers' users across streams (CallGuard).
  hipEvent_t shared_done = nullptr;
  hipStream_t shared_stream = nullptr;
  bool shared_used = false;
};

namespace zntt_detail {

// Holds the context lock for one API call and shields the cache entries that
// call touches from eviction by its own later allocations.
//
// Given the stream the call enqueues its work on, it also orders the call's
// use of the context's shared device buffers (scratch, staging, aux, the
// eight-lane M31 table): when the previous such call enqueued on another
// stream, this stream first waits for the event that call recorded, and
// this call records the event again when it ends.  So asynchronous calls on
// different streams never overlap in those buffers, and a call on the same
// stream as the last one adds only the record.
struct CallGuard {
  std::lock_guard<std::mutex> lock;
  zntt_ctx *ctx;
  hipStream_t stream = nullptr;
  bool ordered = false;
  explicit CallGuard(zntt_ctx *c) : lock(c->mu), ctx(c) { c->protect_from = c->tick + 1; }
  CallGuard(zntt_ctx *c, hipStream_t s) : lock(c->mu), ctx(c), stream(s), ordered(true) {
    c->protect_from = c->tick + 1;
    if (c->shared_used && c->shared_stream != s) (void)hipStreamWaitEvent(s, c->shared_done, 0);
  }
  ~CallGuard() {
    if (!ordered) return;
    (void)hipEventRecord(ctx->shared_done, stream);
    ctx->shared_stream = stream;
    ctx->shared_used = true;
  }
};

// Reference counting of the context (see zntt_ctx::refs).
void ctx_retain(zntt_ctx *c);
void ctx_release(zntt_ctx *c);

bool field_ok(int field);
size_t user_bytes(int field);
size_t dev_bytes(int field);
Elem load_elem(int field, const void *p);
Elem f_mul(int field, const Elem &a, const Elem &b);
Elem f_inv(int field, const Elem &a);
Elem f_from_u64(int field, uint64_t v);
Elem f_one(int field);
Elem f_primitive_root(int field, int log_size);
int two_adicity(int field);
std::array<uint32_t, 9> f_to_dev(int field, const Elem &a);

int hip_status(hipError_t e);
void dev_free(zntt_ctx *c, void *p, size_t bytes);
int dev_alloc(zntt_ctx *c, size_t bytes, void **out);
int ensure(zntt_ctx *c, void **buf, size_t *have, size_t bytes);
// Grows the context's page-locked host buffer to at least `bytes`.
int ensure_pinned(zntt_ctx *c, size_t bytes);
// Host-buffer transfers on `stream` (host_copy.cpp).  Large ones go through
// the context's ring of page-locked slots, copied by a few threads while the
// DMA engine moves the previous slot.  copy_to_device returns once the host
// buffer has been read (the uploads may still be in flight); copy_to_host
// returns with the stream synchronised.  Caller holds the context lock.
bool ring_wanted(size_t bytes);
int ring_ensure(zntt_ctx *c);
void ring_free(zntt_ctx *c);
// A ring of one's own (the serving loop's result reads): null on failure.
// The copies hold no lock; the caller serialises them.
HostRing *ring_create();
void ring_destroy(HostRing *r);
int ring_copy_to_device(HostRing *r, void *dev, const void *host, size_t bytes, hipStream_t stream);
int ring_copy_to_host(HostRing *r, void *host, const void *dev, size_t bytes, hipStream_t stream);
int copy_to_device(zntt_ctx *c, void *dev, const void *host, size_t bytes, hipStream_t stream);
int copy_to_host(zntt_ctx *c, void *host, const void *dev, size_t bytes, hipStream_t stream);
int power_table(zntt_ctx *c, int field, const Elem &root, int L, hipStream_t s, TwiddleEntry **out);
int coset_table(zntt_ctx *c, int field, const Elem &shift, size_t count, hipStream_t s, const uint32_t **out);
// One transform of 2^L user-layout values on the device (tw: table of the
// root's powers, coset: optional shift powers, scale: optional dev-form
// multiplier of every output).
int run_device(zntt_ctx *c, int field, int L, void *d_data, const uint32_t *tw, const uint32_t *coset,
               const std::array<uint32_t, 9> *scale, hipStream_t s, bool tw_one = true);

}  // namespace zntt_detail
