// Lifetime scenarios of the C ABI, run on the CPU against fake_hip.cpp under
// AddressSanitizer.  Each scenario destroys or trims a context while batches
// or servers made on it are alive, in an order a C caller (or a garbage
// collector finalising wrappers) may use, and then keeps using the children.
// A scenario passes when it runs to the end with no ASan report and the
// context's byte accounting balances.  Usage: lifetime <scenario>.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "zntt_gpu.h"

extern "C" void fake_hip_hold(int on);
extern "C" long fake_hip_alloc_calls();
extern "C" long fake_hip_stream_waits();

#define CHECK(x)                                                              \
  do {                                                                        \
    int st_ = (x);                                                            \
    if (st_ != ZNTT_OK) {                                                     \
      fprintf(stderr, "%s:%d: %s -> %s\n", __FILE__, __LINE__, #x,         \
              zntt_status_string(st_));                                       \
      exit(2);                                                                \
    }                                                                         \
  } while (0)
#define EXPECT(c)                                                             \
  do {                                                                        \
    if (!(c)) {                                                               \
      fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #c);      \
      exit(3);                                                                \
    }                                                                         \
  } while (0)

namespace {

const int F = ZNTT_FIELD_BN254_FR;

struct Root {
  uint8_t b[32];
};
Root root_of(size_t log_n) {
  Root r;
  CHECK(zntt_primitive_root(F, log_n, r.b));
  return r;
}
Root small(uint64_t v) {
  // 5 in the storage layout: any element works, the fake kernels do no arithmetic
  Root r;
  memset(r.b, 0, sizeof r.b);
  memcpy(r.b, &v, sizeof v);
  return r;
}

// Three device transforms of one field: forward, inverse, coset-forward.
struct Bufs {
  std::vector<void *> p;
  std::vector<Root> roots;
  Root shift = small(5);
  std::vector<zntt_transform> ts;
  explicit Bufs(std::vector<size_t> logs) {
    int dir = 0;
    for (size_t L : logs) {
      void *d = nullptr;
      if (hipMalloc(&d, ((size_t)1 << L) * 32) != hipSuccess) exit(4);
      memset(d, 0, ((size_t)1 << L) * 32);
      p.push_back(d);
      roots.push_back(root_of(L));
    }
    for (size_t i = 0; i < logs.size(); i++) {
      zntt_transform t = {};
      t.field = F;
      t.direction = dir;
      t.log_n = logs[i];
      t.data = p[i];
      t.root = roots[i].b;
      t.shift = dir == ZNTT_COSET_FORWARD ? shift.b : nullptr;
      ts.push_back(t);
      dir = (dir + 1) % 3;
    }
  }
  ~Bufs() {
    for (void *d : p) (void)hipFree(d);
  }
};

zntt_batch *make_batch(zntt_ctx *c, Bufs &b) {
  zntt_batch *bt = nullptr;
  CHECK(zntt_batch_create(c, b.ts.data(), b.ts.size(), &bt));
  CHECK(zntt_batch_execute(bt, zntt_ctx_stream(c)));  // captures the graph and replays it
  return bt;
}

zntt_server *make_server(zntt_ctx *c) {
  zntt_server_config cfg = {};
  cfg.field = F;
  cfg.hbm_budget = 64ull << 20;
  cfg.max_log_n = 16;
  zntt_server *s = nullptr;
  CHECK(zntt_server_create(c, &cfg, &s));
  return s;
}

std::vector<uint64_t> submit_some(zntt_server *s, int n) {
  Root sh = small(5);
  std::vector<uint64_t> t(n);
  for (int i = 0; i < n; i++) {
    zntt_lde_request r = {};
    r.log_n = 8 + (i % 4);
    r.len = (size_t)1 << r.log_n;
    r.log_blowup = 2 + (i % 2);
    r.shift = sh.b;
    r.seed = 4;
    r.stream_id = (uint64_t)i;
    r.keep_result = 1;
    CHECK(zntt_server_submit(s, &r, &t[i]));
  }
  return t;
}

// The context is destroyed while a batch made on it is alive; the batch is
// replayed and destroyed afterwards.
void batch_outlives_ctx() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  Bufs b({12, 14, 13, 9});
  zntt_batch *bt = make_batch(c, b);
  zntt_ctx_destroy(c);
  CHECK(zntt_batch_execute(bt, nullptr));  // graph replay reads the batch's pinned tables
  CHECK(zntt_batch_status(bt, nullptr));
  zntt_batch_destroy(bt);
}

// trim while a batch holds pinned tables: the tables stay, the batch still
// replays; once it is gone a trim returns every byte.
void trim_with_live_batch() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  Bufs b({12, 14, 13});
  zntt_batch *bt = make_batch(c, b);
  CHECK(zntt_ctx_trim(c));
  EXPECT(zntt_ctx_bytes_in_use(c) > 0);  // the batch's scratch, work list and pinned tables
  CHECK(zntt_batch_execute(bt, zntt_ctx_stream(c)));
  double ms = 0;
  CHECK(zntt_batch_autotune(bt, 1, &ms));
  CHECK(zntt_batch_execute(bt, zntt_ctx_stream(c)));
  zntt_batch_destroy(bt);
  CHECK(zntt_ctx_trim(c));
  EXPECT(zntt_ctx_bytes_in_use(c) == 0);
  zntt_ctx_destroy(c);
}

// The context is destroyed while a server made on it holds kept results.
void server_outlives_ctx() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  zntt_server *s = make_server(c);
  auto t = submit_some(s, 6);
  CHECK(zntt_server_drain(s));
  zntt_ctx_destroy(c);
  auto more = submit_some(s, 4);  // the server keeps working on its context
  CHECK(zntt_server_drain(s));
  std::vector<uint8_t> host(((size_t)1 << 13) * 32);
  zntt_request_info info;
  CHECK(zntt_server_query(s, t[0], &info));
  EXPECT(info.state == ZNTT_REQ_DONE);
  CHECK(zntt_server_read(s, t[0], host.data(), (size_t)1 << (info.log_n + info.log_blowup)));
  for (uint64_t k : t) CHECK(zntt_server_release(s, k));
  zntt_server_destroy(s);
}

// trim while requests are admitted and pinning their tables: completion of
// their first steps is held back, so they are in flight when the trim runs,
// and their later passes and retirement come after it.
void trim_with_live_server() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  zntt_server *s = make_server(c);
  for (int round = 0; round < 3; round++) {
    fake_hip_hold(1);
    submit_some(s, 8);
    // With completions held, the scheduler stops admitting once kDepth steps
    // are in flight; it is enough that some requests are admitted.
    zntt_server_stats st = {};
    for (int i = 0; i < 5000 && st.active == 0; i++) {
      CHECK(zntt_server_stats_get(s, &st));
      if (st.active == 0) usleep(1000);
    }
    EXPECT(st.active > 0);
    usleep(20000);
    CHECK(zntt_ctx_trim(c));
    fake_hip_hold(0);
    CHECK(zntt_server_drain(s));
  }
  zntt_server_destroy(s);
  CHECK(zntt_ctx_trim(c));
  EXPECT(zntt_ctx_bytes_in_use(c) == 0);
  zntt_ctx_destroy(c);
}

// A batch of a destroyed context is destroyed after a new context has been
// made and filled its cache: the new context's accounting and tables must
// not change.
void late_destroy_after_new_ctx() {
  zntt_ctx *a = nullptr, *b = nullptr;
  CHECK(zntt_ctx_create(&a, 0, 1ull << 30));
  Bufs bufs({12, 14, 13});
  zntt_batch *bt = make_batch(a, bufs);
  zntt_ctx_destroy(a);
  CHECK(zntt_ctx_create(&b, 0, 1ull << 30));
  std::vector<uint8_t> data(((size_t)1 << 14) * 32, 0);
  Root r = root_of(14);
  CHECK(zntt_ntt(b, F, data.data(), (size_t)1 << 14, 14, r.b));
  const uint64_t before = zntt_ctx_bytes_in_use(b);
  zntt_batch_destroy(bt);
  EXPECT(zntt_ctx_bytes_in_use(b) == before);
  CHECK(zntt_ntt(b, F, data.data(), (size_t)1 << 14, 14, r.b));
  CHECK(zntt_ctx_trim(b));
  EXPECT(zntt_ctx_bytes_in_use(b) == 0);
  zntt_ctx_destroy(b);
}

// zntt_execute_device_noalloc makes no allocation or free of any kind once
// the tables are cached; it refuses sizes past its static scratch and
// transforms whose tables are not cached, before enqueuing anything.
void noalloc_allocates_nothing() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  Bufs b({16, 13, 15, 20});  // forward 2^16, inverse 2^13, coset 2^15, forward 2^20
  for (auto &t : b.ts) CHECK(zntt_execute_device(c, &t, nullptr));  // builds and caches the tables
  const long before = fake_hip_alloc_calls();
  const uint64_t bytes = zntt_ctx_bytes_in_use(c);
  for (int rep = 0; rep < 2; rep++)
    for (auto &t : b.ts) CHECK(zntt_execute_device_noalloc(c, &t, nullptr));
  EXPECT(fake_hip_alloc_calls() == before);
  EXPECT(zntt_ctx_bytes_in_use(c) == bytes);
  // Fr 2^21 needs 72 MiB of scratch, more than the static 64 MiB.
  void *big = nullptr;
  if (hipMalloc(&big, ((size_t)1 << 21) * 32) != hipSuccess) exit(4);
  Root r21 = root_of(21);
  zntt_transform t = {};
  t.field = F;
  t.direction = ZNTT_FORWARD;
  t.log_n = 21;
  t.data = big;
  t.root = r21.b;
  EXPECT(zntt_execute_device_noalloc(c, &t, nullptr) == ZNTT_OUT_OF_MEMORY);
  // A root whose table is not cached.
  Root other = small(12345);
  t.log_n = 12;
  t.root = other.b;
  const long before2 = fake_hip_alloc_calls();
  EXPECT(zntt_execute_device_noalloc(c, &t, nullptr) == ZNTT_NOT_PREPARED);
  t.root = b.roots[1].b;  // cached for the inverse only: the forward table of that root is another one
  t.log_n = 13;
  t.direction = ZNTT_COSET_FORWARD;
  t.shift = nullptr;
  EXPECT(zntt_execute_device_noalloc(c, &t, nullptr) == ZNTT_INVALID_ARGUMENT);
  Root sh7 = small(7);
  t.shift = sh7.b;
  EXPECT(zntt_execute_device_noalloc(c, &t, nullptr) == ZNTT_NOT_PREPARED);
  EXPECT(fake_hip_alloc_calls() == before2);
  (void)hipFree(big);
  zntt_ctx_destroy(c);
}

// Multi-pass device transforms on two streams of one context: the second
// stream waits once for the first one's work on the shared scratch; calls on
// the same stream add no wait.
void streams_share_scratch_in_order() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  hipStream_t a = nullptr, bs = nullptr;
  (void)hipStreamCreateWithFlags(&a, 0);
  (void)hipStreamCreateWithFlags(&bs, 0);
  Bufs b({16, 16});
  CHECK(zntt_execute_device(c, &b.ts[0], a));
  const long w0 = fake_hip_stream_waits();
  CHECK(zntt_execute_device(c, &b.ts[0], a));
  EXPECT(fake_hip_stream_waits() == w0);
  CHECK(zntt_execute_device(c, &b.ts[1], bs));
  EXPECT(fake_hip_stream_waits() == w0 + 1);
  CHECK(zntt_execute_device(c, &b.ts[1], bs));
  EXPECT(fake_hip_stream_waits() == w0 + 1);
  std::vector<uint8_t> host(((size_t)1 << 14) * 32, 0);
  Root r = root_of(14);
  CHECK(zntt_ntt(c, F, host.data(), (size_t)1 << 14, 14, r.b));  // the context's own stream waits too
  EXPECT(fake_hip_stream_waits() == w0 + 2);
  (void)hipStreamDestroy(a);
  (void)hipStreamDestroy(bs);
  zntt_ctx_destroy(c);
}

// After zntt_ctx_reserve(ctx, 12) the seven zig-ntt functions at sizes up
// to 2^12 make no allocation or free and hold no new device bytes, for
// every field, every root (tables they do not find are built into the
// reservation's slots), registered and foreign twiddle tables, and across a
// trim.  Above the reservation they allocate as before.
void reserved_calls_allocate_nothing() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 1ull << 30));
  CHECK(zntt_ctx_reserve(c, 12));
  const int fields[4] = {ZNTT_FIELD_BN254_FR, ZNTT_FIELD_GOLDILOCKS, ZNTT_FIELD_BABYBEAR, ZNTT_FIELD_KOALABEAR};
  const size_t logs[4] = {1, 5, 11, 12};
  // Registered tables (precomputeTwiddles allocates them: its error set has
  // OutOfMemory), one per field at 2^12.
  std::vector<std::vector<std::vector<uint8_t>>> reg(4);
  std::vector<std::vector<void *>> reg_ptrs(4);
  for (int fi = 0; fi < 4; fi++) {
    const size_t eb = zntt_element_size(fields[fi]);
    uint8_t root[32] = {0};
    CHECK(zntt_primitive_root(fields[fi], 12, root));
    for (size_t s = 0; s < 12; s++) {
      reg[fi].emplace_back(((size_t)1 << s) * eb, 0);
      reg_ptrs[fi].push_back(reg[fi].back().data());
    }
    CHECK(zntt_precompute_twiddles(c, fields[fi], 12, root, reg_ptrs[fi].data(), 12));
  }
  long before = fake_hip_alloc_calls();
  uint64_t bytes = zntt_ctx_bytes_in_use(c);
  for (int rep = 0; rep < 2; rep++) {
    for (int fi = 0; fi < 4; fi++) {
      const int f = fields[fi];
      const size_t eb = zntt_element_size(f);
      for (size_t L : logs) {
        const size_t n = (size_t)1 << L;
        std::vector<uint8_t> data(n * eb, 1);
        uint8_t root[32] = {0}, other[32] = {0};
        CHECK(zntt_primitive_root(f, L, root));
        other[0] = (uint8_t)(3 + L + rep);  // a root no cache holds
        CHECK(zntt_ntt(c, f, data.data(), n, L, root));
        CHECK(zntt_intt(c, f, data.data(), n, L, root));
        CHECK(zntt_ntt(c, f, data.data(), n, L, other));
        CHECK(zntt_intt(c, f, data.data(), n, L, other));
        CHECK(zntt_bit_reverse(c, f, data.data(), n));
        // A foreign table (never registered), both directions.
        std::vector<std::vector<uint8_t>> st;
        std::vector<const void *> sp;
        std::vector<size_t> lens;
        for (size_t s = 0; s < L; s++) {
          st.emplace_back(((size_t)1 << s) * eb, 2);
          sp.push_back(st.back().data());
          lens.push_back((size_t)1 << s);
        }
        CHECK(zntt_ntt_with_twiddles(c, f, data.data(), n, L, sp.data(), lens.data(), L));
        CHECK(zntt_intt_with_twiddles(c, f, data.data(), n, L, sp.data(), lens.data(), L));
      }
      // The registered table, both directions (the inverse is derived into a slot).
      const size_t n = (size_t)1 << 12;
      std::vector<uint8_t> data(n * eb, 1);
      std::vector<size_t> lens;
      for (size_t s = 0; s < 12; s++) lens.push_back((size_t)1 << s);
      const void *const *sp = (const void *const *)reg_ptrs[fi].data();
      CHECK(zntt_ntt_with_twiddles(c, f, data.data(), n, 12, sp, lens.data(), 12));
      CHECK(zntt_intt_with_twiddles(c, f, data.data(), n, 12, sp, lens.data(), 12));
    }
    EXPECT(fake_hip_alloc_calls() == before);
    EXPECT(zntt_ctx_bytes_in_use(c) == bytes);
    if (rep == 0) {
      // A trim frees the cached and registered tables (counted frees), and
      // the reservation outlives it: the second round allocates nothing
      // either (registered tables now go the foreign way).
      CHECK(zntt_ctx_trim(c));
      before = fake_hip_alloc_calls();
      bytes = zntt_ctx_bytes_in_use(c);
    }
  }
  // Above the reservation the calls allocate (and can run out of memory).
  std::vector<uint8_t> big(((size_t)1 << 13) * 32, 1);
  uint8_t r13[32] = {0};
  CHECK(zntt_primitive_root(F, 13, r13));
  CHECK(zntt_ntt(c, F, big.data(), (size_t)1 << 13, 13, r13));
  EXPECT(fake_hip_alloc_calls() > before);
  // A larger reservation grows it; a smaller one changes nothing.
  const uint64_t b12 = zntt_ctx_bytes_in_use(c);
  CHECK(zntt_ctx_reserve(c, 10));
  EXPECT(zntt_ctx_bytes_in_use(c) == b12);
  EXPECT(zntt_ctx_reserve(c, ZNTT_MAX_LOG_N + 1) == ZNTT_UNSUPPORTED);
  for (int fi = 0; fi < 4; fi++)
    CHECK(zntt_free_twiddles(c, (const void *const *)reg_ptrs[fi].data(), 12));
  zntt_ctx_destroy(c);
}

// Host buffers larger than one 16 MiB slot of the context's page-locked
// ring go through it in chunks (2 and 8 chunks here, the second reusing
// every slot twice); the stand-in kernels change nothing, so each buffer must
// come back as it went, and an LDE's output must be its coefficients
// followed by the zero padding.  Under AddressSanitizer this also checks the
// chunk arithmetic and the copy threads' bounds.
void host_copies_through_the_ring() {
  zntt_ctx *c = nullptr;
  CHECK(zntt_ctx_create(&c, 0, 4ull << 30));
  for (size_t L : {20, 22}) {
    const size_t bytes = ((size_t)1 << L) * 32;
    std::vector<uint8_t> data(bytes), keep;
    for (size_t i = 0; i < bytes; i++) data[i] = (uint8_t)(i * 131 + (i >> 13) * 7 + L);
    keep = data;
    Root r = root_of(L);
    CHECK(zntt_ntt(c, F, data.data(), (size_t)1 << L, L, r.b));
    EXPECT(data == keep);
  }
  const size_t log_k = 21, log_b = 1, k = (size_t)1 << log_k;
  std::vector<uint8_t> coeffs(k * 32), out((k << log_b) * 32, 0xAB);
  for (size_t i = 0; i < coeffs.size(); i++) coeffs[i] = (uint8_t)(i * 29 + 3);
  Root sh = small(5);
  CHECK(zntt_lde(c, F, coeffs.data(), log_k, log_b, sh.b, out.data()));
  EXPECT(memcmp(out.data(), coeffs.data(), coeffs.size()) == 0);
  for (size_t i = coeffs.size(); i < out.size(); i++) EXPECT(out[i] == 0);
  // Lengths that end in a partial slot, both ways (a polynomial product:
  // 600,001 and 700,003 coefficients in, 1,300,003 out).
  const size_t na = 600001, nb = 700003;
  std::vector<uint8_t> a(na * 32, 1), b(nb * 32, 2), prod((na + nb - 1) * 32);
  a[na * 32 - 1] = b[nb * 32 - 1] = 3;  // nonzero leading coefficients
  size_t got = 0;
  CHECK(zntt_poly_mul(c, F, a.data(), na, b.data(), nb, prod.data(), na + nb - 1, &got));
  EXPECT(got == na + nb - 1);
  zntt_ctx_destroy(c);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return 1;
  const std::string s = argv[1];
  if (s == "batch_outlives_ctx") batch_outlives_ctx();
  else if (s == "trim_with_live_batch") trim_with_live_batch();
  else if (s == "server_outlives_ctx") server_outlives_ctx();
  else if (s == "trim_with_live_server") trim_with_live_server();
  else if (s == "late_destroy_after_new_ctx") late_destroy_after_new_ctx();
  else if (s == "noalloc_allocates_nothing") noalloc_allocates_nothing();
  else if (s == "streams_share_scratch_in_order") streams_share_scratch_in_order();
  else if (s == "reserved_calls_allocate_nothing") reserved_calls_allocate_nothing();
  else if (s == "host_copies_through_the_ring") host_copies_through_the_ring();
  else return 1;
  printf("ok %s\n", s.c_str());
  return 0;
}
