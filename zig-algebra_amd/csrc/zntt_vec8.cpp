// C ABI of the eight-lane Mersenne-31 transform (Vec8NttM31,
// libs/field/src/lib.zig:39-137): argument checks in the reference's order,
// the twiddle table, the lane (de)interleaving and the eight transforms
// (pass_m31.hip).
#include <hip/hip_runtime.h>

#include "ctx_internal.hpp"

using namespace zntt_detail;

namespace {

constexpr uint64_t kP = 0x7fffffffull;  // M31, libs/field/src/predef/m31.zig:6

uint64_t m31_mul(uint64_t a, uint64_t b) { return (a % kP) * (b % kP) % kP; }
uint64_t m31_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1, b = a % kP;
  for (; e; e >>= 1) {
    if (e & 1) r = m31_mul(r, b);
    b = m31_mul(b, b);
  }
  return r;
}
// Total inverse, as the reference's (zero maps to zero).
uint64_t m31_inv(uint64_t a) { return a % kP == 0 ? 0 : m31_pow(a, kP - 2); }

// transformSize (lib.zig:131-136): log_n >= bit width - 3 is InvalidLength.
int check_size(size_t log_n) { return log_n >= 61 ? ZNTT_INVALID_LENGTH : ZNTT_OK; }

// The eight transforms of the 8 * 2^L u64 values at `data` (device), in
// place on stream s.  Caller holds the context lock.
int vec8_device(zntt_ctx *c, bool inverse, void *data, int L, uint64_t root, hipStream_t s) {
  if (L == 0) return ZNTT_OK;  // one point per lane: bit reversal and n^-1 = 1 leave it
  const uint64_t n = 1ull << L;
  // The inverse runs the forward DAG with root^-1, then scales by n^-1
  // (lib.zig:114-129).
  const uint64_t r = inverse ? m31_inv(root) : root;
  // From 2^6 values a lane the passes work on the interleaved buffer itself
  // (scratch: 8 * n u32); smaller transforms gather the lanes first.
  // Context memory: eight lane-major transforms (u64, gathered form only),
  // the pass scratch (u32) and the stage powers; the twiddle table (u32) has
  // a buffer of its own and is rebuilt only when (root, L) changes.
  const bool fused = L >= zntt::kM31LanesMinLog;
  const size_t lanes_b = fused ? 0 : 8 * n * 8, scratch_b = (fused ? 8 : 1) * n * 4, table_b = n * 4,
               pows_b = 64 * 4;
  int st = ensure(c, &c->aux, &c->aux_bytes, lanes_b + scratch_b + pows_b);
  if (st) return st;
  uint8_t *base = (uint8_t *)c->aux;
  uint64_t *lanes = (uint64_t *)base;
  uint32_t *scratch = (uint32_t *)(base + lanes_b);
  uint32_t *pows_d = (uint32_t *)(base + lanes_b + scratch_b);
  const uint64_t rk = r % kP;
  const bool cached = c->m31_table && c->m31_log_n == L && c->m31_root == rk;
  hipError_t e = hipSuccess;
  if (!cached) {
    c->m31_log_n = -1;  // no valid table until this call's build has completed
    if ((st = ensure(c, &c->m31_table, &c->m31_table_bytes, table_b))) return st;
    if ((st = ensure_pinned(c, pows_b))) return st;
    uint32_t *pows = (uint32_t *)c->pinned;
    uint64_t x = rk;
    for (int k = 0; k < L; k++) {
      pows[k] = (uint32_t)x;
      x = m31_mul(x, x);
    }
    e = hipMemcpyAsync(pows_d, pows, (size_t)L * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = zntt::launch_m31_table(L, (uint32_t *)c->m31_table, pows_d, s);
  }
  const uint32_t *table = (const uint32_t *)c->m31_table;
  uint32_t scale[9] = {(uint32_t)m31_inv(n % kP)};
  if (fused) {
    if (e == hipSuccess)
      e = zntt::launch_transform_m31_lanes(L, (uint32_t *)data, scratch, table, inverse ? scale : nullptr, s);
  } else {
    if (e == hipSuccess) e = zntt::launch_m31_lanes(true, data, lanes, n, s);
    for (int lane = 0; lane < 8 && e == hipSuccess; lane++)
      e = zntt::launch_transform_m31(L, (uint32_t *)(lanes + (size_t)lane * n), scratch, table,
                                     inverse ? scale : nullptr, s);
    if (e == hipSuccess) e = zntt::launch_m31_lanes(false, lanes, data, n, s);
  }
  // The staged powers are read by the copy above; keep them until it is done.
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess && !cached) {
    c->m31_root = rk;
    c->m31_log_n = L;
  }
  return hip_status(e);
}

int vec8_host(zntt_ctx *c, bool inverse, void *data, size_t len, size_t log_n, const void *root) {
  if (!c || !data || !root) return ZNTT_INVALID_ARGUMENT;
  int st = check_size(log_n);
  if (st) return st;
  const size_t n = (size_t)1 << log_n;
  if (len != 8 * n) return ZNTT_INVALID_LENGTH;  // lib.zig:54-55
  if (log_n > ZNTT_VEC8_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  uint64_t rv;
  memcpy(&rv, root, 8);
  CallGuard g(c, c->stream);
  if ((st = ensure(c, &c->staging, &c->staging_bytes, len * 8))) return st;
  st = copy_to_device(c, c->staging, data, len * 8, c->stream);
  if (!st) st = vec8_device(c, inverse, c->staging, (int)log_n, rv, c->stream);
  if (st) {
    hipStreamSynchronize(c->stream);
    return st;
  }
  return copy_to_host(c, data, c->staging, len * 8, c->stream);
}

}  // namespace

extern "C" {

int zntt_ntt_vec8_m31(zntt_ctx *c, void *data, size_t len, size_t log_n, const void *root) {
  return vec8_host(c, false, data, len, log_n, root);
}

int zntt_intt_vec8_m31(zntt_ctx *c, void *data, size_t len, size_t log_n, const void *root) {
  return vec8_host(c, true, data, len, log_n, root);
}

int zntt_vec8_m31_device(zntt_ctx *c, int inverse, void *data, size_t log_n, const void *root, void *stream) {
  if (!c || !data || !root) return ZNTT_INVALID_ARGUMENT;
  int st = check_size(log_n);
  if (st) return st;
  if (log_n > ZNTT_VEC8_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  uint64_t rv;
  memcpy(&rv, root, 8);
  CallGuard g(c, (hipStream_t)stream);
  return vec8_device(c, inverse != 0, data, (int)log_n, rv, (hipStream_t)stream);
}

}  // extern "C"
