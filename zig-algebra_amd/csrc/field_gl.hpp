// Goldilocks p = 2^64 - 2^32 + 1 on gfx950 VALU.
//
// Storage format is the reference's SmallField element: the canonical value as
// one little-endian u64 (libs/field/src/field.zig:142).  The reference
// multiplies with a 128/64-bit `%` (field.zig:367-368); here the 128-bit
// product is formed from four v_mad_u64_u32 and reduced with the identities
// 2^64 = 2^32 - 1 and 2^96 = -1 (mod p), which yields the same canonical value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zntt {

constexpr uint64_t kGlP = 0xFFFFFFFF00000001ull;
constexpr uint64_t kGlEps = 0xFFFFFFFFull;  // 2^64 mod p

__device__ __forceinline__ uint64_t gl_canon(uint64_t x) { return x >= kGlP ? x - kGlP : x; }

// (hi * 2^64 + lo) mod p, canonical.
__device__ __forceinline__ uint64_t gl_reduce128(uint64_t lo, uint64_t hi) {
  const uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
  // lo - h1 * 2^96-part: 2^96 = -1
  uint64_t t0 = lo - h1;
  if (lo < (uint64_t)h1) t0 -= kGlEps;  // borrowed 2^64 = 2^32 - 1
  // h0 * 2^64 = h0 * (2^32 - 1)
  const uint64_t t1 = ((uint64_t)h0 << 32) - h0;
  uint64_t r = t0 + t1;
  if (r < t1) r += kGlEps;  // carried 2^64; cannot carry again (r < t1 <= 2^64 - 2^33 + 1)
  return gl_canon(r);
}

__device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
  const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;
  const uint64_t hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
  const uint64_t lo = (u << 32) | (uint32_t)p00;
  return gl_reduce128(lo, hi);
}

// Lazy forms for the transform's inner loop.  Values there are only "weakly
// reduced": any u64, congruent to the true value mod p; stores canonicalise.
//
// (hi * 2^64 + lo) mod p up to one multiple of p, for any lo, hi: with
// hi = h1 * 2^32 + h0, x = lo + h0 * 2^64 + h1 * 2^96 = lo + h0 * eps - h1.
// lo - h1 borrows only when lo < h1, and then the wrapped value exceeds eps,
// so subtracting eps (= adding back 2^64 - p) cannot borrow again; t0 + t1
// with t1 = h0 * eps <= (2^32 - 1)^2 carries to a value below t1, so adding
// eps cannot carry again.
__device__ __forceinline__ uint64_t gl_reduce_lazy(uint64_t lo, uint64_t hi) {
  const uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
  uint64_t t0 = lo - h1;
  t0 -= (lo < (uint64_t)h1) ? kGlEps : 0;
  const uint64_t t1 = (uint64_t)h0 * kGlEps;
  uint64_t r = t0 + t1;
  r += (r < t1) ? kGlEps : 0;
  return r;
}

// a * b mod p, weakly reduced; a, b any u64.
__device__ __forceinline__ uint64_t gl_mul_lazy(uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
  const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;
  const uint64_t hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
  const uint64_t lo = (u << 32) | (uint32_t)p00;
  return gl_reduce_lazy(lo, hi);
}

// a, b canonical.
__device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  if (s < a) s += kGlEps;  // a + b < 2p: no second carry
  return gl_canon(s);
}

__device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  if (a < b) d -= kGlEps;  // + p == - (2^32 - 1) after the wrap
  return d;
}

__device__ __forceinline__ uint64_t gl_neg(uint64_t a) { return a == 0 ? 0 : kGlP - a; }

// Radix-2 DIT butterfly on weakly reduced values: (u, v) <- (u + w v, u - w v),
// outputs weakly reduced.  With t = w v reduced below p, u + t carries to a
// value below t (< p), so adding eps cannot carry again; u - t borrows to a
// value above 2^64 - p = eps, so subtracting eps cannot borrow again.
__device__ __forceinline__ void gl_butterfly(uint64_t &u, uint64_t &v, uint64_t w) {
  const uint64_t t = gl_canon(gl_mul_lazy(v, w));
  uint64_t s = u + t;
  s += (s < t) ? kGlEps : 0;
  uint64_t d = u - t;
  d -= (u < t) ? kGlEps : 0;
  u = s;
  v = d;
}

// Butterfly with twiddle 1 on weakly reduced values.
__device__ __forceinline__ void gl_butterfly_one(uint64_t &u, uint64_t &v) {
  const uint64_t t = gl_canon(v);
  uint64_t s = u + t;
  s += (s < t) ? kGlEps : 0;
  uint64_t d = u - t;
  d -= (u < t) ? kGlEps : 0;
  u = s;
  v = d;
}

}  // namespace zntt
