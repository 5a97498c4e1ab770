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

// (hi * 2^64 + lo) mod p, canonical, for any lo, hi.  With hi = h1 * 2^32 + h0,
// x = lo + h0 * 2^64 + h1 * 2^96 = lo + h0 * eps - h1.  The term
// z = (h0 + 1) * eps - h1 = h0 * eps + (eps - h1) is one multiply-add whose
// 32-bit addend is ~h1, never negative and below 2^64 - 2^32, and
// x = lo + z - eps.  r = lo + z (mod 2^64):
// * carried: lo + z = r + 2^64 = r + eps, so x = r, and r < z < p;
// * otherwise x = r - eps, in [0, p) when r >= eps, else r - eps + p, which
//   is r - 2 eps mod 2^64.
// One carry test and one compare against eps pick the multiple of eps to
// subtract; there is no separate canonicalising step.
__device__ __forceinline__ uint64_t gl_reduce128(uint64_t lo, uint64_t hi) {
  const uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
  const uint64_t z = (uint64_t)h0 * kGlEps + (uint32_t)~h1;
  const uint64_t r = lo + z;
  const bool carried = r < z;
  const uint64_t k = carried ? 0 : (r < kGlEps ? 2 * kGlEps : kGlEps);
  return r - k;
}

// a * b mod p, canonical; a, b any u64 (the 128-bit product is exact).
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

// Values in the transform's inner loop are only "weakly reduced": any u64,
// congruent to the true value mod p; stores canonicalise.

// Radix-2 DIT butterfly on weakly reduced values: (u, v) <- (u + w v, u - w v),
// outputs weakly reduced.  With t = w v reduced below p, u + t carries to a
// value below t (< p), so adding eps cannot carry again; u - t borrows to a
// value above 2^64 - p = eps, so subtracting eps cannot borrow again.
__device__ __forceinline__ void gl_butterfly(uint64_t &u, uint64_t &v, uint64_t w) {
  const uint64_t t = gl_mul(v, w);
  uint64_t s = u + t;
  s += (s < t) ? kGlEps : 0;
  uint64_t d = u - t;
  d -= (u < t) ? kGlEps : 0;
  u = s;
  v = d;
}

// Products by the powers of two that are 8th roots of unity here (2 has
// order 192 mod p, so the primitive 8th roots are +-2^24 and +-2^72): a
// shift and one reduction instead of the four-multiply product.  Any u64 x
// in, canonical out.
__device__ __forceinline__ uint64_t gl_shl24(uint64_t x) { return gl_reduce128(x << 24, x >> 40); }
__device__ __forceinline__ uint64_t gl_shl48(uint64_t x) { return gl_reduce128(x << 48, x >> 16); }
// x 2^72 = y 2^64 with y = x 2^8 (72 bits): (y mod 2^64) 2^64 + (x >> 56) 2^128,
// and 2^128 = -2^32 (mod p).
__device__ __forceinline__ uint64_t gl_shl72(uint64_t x) {
  const uint64_t c = gl_reduce128(0, x << 8);
  const uint64_t d = (x >> 56) << 32;  // below 2^40
  uint64_t r = c - d;
  if (c < d) r -= kGlEps;  // c - d + p
  return r;
}
// 2^72 mod p = 2^8 (2^32 - 1).
constexpr uint64_t kGl2p72 = 0xFFFFFFFF00ull;

// Butterfly with the product t (< p) already formed: (u, v) <- (u + t, u - t),
// u any u64, outputs weakly reduced (the carry and borrow fixes of
// gl_butterfly).
__device__ __forceinline__ void gl_butterfly_t(uint64_t &u, uint64_t &v, uint64_t t) {
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
