// Mersenne-31 (p = 2^31 - 1) on gfx950 VALU, for the eight-lane M31
// transform (Vec8NttM31, libs/field/src/lib.zig:39-137).
//
// Storage is the reference's SmallField element: the canonical value as one
// little-endian u64 (field.zig:142).  On the device a value is one u32 below
// p; constants (twiddles, the n^-1 scale) are plain values too, since the
// Mersenne product needs no Montgomery form:
//   a * b = hi * 2^31 + lo = hi + lo (mod p),  hi = ab >> 31, lo = ab & p,
// and hi + lo < 2p for a, b < p (it would reach 2p only for ab = p(p + 2)),
// so one min(s, s - p) -- s - p wraps to a huge u32 exactly when s < p --
// leaves the canonical value.  That is the reference's mul (field.zig:358-366):
// the same residue, reduced fully.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_traits.hpp"

namespace zntt {

struct M31Arith {
  static constexpr uint32_t P = 0x7fffffffu;
  // x < 2p -> x mod p
  __device__ static __forceinline__ uint32_t canon(uint32_t x) { return min(x, x - P); }
  __device__ static __forceinline__ uint32_t mul(uint32_t a, uint32_t b) {
    const uint64_t t = (uint64_t)a * b;
    return canon(((uint32_t)t & P) + (uint32_t)(t >> 31));
  }
  // (u, v) <- (u + w v, u - w v), all canonical (lib.zig:87-96).
  __device__ static __forceinline__ void butterfly(uint32_t &u, uint32_t &v, uint32_t w) {
    const uint32_t t = mul(v, w);
    const uint32_t s = canon(u + t);
    v = canon(u + P - t);
    u = s;
  }
  __device__ static __forceinline__ void butterfly_one(uint32_t &u, uint32_t &v) {
    const uint32_t s = canon(u + v);
    v = canon(u + P - v);
    u = s;
  }
};

// Traits for the generic pass kernels (the same shape as F31T: one word per
// element in registers, LDS and scratch; u64 in the caller's buffer).
struct M31T {
  using A = M31Arith;
  using E = uint32_t;
  static constexpr int W = 1;
  static constexpr int UW = 2;
  static constexpr uint32_t kP = A::P;
  __device__ static E load_user(const uint32_t *base, size_t i) { return gld<uint32_t>(base)[2 * i]; }
  __device__ static void store_user(uint32_t *base, size_t i, E x) { gst<zu32x2>(base)[i] = zu32x2{x, 0u}; }
  __device__ static E load_dev(const uint32_t *base, size_t i) { return gld<uint32_t>(base)[i]; }
  __device__ static void store_dev(uint32_t *base, size_t i, const E &x) { gst<uint32_t>(base)[i] = x; }
  __device__ static E lds_get(const uint32_t *lds, int, int i) { return lds[i]; }
  __device__ static void lds_put(uint32_t *lds, int, int i, const E &x) { lds[i] = x; }
  __device__ static E from_words(const uint32_t *w) { return w[0]; }
  __device__ static E zero() { return 0; }
  __device__ static E mul(const E &a, const E &b) { return A::mul(a, b); }
  __device__ static void mul2(E &x0, const E &b0, E &x1, const E &b1) {
    x0 = A::mul(x0, b0);
    x1 = A::mul(x1, b1);
  }
  template <bool NORM = true>
  __device__ static void bfly(E &u, E &v, const E &w) {
    A::butterfly(u, v, w);
  }
  template <bool NORM = true>
  __device__ static void bfly2(E &u0, E &v0, const E &w0, E &u1, E &v1, const E &w1) {
    A::butterfly(u0, v0, w0);
    A::butterfly(u1, v1, w1);
  }
  template <bool NORM = true, bool LATE = false>
  __device__ static void bfly_one(E &u, E &v) {
    A::butterfly_one(u, v);
  }
  static constexpr bool one_ok(int) { return true; }
  template <bool NORM = true>
  __device__ static void bfly_zero(E &u, E &v) {
    v = u;
  }
  __device__ static E canon(E x) { return A::canon(x); }
  __device__ static E neg(const E &x) { return x == 0 ? 0 : A::P - x; }
};

}  // namespace zntt
