// 31-bit prime fields (BabyBear p = 2^31 - 2^27 + 1, KoalaBear
// p = 2^31 - 2^24 + 1) on gfx950 VALU.
//
// Storage format is the reference's SmallField element: the canonical value
// as one little-endian u64 (libs/field/src/field.zig:142); the reference
// multiplies with a u128 `%` (field.zig:358-369).  On the device a value is
// one u32 below p.  Constants (twiddles, coset powers, scale factors) are
// kept in Montgomery form c * 2^32 mod p, so the Montgomery product of a
// plain value and a constant is the plain product:
//     mont(x, c * 2^32) = x * c * 2^32 / 2^32 = x * c   (mod p).
// mont(a, b) = (a*b + m*p) / 2^32 with m = (a*b mod 2^32) * (-p^-1) mod 2^32;
// for a < 2p and b < p it is below 2p (a*b < 2p^2 and m*p < 2^32 p, and
// 2p < 2^32), and one min(t, t - p) makes it canonical: t - p wraps to a
// huge u32 exactly when t < p.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zntt {

template <uint32_t P, uint32_t NPRIME>
struct F31Arith {
  // x < 2p -> x mod p
  __device__ static __forceinline__ uint32_t canon(uint32_t x) { return min(x, x - P); }
  __device__ static __forceinline__ uint32_t mont(uint32_t a, uint32_t b) {
    const uint64_t x = (uint64_t)a * b;
    const uint32_t m = (uint32_t)x * NPRIME;
    const uint64_t t = x + (uint64_t)m * P;
    return canon((uint32_t)(t >> 32));
  }
  __device__ static __forceinline__ uint32_t add(uint32_t a, uint32_t b) { return canon(a + b); }
  __device__ static __forceinline__ uint32_t sub(uint32_t a, uint32_t b) { return canon(a + P - b); }
  // (u, v) <- (u + w v, u - w v); u, v canonical, w a Montgomery constant.
  __device__ static __forceinline__ void butterfly(uint32_t &u, uint32_t &v, uint32_t w) {
    const uint32_t t = mont(v, w);
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

constexpr uint32_t kBbP = 0x78000001u, kBbNPrime = 0x77ffffffu;
constexpr uint32_t kKbP = 0x7f000001u, kKbNPrime = 0x7effffffu;

}  // namespace zntt
