// BN254 scalar field r on gfx950 VALU, radix 2^29.
//
// Storage (user) format: the reference's in-memory `Fr` -- four little-endian
// u64 limbs of the Montgomery form x * 2^256 mod r (libs/field/src/field.zig:884,
// libs/field/src/montgomery.zig:23-79).  That is byte-for-byte eight
// little-endian u32 words, which is how the kernels read and write it.
//
// Register format: nine 29-bit limbs.  On gfx950 v_mad_u64_u32 issues at the
// same rate as a 32-bit add, so the cost of a modular multiply is its total
// VALU instruction count.  With 29-bit limbs a column of the product-scanning
// Montgomery multiply sums at most 9 products a_i*b_j plus 9 products m_i*r_j
// and a carry, which stays below 2^64: every partial product is one
// accumulate-only v_mad_u64_u32 with no carry-out handling.  (With 32-bit
// limbs each product needs a second instruction to catch its carry.)
//
// Montgomery radix: R' = 2^261 (9 * 29).  Data stays in the reference's
// 2^256 form throughout; only constants -- twiddles, coset powers, scale
// factors -- are kept in the 2^261 form ("dev form", c * 2^261 mod r).  Then
//     mont(x * 2^256, c * 2^261) = x * c * 2^256 * 2^261 / 2^261 = x * c * 2^256,
// so a product of a datum and a constant is again in the reference's form.
//
// Lazy reduction.  Values are not reduced modulo r between butterflies:
//   * mont(a, b) < a*b/2^261 + r < 2r whenever a < 2^261 and b < r;
//   * the butterfly writes u + t and u + 4r - t, so each stage adds at most
//     4r to the bound.  Inputs are below 2r (a coset pre-scale is a
//     product); the first two stages' twiddle-1 butterflies give at most 12r
//     after stage 2 (u + 8r - v there), so a transform of 2^L points keeps
//     every value below 4r(L + 1) < 2^261 for L <= 28;
//   * limbs of the sum/difference may exceed 29 bits (up to 1.5 * 2^30 after
//     one stage); mont() accepts a-limbs below 3 * 2^30 (the column bound), so
//     only the `u` input of each butterfly is carry-normalised first.
// The last pass reduces each value fully (quotient estimate from the top limb,
// then one conditional subtraction) and repacks it into the 8-word layout, so
// results are canonical and bit-identical to the reference's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zntt {

constexpr uint32_t kM29 = 0x1FFFFFFFu;

// r in radix 2^29.
#define ZNTT_FR_P29 {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, \
                     0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}
// 4r in a redundant radix-2^29 form whose limbs 0..7 lie in [2^29, 2^30), so
// that u + 4r - t never borrows limb-wise when t is carry-normalised.
#define ZNTT_FR_C4R {0x20000004u, 0x3c3eb27du, 0x39709142u, 0x3f4243ccu, 0x36174a0bu, \
                     0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u}
// -r^-1 mod 2^29
constexpr uint32_t kFrN0 = 0x0FFFFFFFu;
// Quotient-estimate multiplier for fr_canon: just below 2^40 / (r_8 + 1) =
// 346695.2, with r_8 = r >> 232 = 0x30644e.
constexpr uint32_t kFrQMagic = 346694u;

struct Fr9 {
  uint32_t v[9];
};

// Hides a constant from the optimiser so the multiply by r_0 = 2^28 + 1 stays a
// single v_mad_u64_u32 instead of being strength-reduced into two 64-bit ops.
__device__ __forceinline__ uint32_t fr_opaque(uint32_t x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Pins the accumulator to one register pair between multiply-adds.  Without
// it the compiler splits each column into parallel partial sums and merges
// them with extra 64-bit adds (one more VALU op per column); a single chain
// is cheaper because the butterfly already carries enough independent work.
__device__ __forceinline__ void fr_chain(uint64_t &acc) { asm volatile("" : "+v"(acc)); }

// Montgomery product a * b * 2^-261 mod r (result < 2r, carry-normalised).
// a: limbs < 3 * 2^30, value < 2^261.   b: carry-normalised, value < r.
// Column bound: 9 * 3*2^30 * 2^29 + 9 * 2^58 + 2^36 < 2^64.
__device__ __forceinline__ Fr9 fr_mul(const Fr9 &a, const Fr9 &b) {
  constexpr uint32_t P[9] = ZNTT_FR_P29;
  const uint32_t p0 = fr_opaque(P[0]);
  uint32_t m[9];
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i <= k; i++) {
      acc += (uint64_t)a.v[i] * b.v[k - i];
      fr_chain(acc);
    }
#pragma unroll
    for (int i = 0; i < k; i++) {
      acc += (uint64_t)m[i] * P[k - i];
      fr_chain(acc);
    }
    m[k] = ((uint32_t)acc * kFrN0) & kM29;
    acc += (uint64_t)m[k] * p0;
    acc >>= 29;
  }
  Fr9 o;
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i < 9; i++) {
      acc += (uint64_t)a.v[i] * b.v[k - i];
      fr_chain(acc);
      acc += (uint64_t)m[i] * P[k - i];
      fr_chain(acc);
    }
    o.v[k - 9] = (uint32_t)acc & kM29;
    acc >>= 29;
  }
  o.v[8] = (uint32_t)acc;
  return o;
}

// Two independent Montgomery products, a0*b0 and a1*b1, with their
// multiply-add chains interleaved instruction by instruction.  A
// v_mad_u64_u32 that reads the accumulator written by the instruction just
// before it costs a wait state on gfx950; alternating two chains gives every
// multiply-add an independent predecessor, so none is needed.  Bounds as for
// fr_mul.
__device__ __forceinline__ void fr_mul2(const Fr9 &a0, const Fr9 &b0, const Fr9 &a1, const Fr9 &b1, Fr9 &o0,
                                        Fr9 &o1) {
  constexpr uint32_t P[9] = ZNTT_FR_P29;
  const uint32_t p0 = fr_opaque(P[0]);
  uint32_t m0[9], m1[9];
  uint64_t c0 = 0, c1 = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i <= k; i++) {
      c0 += (uint64_t)a0.v[i] * b0.v[k - i];
      fr_chain(c0);
      c1 += (uint64_t)a1.v[i] * b1.v[k - i];
      fr_chain(c1);
    }
#pragma unroll
    for (int i = 0; i < k; i++) {
      c0 += (uint64_t)m0[i] * P[k - i];
      fr_chain(c0);
      c1 += (uint64_t)m1[i] * P[k - i];
      fr_chain(c1);
    }
    m0[k] = ((uint32_t)c0 * kFrN0) & kM29;
    m1[k] = ((uint32_t)c1 * kFrN0) & kM29;
    c0 += (uint64_t)m0[k] * p0;
    fr_chain(c0);
    c1 += (uint64_t)m1[k] * p0;
    fr_chain(c1);
    c0 >>= 29;
    c1 >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i < 9; i++) {
      c0 += (uint64_t)a0.v[i] * b0.v[k - i];
      fr_chain(c0);
      c1 += (uint64_t)a1.v[i] * b1.v[k - i];
      fr_chain(c1);
      c0 += (uint64_t)m0[i] * P[k - i];
      fr_chain(c0);
      c1 += (uint64_t)m1[i] * P[k - i];
      fr_chain(c1);
    }
    o0.v[k - 9] = (uint32_t)c0 & kM29;
    o1.v[k - 9] = (uint32_t)c1 & kM29;
    c0 >>= 29;
    c1 >>= 29;
  }
  o0.v[8] = (uint32_t)c0;
  o1.v[8] = (uint32_t)c1;
}

// Carry-propagate limbs 0..7 into 29 bits; the top limb absorbs the rest.
__device__ __forceinline__ void fr_norm(Fr9 &x) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    x.v[i + 1] += x.v[i] >> 29;
    x.v[i] &= kM29;
  }
}

// Radix-2 DIT butterfly with lazy reduction: (u, v) <- (u + w v, u - w v).
// With NORM the u input is carry-normalised first.  Output limbs are below
// (u-limb bound) + 2^30, so normalising at every other stage keeps every limb
// that reaches fr_mul below 2.5 * 2^30 (< 3 * 2^30).
template <bool NORM = true>
__device__ __forceinline__ void fr_butterfly(Fr9 &u, Fr9 &v, const Fr9 &w) {
  constexpr uint32_t C[9] = ZNTT_FR_C4R;
  Fr9 t = fr_mul(v, w);
  if (NORM) fr_norm(u);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    uint32_t ui = u.v[i];
    u.v[i] = ui + t.v[i];
    v.v[i] = (ui + C[i]) - t.v[i];
  }
}

// Two independent butterflies with their products interleaved (fr_mul2).
template <bool NORM = true>
__device__ __forceinline__ void fr_butterfly2(Fr9 &u0, Fr9 &v0, const Fr9 &w0, Fr9 &u1, Fr9 &v1, const Fr9 &w1) {
  constexpr uint32_t C[9] = ZNTT_FR_C4R;
  Fr9 t0, t1;
  fr_mul2(v0, w0, v1, w1, t0, t1);
  if (NORM) {
    fr_norm(u0);
    fr_norm(u1);
  }
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint32_t a = u0.v[i], b = u1.v[i];
    u0.v[i] = a + t0.v[i];
    v0.v[i] = (a + C[i]) - t0.v[i];
    u1.v[i] = b + t1.v[i];
    v1.v[i] = (b + C[i]) - t1.v[i];
  }
}

// Butterfly with twiddle 1: (u, v) <- (u + v, u - v).  v must meet what
// fr_mul's output meets (carry-normalised, value < 2r): a freshly loaded or
// multiplied value.
template <bool NORM = true>
__device__ __forceinline__ void fr_butterfly_one(Fr9 &u, Fr9 &v) {
  constexpr uint32_t C[9] = ZNTT_FR_C4R;
  if (NORM) fr_norm(u);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint32_t ui = u.v[i], ti = v.v[i];
    u.v[i] = ui + ti;
    v.v[i] = (ui + C[i]) - ti;
  }
}

// 8r in a redundant radix-2^29 form with limbs 0..7 in [2^29 - 1, 2^30) and
// top limb 8r_8 + 2, above the top limb of any value below 4r.
#define ZNTT_FR_C8R {0x20000008u, 0x387d64fbu, 0x32e12286u, 0x3e84879au, 0x2c2e9418u, \
                     0x36da0604u, 0x25370a07u, 0x32e1319fu, 0x01832272u}

// Butterfly with twiddle 1 whose v is a sum of two loaded or multiplied
// values (value < 4r, limbs < 2^30): the second stage's j = 0 butterflies.
// v is carry-normalised first; 4r's redundant top limb can fall one short of
// such a v's, so the difference is u + 8r - v.  (Cheaper than a product by
// the twiddle 1: a normalisation instead of a Montgomery multiply.)
template <bool NORM = true>
__device__ __forceinline__ void fr_butterfly_one_c8(Fr9 &u, Fr9 &v) {
  constexpr uint32_t C[9] = ZNTT_FR_C8R;
  fr_norm(v);
  if (NORM) fr_norm(u);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint32_t ui = u.v[i], ti = v.v[i];
    u.v[i] = ui + ti;
    v.v[i] = (ui + C[i]) - ti;
  }
}

// x <- x - r if x >= r (x carry-normalised, limbs < 2^29, value < 2^261).
__device__ __forceinline__ void fr_csub_r(Fr9 &x) {
  constexpr uint32_t P[9] = ZNTT_FR_P29;
  Fr9 d;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    uint32_t s = x.v[i] - P[i] - borrow;
    borrow = s >> 31;
    d.v[i] = s & kM29;
  }
  // top limb: s = x8 - P8 - b is exact in 32 bits; no borrow => x >= r
  const bool ge = borrow == 0;
#pragma unroll
  for (int i = 0; i < 9; i++) x.v[i] = ge ? d.v[i] : x.v[i];
}

// Full reduction to the canonical representative in [0, r).
// Requires value < 2^261 (guaranteed by the lazy-reduction bound), so that
// the normalised top limb x8 = floor(x / 2^232) is below 2^29.
//
// Quotient estimate q = floor(x8 * M / 2^40), M = kFrQMagic:
//   * q <= floor(x / r): M * (r8 + 1) < 2^40 (r8 = floor(r / 2^232)), so
//     q <= x8 / (r8 + 1) < x8 * 2^232 / r <= x / r;
//   * floor(x / r) - q <= 1: checked exactly for every x8 < 2^29 at each
//     point where floor(x / r) steps up (the largest x of a top limb x8 has
//     floor(x / r) = floor(((x8 + 1) 2^232 - 1) / r), and q is
//     nondecreasing in x8).
// So x - q r lies in [0, 2r) and one conditional subtraction finishes.
__device__ __forceinline__ void fr_canon(Fr9 &x) {
  constexpr uint32_t P[9] = ZNTT_FR_P29;
  fr_norm(x);
  const uint32_t q = (uint32_t)(((uint64_t)x.v[8] * kFrQMagic) >> 40);
  int64_t acc = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    acc += (int64_t)x.v[i] - (int64_t)((uint64_t)q * P[i]);
    x.v[i] = (uint32_t)acc & kM29;
    acc >>= 29;  // arithmetic shift: signed carry
  }
  x.v[8] = (uint32_t)(x.v[8] | ((uint32_t)acc << 29));  // acc == 0 here: x - q r >= 0
  fr_csub_r(x);
}

// 8 little-endian u32 words (value < 2^256) -> nine 29-bit limbs.
__device__ __forceinline__ Fr9 fr_unpack8(const uint32_t w[8]) {
  Fr9 x;
  x.v[0] = w[0] & kM29;
  x.v[1] = __builtin_amdgcn_alignbit(w[1], w[0], 29) & kM29;
  x.v[2] = __builtin_amdgcn_alignbit(w[2], w[1], 26) & kM29;
  x.v[3] = __builtin_amdgcn_alignbit(w[3], w[2], 23) & kM29;
  x.v[4] = __builtin_amdgcn_alignbit(w[4], w[3], 20) & kM29;
  x.v[5] = __builtin_amdgcn_alignbit(w[5], w[4], 17) & kM29;
  x.v[6] = __builtin_amdgcn_alignbit(w[6], w[5], 14) & kM29;
  x.v[7] = __builtin_amdgcn_alignbit(w[7], w[6], 11) & kM29;
  x.v[8] = w[7] >> 8;
  return x;
}

// Canonical carry-normalised limbs (value < 2^256) -> 8 u32 words.
__device__ __forceinline__ void fr_pack8(const Fr9 &x, uint32_t w[8]) {
  w[0] = x.v[0] | (x.v[1] << 29);
  w[1] = (x.v[1] >> 3) | (x.v[2] << 26);
  w[2] = (x.v[2] >> 6) | (x.v[3] << 23);
  w[3] = (x.v[3] >> 9) | (x.v[4] << 20);
  w[4] = (x.v[4] >> 12) | (x.v[5] << 17);
  w[5] = (x.v[5] >> 15) | (x.v[6] << 14);
  w[6] = (x.v[6] >> 18) | (x.v[7] << 11);
  w[7] = (x.v[7] >> 21) | (x.v[8] << 8);
}

// Canonical negation r - x for a canonical x (0 stays 0).
__device__ __forceinline__ Fr9 fr_neg_canon(const Fr9 &x) {
  constexpr uint32_t P[9] = ZNTT_FR_P29;
  Fr9 d;
  uint32_t borrow = 0, nz = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    uint32_t s = P[i] - x.v[i] - borrow;
    borrow = s >> 31;
    d.v[i] = s & kM29;
    nz |= x.v[i];
  }
#pragma unroll
  for (int i = 0; i < 9; i++) d.v[i] = nz ? d.v[i] : 0u;
  return d;
}

}  // namespace zntt
