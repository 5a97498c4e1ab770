// The FRI fold of one antipodal pair, shared by the fold kernel
// (fri_kernels.hip) and the fold-and-commit launch (merkle_kernels.hip).
//
// next[j] = (f[j] + f[j+h]) / 2 + alpha (f[j] - f[j+h]) / (2 x_j)
// (libs/fri/src/root.zig:392-403), computed as
//   next[j] = half((f[j] + f[j+h]) + ((f[j] - f[j+h]) * alpha) * g^-j)
// with x_j = g^j and half(y) = y / 2 mod p: y even -> y >> 1, y odd ->
// (y + p) >> 1.  Two field products per value; halving is linear, so it acts
// on BN254 Fr's Montgomery form x * 2^256 as on a plain value.  The
// arithmetic is exact, so this grouping gives the reference's values.
// Outputs are canonical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_traits.hpp"

namespace zntt {

struct FoldConsts {
  uint32_t alpha[9];  // alpha, dev form
};

template <class F>
struct FoldOps;

template <>
struct FoldOps<GlT> {
  __device__ static uint64_t half(uint64_t y) {  // y < p
    return (y >> 1) + ((y & 1) ? (kGlP >> 1) + 1 : 0);
  }
  __device__ static uint64_t fold(uint64_t a, uint64_t b, uint64_t w, uint64_t alpha) {
    const uint64_t s = gl_add(a, b), d = gl_sub(a, b);
    return half(gl_add(s, gl_mul(gl_mul(d, alpha), w)));
  }
  __device__ static void store(uint32_t *base, size_t i, uint64_t v) { GlT::store_user(base, i, v); }
};

template <uint32_t P, uint32_t NP>
struct FoldOps<F31T<P, NP>> {
  using A = F31Arith<P, NP>;
  __device__ static uint32_t half(uint32_t y) { return (y + ((y & 1) ? P : 0)) >> 1; }  // y < p < 2^31
  // w and alpha are Montgomery forms c * 2^32, the data plain, so
  // mont(data, constant) is plain.
  __device__ static uint32_t fold(uint32_t a, uint32_t b, uint32_t w, uint32_t alpha) {
    const uint32_t s = A::add(a, b), d = A::sub(a, b);
    return half(A::add(s, A::mont(A::mont(d, alpha), w)));
  }
  __device__ static void store(uint32_t *base, size_t i, uint32_t v) { F31T<P, NP>::store_user(base, i, v); }
};

// BN254 Fr.  s = a + b and d = a + 4r - b stay below 2^261 with limbs under
// 3 * 2^30 (a and b freshly unpacked), the bounds of fr_mul's first operand;
// d * alpha (< 2r, carry-normalised) is again a valid first operand against
// the canonical table entry.  The sum (< 4r) is reduced fully, then halved:
// (y + r * (y & 1)) >> 1 across the 29-bit limbs.
template <>
struct FoldOps<FrT> {
  __device__ static Fr9 half(Fr9 y) {  // y canonical
    constexpr uint32_t P[9] = ZNTT_FR_P29;
    const uint32_t odd = 0u - (y.v[0] & 1u);
#pragma unroll
    for (int i = 0; i < 9; i++) y.v[i] += P[i] & odd;
    fr_norm(y);
#pragma unroll
    for (int i = 0; i < 8; i++) y.v[i] = (y.v[i] >> 1) | ((y.v[i + 1] & 1u) << 28);
    y.v[8] >>= 1;
    return y;
  }
  __device__ static Fr9 fold(const Fr9 &a, const Fr9 &b, const Fr9 &w, const Fr9 &alpha) {
    constexpr uint32_t C[9] = ZNTT_FR_C4R;
    Fr9 s, d;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      s.v[i] = a.v[i] + b.v[i];
      d.v[i] = (a.v[i] + C[i]) - b.v[i];
    }
    const Fr9 t = fr_mul(fr_mul(d, alpha), w);
#pragma unroll
    for (int i = 0; i < 9; i++) s.v[i] += t.v[i];
    fr_canon(s);
    return half(s);
  }
  // v canonical: repack without another reduction.
  __device__ static void store(uint32_t *base, size_t i, const Fr9 &v) {
    uint32_t w[8];
    fr_pack8(v, w);
    uint4 *p = reinterpret_cast<uint4 *>(base + i * 8);
    p[0] = make_uint4(w[0], w[1], w[2], w[3]);
    p[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
};

}  // namespace zntt
