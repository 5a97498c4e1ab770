// FRI folding and proof-gathering kernels on gfx950.
//
// The fold of one round (libs/fri/src/root.zig:392-403) maps a layer f of
// 2h values to h values:
//   next[j] = (f[j] + f[j+h]) / 2 + alpha * (f[j] - f[j+h]) / (2 x_j),
// x_j = g^j, g the generator of the layer's domain.  The kernel reads
// g^-j from a table of powers (dev layout) and multiplies by two constants,
// alpha/2 and 1/2, so each output costs three field products:
//   w = g^-j * (alpha/2),  next[j] = (f[j] + f[j+h]) * (1/2) + (f[j] - f[j+h]) * w.
// Field arithmetic is exact, so this grouping gives the reference's values.
// Per output it moves 2 values in, 1 value out and one table entry: HBM-bound
// for every field (BN254 Fr: 132 B against three radix-2^29 products).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_traits.hpp"
#include "kernels.hpp"

namespace zntt {
namespace {

constexpr int kFoldBlock = 256;

struct FoldConsts {
  uint32_t alpha_half[9];  // alpha / 2, dev form
  uint32_t half[9];        // 1 / 2, dev form
};

// Canonical-input fold for the single-word fields (Goldilocks, 31-bit).
template <class F>
struct FoldOps {
  using E = typename F::E;
  __device__ static E fold(E a, E b, E w, E ah, E hf);
};

template <>
struct FoldOps<GlT> {
  __device__ static uint64_t fold(uint64_t a, uint64_t b, uint64_t w, uint64_t ah, uint64_t hf) {
    const uint64_t s = gl_add(a, b), d = gl_sub(a, b);
    return gl_add(gl_mul(s, hf), gl_mul(d, gl_mul(w, ah)));
  }
};

template <uint32_t P, uint32_t NP>
struct FoldOps<F31T<P, NP>> {
  using A = F31Arith<P, NP>;
  // w and the constants are Montgomery forms c * 2^32, the data plain, so
  // mont(w, ah) is again a Montgomery constant and mont(data, constant) plain.
  __device__ static uint32_t fold(uint32_t a, uint32_t b, uint32_t w, uint32_t ah, uint32_t hf) {
    const uint32_t s = A::add(a, b), d = A::sub(a, b);
    return A::add(A::mont(s, hf), A::mont(d, A::mont(w, ah)));
  }
};

// BN254 Fr with lazy sums: s = a + b and d = a + 4r - b stay below 2^261
// with limbs under 3 * 2^30 (a and b are freshly unpacked, limbs < 2^29), the
// bounds fr_mul accepts for its first operand.  mont(w, ah) < 2r is brought
// below r (fr_mul's second operand), and the sum of the two products (< 4r)
// is reduced fully by the store.
template <>
struct FoldOps<FrT> {
  __device__ static Fr9 fold(const Fr9 &a, const Fr9 &b, const Fr9 &w, const Fr9 &ah, const Fr9 &hf) {
    constexpr uint32_t C[9] = ZNTT_FR_C4R;
    Fr9 s, d;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      s.v[i] = a.v[i] + b.v[i];
      d.v[i] = (a.v[i] + C[i]) - b.v[i];
    }
    Fr9 wc = fr_mul(w, ah);
    fr_csub_r(wc);
    Fr9 u, t;
    fr_mul2(s, hf, d, wc, u, t);
    Fr9 o;
#pragma unroll
    for (int i = 0; i < 9; i++) o.v[i] = u.v[i] + t.v[i];
    return o;
  }
};

template <class F>
__global__ __launch_bounds__(kFoldBlock) void fri_fold_kernel(const uint32_t *cur, uint32_t *next, uint64_t half,
                                                              const uint32_t *tab, FoldConsts k) {
  const uint64_t j = (uint64_t)blockIdx.x * kFoldBlock + threadIdx.x;
  if (j >= half) return;
  const auto a = F::load_user(cur, j), b = F::load_user(cur, j + half);
  const auto w = F::load_dev(tab, j);
  F::store_user(next, j, FoldOps<F>::fold(a, b, w, F::from_words(k.alpha_half), F::from_words(k.half)));
}

// out[i] = the `words` u32 words at device address addrs[i].
__global__ __launch_bounds__(256) void gather_kernel(const uint64_t *addrs, uint64_t count, int words,
                                                     uint32_t *out) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(addrs[i]);
  for (int k = 0; k < words; k++) out[i * words + k] = src[k];
}

}  // namespace

hipError_t launch_fri_fold(int field, const void *cur, void *next, uint64_t half, const uint32_t *inv_pows,
                           const uint32_t alpha_half_dev[9], const uint32_t half_dev[9], hipStream_t s) {
  FoldConsts k;
  for (int i = 0; i < 9; i++) {
    k.alpha_half[i] = alpha_half_dev[i];
    k.half[i] = half_dev[i];
  }
  const unsigned grid = (unsigned)((half + kFoldBlock - 1) / kFoldBlock);
  const uint32_t *c = static_cast<const uint32_t *>(cur);
  uint32_t *nx = static_cast<uint32_t *>(next);
  with_field(field, [&](auto f) {
    using F = decltype(f);
    hipLaunchKernelGGL(fri_fold_kernel<F>, dim3(grid), dim3(kFoldBlock), 0, s, c, nx, half, inv_pows, k);
  });
  return hipGetLastError();
}

hipError_t launch_gather(const uint64_t *addrs, uint64_t count, int words, uint32_t *out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const unsigned grid = (unsigned)((count + 255) / 256);
  hipLaunchKernelGGL(gather_kernel, dim3(grid), dim3(256), 0, s, addrs, count, words, out);
  return hipGetLastError();
}

}  // namespace zntt
