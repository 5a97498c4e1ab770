// FRI folding and query-opening kernels on gfx950.
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

// Query openings (buildQueries, libs/fri/src/root.zig:556-569): thread t
// serves query k = t / R in round r = t % R.  The pair slot is
// j = e0 mod half_r (half_r = 2^(ld - r - 1) divides every earlier half, so
// reducing the previous round's slot gives the same), the values are
// elements j and j + half_r of layer r, the path the siblings of leaf j from
// the leaf level up (MerkleTree.prove).  Paths of query k are stored round
// by round; round r starts sum_{r' < r} (ld - r' - 1) nodes in.
__global__ __launch_bounds__(256) void fri_open_kernel(FriOpenArgs a, const uint64_t *e0, uint64_t q,
                                                       uint32_t *vals, uint32_t *paths) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= q * a.rounds) return;
  const uint64_t k = t / a.rounds;
  const uint32_t r = (uint32_t)(t % a.rounds);
  const uint64_t half = 1ull << (a.log_domain - r - 1);
  const uint64_t j = e0[k] & (half - 1);
  const uint32_t w = a.elem_words;
  const uint32_t *layer = reinterpret_cast<const uint32_t *>(a.layer[r]);
  uint32_t *v = vals + t * 2 * w;
  for (uint32_t i = 0; i < w; i++) {
    v[i] = layer[j * w + i];
    v[w + i] = layer[(j + half) * w + i];
  }
  const uint64_t per_query = (uint64_t)a.rounds * (a.log_domain - 1) - (uint64_t)a.rounds * (a.rounds - 1) / 2;
  const uint64_t before = (uint64_t)r * (a.log_domain - 1) - (uint64_t)r * (r - 1) / 2;
  uint4 *p = reinterpret_cast<uint4 *>(paths) + 2 * (k * per_query + before);
  const uint4 *tree = reinterpret_cast<const uint4 *>(a.tree[r]);
  for (uint64_t pos = half + j; pos > 1; pos >>= 1) {
    p[0] = tree[2 * (pos ^ 1)];
    p[1] = tree[2 * (pos ^ 1) + 1];
    p += 2;
  }
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

hipError_t launch_fri_open(const FriOpenArgs &a, const uint64_t *e0, uint64_t q, uint32_t *vals, uint32_t *paths,
                           hipStream_t s) {
  const uint64_t threads = q * a.rounds;
  if (threads == 0) return hipSuccess;
  const unsigned grid = (unsigned)((threads + 255) / 256);
  hipLaunchKernelGGL(fri_open_kernel, dim3(grid), dim3(256), 0, s, a, e0, q, vals, paths);
  return hipGetLastError();
}

}  // namespace zntt
