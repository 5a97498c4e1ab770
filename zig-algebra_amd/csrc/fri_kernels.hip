// FRI folding and query-opening kernels on gfx950.
//
// The fold of one round (libs/fri/src/root.zig:392-403) maps a layer of 2h
// values to h values; the arithmetic, two field products and a halving per
// value, is in fri_fold.hpp.  Per output the kernel moves two values in, one
// out and one table entry (g^-j, dev layout).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_traits.hpp"
#include "fri_fold.hpp"
#include "kernels.hpp"

namespace zntt {
namespace {

constexpr int kFoldBlock = 256;

template <class F>
__global__ __launch_bounds__(kFoldBlock) void fri_fold_kernel(const uint32_t *cur, uint32_t *next, uint64_t half,
                                                              const uint32_t *tab, FoldConsts k) {
  const uint64_t j = (uint64_t)blockIdx.x * kFoldBlock + threadIdx.x;
  if (j >= half) return;
  const auto a = F::load_user(cur, j), b = F::load_user(cur, j + half);
  const auto w = F::load_dev(tab, j);
  FoldOps<F>::store(next, j, FoldOps<F>::fold(a, b, w, F::from_words(k.alpha)));
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
                           const uint32_t alpha_dev[9], hipStream_t s) {
  FoldConsts k;
  for (int i = 0; i < 9; i++) k.alpha[i] = alpha_dev[i];
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
