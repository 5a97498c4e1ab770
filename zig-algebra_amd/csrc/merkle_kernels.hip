// Merkle commitment of a FRI layer on gfx950: BLAKE3 leaves over antipodal
// pairs and the binary tree above them.
//
// Reference: the FRI prover commits a layer f of n = 2^k values by
//   leaf j = BLAKE3(toBytes(f[j]) || toBytes(f[j + n/2])),  j < n/2
// (libs/fri/src/root.zig:283-292, :378-384) and builds
// MerkleTree.initFromHashes over the leaves
// (libs/merkle/src/merkle_tree.zig:147-182): a heap of 2 * cap nodes, cap the
// leaf count rounded up to a power of two, nodes[cap + i] = leaf i (BLAKE3 of
// the empty string past the last leaf), nodes[s + i] = BLAKE3(nodes[2(s+i)] ||
// nodes[2(s+i)+1]) for every level start s, root nodes[1], nodes[0] zero.
// toBytes is the canonical value in little-endian bytes: 32 for BN254 Fr (out
// of the Montgomery form, field.zig:1002-1016), 8 for Goldilocks, 4 for the
// 31-bit fields (field.zig:108, :244-252).
//
// Every hash input here is at most 64 bytes, so each BLAKE3 call is a single
// compression of one block (the first and last block of the only chunk,
// flagged ROOT).  That compression -- 7 rounds of 8 quarter-rounds, 12 VALU
// instructions each with v_add3_u32 and v_alignbit_b32 rotates -- is the whole
// cost: about n compressions per layer of n values against 2 to 32 bytes of
// HBM traffic per value, so the kernel is VALU-bound.
//
// Layout of the work: one workgroup of 256 threads owns up to 512 consecutive
// children of the level it starts from.  Each thread forms two children (leaf
// hashes, or stored nodes) and their parent in registers; further levels
// reduce through LDS.  A launch over a level larger than 512 nodes climbs
// three levels (512 -> 64 nodes per workgroup, whole waves busy at every
// level); the launch over the last 512 or fewer climbs to the root, so a
// layer of 2^23 leaves takes six launches (2^23 -> 2^20 -> 2^17 -> 2^14 ->
// 2^11 -> 2^8 -> 1).  Nodes are written to the caller's heap when it asks for
// the whole tree; otherwise only the first launch's top level and what lies
// above it are kept, in a heap of 2 * 2^(k - 1 - 3) nodes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_traits.hpp"
#include "fri_fold.hpp"
#include "kernels.hpp"

namespace zntt {
namespace {

constexpr int kMerkleBlock = 256;
constexpr int kMerkleMaxLevels = 9;  // log2(2 * kMerkleBlock)

constexpr uint32_t kIv[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                             0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
// CHUNK_START | CHUNK_END | ROOT: the one block of a hash of at most 64 bytes.
constexpr uint32_t kOneBlockFlags = 1u | 2u | 8u;

// Message word that feeds input slot i of round r: the BLAKE3 message
// permutation applied r times.
struct Schedule {
  uint8_t w[7][16];
};
constexpr Schedule make_schedule() {
  constexpr uint8_t perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  Schedule s{};
  for (int i = 0; i < 16; i++) s.w[0][i] = (uint8_t)i;
  for (int r = 1; r < 7; r++)
    for (int i = 0; i < 16; i++) s.w[r][i] = s.w[r - 1][perm[i]];
  return s;
}

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return __builtin_amdgcn_alignbit(x, x, n); }

__device__ __forceinline__ void quarter(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t x, uint32_t y) {
  a = a + b + x;
  d = rotr(d ^ a, 16);
  c = c + d;
  b = rotr(b ^ c, 12);
  a = a + b + y;
  d = rotr(d ^ a, 8);
  c = c + d;
  b = rotr(b ^ c, 7);
}

template <int R>
__device__ __forceinline__ void round(uint32_t (&v)[16], const uint32_t (&m)[16]) {
  constexpr Schedule s = make_schedule();
  quarter(v[0], v[4], v[8], v[12], m[s.w[R][0]], m[s.w[R][1]]);
  quarter(v[1], v[5], v[9], v[13], m[s.w[R][2]], m[s.w[R][3]]);
  quarter(v[2], v[6], v[10], v[14], m[s.w[R][4]], m[s.w[R][5]]);
  quarter(v[3], v[7], v[11], v[15], m[s.w[R][6]], m[s.w[R][7]]);
  quarter(v[0], v[5], v[10], v[15], m[s.w[R][8]], m[s.w[R][9]]);
  quarter(v[1], v[6], v[11], v[12], m[s.w[R][10]], m[s.w[R][11]]);
  quarter(v[2], v[7], v[8], v[13], m[s.w[R][12]], m[s.w[R][13]]);
  quarter(v[3], v[4], v[9], v[14], m[s.w[R][14]], m[s.w[R][15]]);
}

// BLAKE3 (32-byte output) of a message of `len` <= 64 bytes held as 16
// little-endian words, zero past `len`.
__device__ __forceinline__ void hash_block(const uint32_t (&m)[16], uint32_t len, uint32_t out[8]) {
  uint32_t v[16] = {kIv[0], kIv[1], kIv[2], kIv[3], kIv[4], kIv[5], kIv[6], kIv[7],
                    kIv[0], kIv[1], kIv[2], kIv[3], 0u, 0u, len, kOneBlockFlags};
  round<0>(v, m);
  round<1>(v, m);
  round<2>(v, m);
  round<3>(v, m);
  round<4>(v, m);
  round<5>(v, m);
  round<6>(v, m);
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = v[i] ^ v[8 + i];
}

__device__ __forceinline__ void load8(const uint32_t *p, uint32_t w[8]) {
  const uint4 a = reinterpret_cast<const uint4 *>(p)[0], b = reinterpret_cast<const uint4 *>(p)[1];
  w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
  w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}

__device__ __forceinline__ void store8(uint32_t *p, const uint32_t w[8]) {
  reinterpret_cast<uint4 *>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
  reinterpret_cast<uint4 *>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// toBytes of a BN254 Fr element stored in the reference's Montgomery form
// x * 2^256: mont(x * 2^256, 32) = x * 2^261 / 2^261 = x with R' = 2^261.
// The product is below (2^256 * 32) / 2^261 + r = r + 1, so one conditional
// subtraction of r leaves the canonical representative (r itself only for a
// stored value of r, the unreduced zero).
__device__ __forceinline__ void fr_canonical_words(const Fr9 &mont, uint32_t w[8]) {
  Fr9 k;
#pragma unroll
  for (int q = 0; q < 9; q++) k.v[q] = 0;
  k.v[0] = 32;
  Fr9 x = fr_mul(mont, k);
  fr_csub_r(x);
  fr_pack8(x, w);
}

// Children of the first launch over a FRI layer: leaf j of the pairs
// (j, j + half), half = the leaf count.  get2 forms leaves j and j + 1 (the
// two children of one parent) with every load issued before any hashing.
template <class F>
struct LayerLeaves;

template <>
struct LayerLeaves<FrT> {
  const uint32_t *evals;
  uint64_t half;
  __device__ void get(uint64_t j, uint32_t h[8]) const {
    uint32_t m[16];
    fr_canonical_words(FrT::load_user(evals, j), m);
    fr_canonical_words(FrT::load_user(evals, j + half), m + 8);
    hash_block(m, 64, h);
  }
  __device__ void get2(uint64_t j, uint32_t (&out)[16]) const {
    const Fr9 a0 = FrT::load_user(evals, j), b0 = FrT::load_user(evals, j + half);
    const Fr9 a1 = FrT::load_user(evals, j + 1), b1 = FrT::load_user(evals, j + 1 + half);
    uint32_t m[16];
    fr_canonical_words(a0, m);
    fr_canonical_words(b0, m + 8);
    hash_block(m, 64, out);
    fr_canonical_words(a1, m);
    fr_canonical_words(b1, m + 8);
    hash_block(m, 64, out + 8);
  }
};

template <>
struct LayerLeaves<GlT> {
  const uint32_t *evals;
  uint64_t half;
  __device__ static void leaf(uint2 a, uint2 b, uint32_t h[8]) {
    const uint32_t m[16] = {a.x, a.y, b.x, b.y, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    hash_block(m, 16, h);
  }
  __device__ void get(uint64_t j, uint32_t h[8]) const {
    const uint2 *e = reinterpret_cast<const uint2 *>(evals);
    leaf(e[j], e[j + half], h);
  }
  __device__ void get2(uint64_t j, uint32_t (&out)[16]) const {
    const uint4 a = *reinterpret_cast<const uint4 *>(evals + 2 * j);
    const uint4 b = *reinterpret_cast<const uint4 *>(evals + 2 * (j + half));
    leaf(make_uint2(a.x, a.y), make_uint2(b.x, b.y), out);
    leaf(make_uint2(a.z, a.w), make_uint2(b.z, b.w), out + 8);
  }
};

// BabyBear and KoalaBear: 4 bytes each, the low word of the stored u64.
struct LayerLeaves31 {
  const uint32_t *evals;
  uint64_t half;
  __device__ static void leaf(uint32_t a, uint32_t b, uint32_t h[8]) {
    const uint32_t m[16] = {a, b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    hash_block(m, 8, h);
  }
  __device__ void get(uint64_t j, uint32_t h[8]) const { leaf(evals[2 * j], evals[2 * (j + half)], h); }
  __device__ void get2(uint64_t j, uint32_t (&out)[16]) const {
    const uint4 a = *reinterpret_cast<const uint4 *>(evals + 2 * j);
    const uint4 b = *reinterpret_cast<const uint4 *>(evals + 2 * (j + half));
    leaf(a.x, b.x, out);
    leaf(a.z, b.z, out + 8);
  }
};

// Leaf words of a fold output (canonical): its toBytes words (BN254 Fr: out
// of the Montgomery form).
__device__ __forceinline__ int leaf_words(const FrT::E &v, uint32_t *w) {
  fr_canonical_words(v, w);
  return 8;
}
__device__ __forceinline__ int leaf_words(uint64_t v, uint32_t *w) {
  w[0] = (uint32_t)v;
  w[1] = (uint32_t)(v >> 32);
  return 2;
}
__device__ __forceinline__ int leaf_words(uint32_t v, uint32_t *w) {
  w[0] = v;
  return 1;
}

// Children of a launch that folds the previous FRI layer and commits the
// new one: leaf j of the new layer (2 * half values) hashes new values j and
// j + half, each folded here from the previous layer's pair (x, x + 2 half)
// and stored to the new layer on the way.  Every new value is folded once:
// get2(j) covers j, j + 1, j + half and j + 1 + half.
template <class F>
struct FoldLeaves {
  using E = typename F::E;
  const uint32_t *prev;
  uint32_t *next;
  const uint32_t *inv_pows;  // g^-x, dev layout
  uint64_t half;             // leaf count of the new layer
  FoldConsts k;
  __device__ E fold_at(uint64_t x) const {
    const E a = F::load_user(prev, x), b = F::load_user(prev, x + 2 * half);
    const E v = FoldOps<F>::fold(a, b, F::load_dev(inv_pows, x), F::from_words(k.alpha));
    FoldOps<F>::store(next, x, v);
    return v;
  }
  __device__ static void leaf(const E &u, const E &v, uint32_t h[8]) {
    uint32_t m[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int w = leaf_words(u, m);
    leaf_words(v, m + w);
    hash_block(m, 8 * w, h);
  }
  __device__ void get(uint64_t j, uint32_t h[8]) const { leaf(fold_at(j), fold_at(j + half), h); }
  __device__ void get2(uint64_t j, uint32_t (&out)[16]) const {
    get(j, out);
    get(j + 1, out + 8);
  }
};

// Children that are already hashes (count of them, then BLAKE3("") padding).
struct StoredNodes {
  const uint32_t *nodes;
  uint64_t count;
  __device__ void get(uint64_t j, uint32_t h[8]) const {
    if (j < count) {
      load8(nodes + 8 * j, h);
    } else {
      // BLAKE3 of the empty string, little-endian words.
      constexpr uint32_t kEmpty[8] = {0xb94913afu, 0xa6a1f9f5u, 0xea4d40a0u, 0x49c9dc36u,
                                      0xc925cb9bu, 0xb712c1adu, 0xca939accu, 0x62321fe4u};
#pragma unroll
      for (int i = 0; i < 8; i++) h[i] = kEmpty[i];
    }
  }
  __device__ void get2(uint64_t j, uint32_t (&out)[16]) const {
    get(j, out);
    get(j + 1, out + 8);
  }
};

// One launch over a level of 2^child_log nodes: each workgroup owns 2^span_log
// consecutive children and climbs `levels` levels above them, ending with
// 2^(span_log - levels) nodes.  heap is the reference's node array (node s + i
// at heap + 8 * (s + i)); children are written at their heap positions when
// write_children, inner nodes when write_inner, and the nodes of the last
// level climbed always.  levels == 0: the single child is the root.
template <class Src>
__global__ __launch_bounds__(kMerkleBlock) void merkle_kernel(Src src, uint32_t *heap, int child_log, int span_log,
                                                              int levels, int write_children, int write_inner,
                                                              int zero_node0) {
  __shared__ uint32_t lds[8 * kMerkleBlock];
  const uint32_t t = threadIdx.x;
  const uint64_t C = 1ull << child_log;
  if (zero_node0 && blockIdx.x == 0 && t < 8) heap[t] = 0;
  if (levels == 0) {
    if (blockIdx.x == 0 && t == 0) {
      uint32_t h[8];
      src.get(0, h);
      store8(heap + 8, h);
    }
    return;
  }
  const uint64_t base = (uint64_t)blockIdx.x << span_log;  // first child of this workgroup
  uint32_t cnt = 1u << (span_log - 1);                       // nodes one level up
  uint32_t h[8];
  if (t < cnt) {
    uint32_t m[16];
    const uint64_t j = base + 2 * t;
    src.get2(j, m);
    if (write_children) {
      store8(heap + 8 * (C + j), m);
      store8(heap + 8 * (C + j + 1), m + 8);
    }
    hash_block(m, 64, h);
    if (write_inner || levels == 1) store8(heap + 8 * ((C >> 1) + (base >> 1) + t), h);
  }
  for (int l = 2; l <= levels; l++) {
    if (t < cnt) {
#pragma unroll
      for (int k = 0; k < 8; k++) lds[k * kMerkleBlock + t] = h[k];
    }
    __syncthreads();
    cnt >>= 1;
    if (t < cnt) {
      uint32_t m[16];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        m[k] = lds[k * kMerkleBlock + 2 * t];
        m[8 + k] = lds[k * kMerkleBlock + 2 * t + 1];
      }
      hash_block(m, 64, h);
      if (write_inner || l == levels) store8(heap + 8 * ((C >> l) + (base >> l) + t), h);
    }
    __syncthreads();
  }
}

template <class Src>
hipError_t launch_level(const Src &src, uint32_t *heap, int child_log, int span_log, int levels,
                        bool write_children, bool write_inner, bool zero_node0, hipStream_t s) {
  const unsigned grid = (unsigned)(1ull << (child_log - span_log));
  hipLaunchKernelGGL(merkle_kernel<Src>, dim3(grid), dim3(kMerkleBlock), 0, s, src, heap, child_log, span_log,
                     levels, (int)write_children, (int)write_inner, (int)zero_node0);
  return hipGetLastError();
}

// Levels one launch climbs over a level of 2^log nodes.  A level that one
// workgroup covers is finished by it.  A larger one stops where each
// workgroup still has 64 nodes (512 -> 256 -> 128 -> 64): every level climbed
// fills whole waves, and the next launch starts from an eighth of the nodes.
// Climbing on to one node per workgroup would run six levels on waves with
// 32, 16, ... 1 busy lanes.
int climb_levels(int log) { return log <= kMerkleMaxLevels ? log : 3; }

// Finishes a tree whose level of 2^top_log nodes is already in heap.
hipError_t finish_tree(uint32_t *heap, int top_log, hipStream_t s) {
  while (top_log > 0) {
    const int span = top_log < kMerkleMaxLevels ? top_log : kMerkleMaxLevels;
    const int lv = climb_levels(top_log);
    StoredNodes src{heap + 8 * (1ull << top_log), 1ull << top_log};
    hipError_t e = launch_level(src, heap, top_log, span, lv, false, true, top_log == lv, s);
    if (e != hipSuccess) return e;
    top_log -= lv;
  }
  return hipSuccess;
}

template <class Src>
hipError_t build_tree(const Src &src, int cap_log, uint32_t *heap, bool full, hipStream_t s) {
  const int span = cap_log < kMerkleMaxLevels ? cap_log : kMerkleMaxLevels;
  const int lv = climb_levels(cap_log);
  hipError_t e = launch_level(src, heap, cap_log, span, lv, full, full, cap_log == lv, s);
  if (e != hipSuccess) return e;
  return finish_tree(heap, cap_log - lv, s);
}

}  // namespace

int merkle_heap_log(int cap_log, bool full) { return full ? cap_log + 1 : cap_log - climb_levels(cap_log) + 1; }

hipError_t launch_merkle_layer(int field, const void *evals, int leaf_log, uint32_t *heap, bool full,
                               hipStream_t s) {
  const uint32_t *e = static_cast<const uint32_t *>(evals);
  const uint64_t half = 1ull << leaf_log;
  if (field == kFr) return build_tree(LayerLeaves<FrT>{e, half}, leaf_log, heap, full, s);
  if (field == kGl) return build_tree(LayerLeaves<GlT>{e, half}, leaf_log, heap, full, s);
  return build_tree(LayerLeaves31{e, half}, leaf_log, heap, full, s);
}

hipError_t launch_merkle_fold(int field, const void *prev, void *next, int leaf_log, uint32_t *heap,
                              const uint32_t *inv_pows, const uint32_t alpha_dev[9], hipStream_t s) {
  FoldConsts k;
  for (int i = 0; i < 9; i++) k.alpha[i] = alpha_dev[i];
  const uint32_t *p = static_cast<const uint32_t *>(prev);
  uint32_t *nx = static_cast<uint32_t *>(next);
  const uint64_t half = 1ull << leaf_log;
  return with_field(field, [&](auto f) {
    using F = decltype(f);
    return build_tree(FoldLeaves<F>{p, nx, inv_pows, half, k}, leaf_log, heap, true, s);
  });
}

hipError_t launch_merkle_from_hashes(const void *hashes, uint64_t count, int cap_log, uint32_t *heap, bool full,
                                     hipStream_t s) {
  return build_tree(StoredNodes{static_cast<const uint32_t *>(hashes), count}, cap_log, heap, full, s);
}

}  // namespace zntt
