// Goldilocks instantiation of the pass kernels.
#include <algorithm>
#include <mutex>

#include "pass_kernels.hpp"

namespace zntt {

hipError_t launch_transform_gl(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw, const uint32_t *coset,
                               const uint32_t *scale_dev, bool tw_one, hipStream_t stream) {
  static std::once_flag once;
  static hipError_t configured = hipSuccess;
  std::call_once(once, [] { configured = configure_pass_lds<GlT>(2 * 4 * (1 << 12)); });
  if (configured != hipSuccess) return configured;
  return run_passes_t<GlT>(kGl, L, data, scratch, tw, coset, scale_dev, tw_one, stream);
}

// ---------------------------------------------------------------------------
// Goldilocks whole transforms (2^9, 2^12) as radix-8 steps.
//
// After the first 3r stages of a DIT over bit-reversed input, each block of
// 2^l0 positions (l0 = 3r) holds a 2^l0-point transform; the next three stages
// combine eight blocks.  With Y_m the block at position brev3(k) = m of the
// eight and b < 2^l0, the combined transform is
//     X[b + 2^l0 q] = sum_m w8^(q m) (w_(2^(l0+3))^(b m) Y_m[b]),
// an 8-point DFT of the block values after a pre-twiddle
// w_(2^(l0+3))^(b m) = w^((b m) 2^(K-3-l0)).  The 8-point DFT's twiddles
// are the powers of w8, and every primitive 8th root of unity mod p is
// +-2^24 or +-2^72: its products are shifts and one reduction
// (gl_shl24/48/72), and only the seven pre-twiddles per 8 values and round
// are general products (twelve in the plain DIT round).  Same DAG values,
// so the result is the DIT result bit for bit.
// ---------------------------------------------------------------------------

// Power tables in LDS, after the tile's two planes: P[e] = w^e for e <
// 2^K (two u32 planes of 2^K), then for each round l0 < K - 3 a compact
// table T_l0[j] = w^(j 2^(K-3-l0)) = P[j << (K-3-l0)], j < 2^(l0+3) (two
// planes of 2^(l0+3)).  A round's pre-twiddle reads, lane by lane, entry
// b m of its table, b the lane's low position bits: consecutive b, so the
// lanes of a read spread over the banks (in P itself, the reads of l0 = 3
// would be multiples of 64 entries apart, all in one bank).
__host__ __device__ constexpr int gl8_table_off(int K, int l0) {
  // u32 offset of the round's table from the start of P
  return l0 + 3 >= K ? 0 : 2 * (1 << K) + (l0 == 3 ? 0 : 2 * 64);
}
__host__ __device__ constexpr int gl8_tables_words(int K) {
  return 2 * (1 << K) + (K >= 9 ? 2 * 64 : 0) + (K >= 12 ? 2 * 512 : 0);
}

// Butterfly (u, v) <- (u + s 2^E v, u - s 2^E v), s = -1 when NEG.
template <int E, bool NEG>
__device__ __forceinline__ void gl_bfly_pow2(uint64_t &u, uint64_t &v) {
  const uint64_t t = E == 24 ? gl_shl24(v) : (E == 48 ? gl_shl48(v) : gl_shl72(v));
  uint64_t a = u, b;
  gl_butterfly_t(a, b, t);
  u = NEG ? b : a;
  v = NEG ? a : b;
}

// Butterfly with twiddle 1, v any u64.
__device__ __forceinline__ void gl_bfly_one(uint64_t &u, uint64_t &v) {
  const uint64_t t = gl_canon(v);
  gl_butterfly_t(u, v, t);
}

// Butterfly with a general twiddle w (canonical).
__device__ __forceinline__ void gl_bfly_w(uint64_t &u, uint64_t &v, uint64_t w) {
  const uint64_t t = gl_mul(v, w);
  gl_butterfly_t(u, v, t);
}

// 8-point DIT in registers: slot k holds the input of index brev3(k), slot q
// receives output q.  V: w8 = 2^24 (0), -2^24 (1), 2^72 (2), -2^72 (3), or
// the twiddles w8, w4 = w8^2, w83 = w8^3 as general products (4).  In V 0
// and 1, w4 = 2^48 and w8^3 = +-2^72; in V 2 and 3, w4 = -2^48 and w8^3 =
// +-2^24.  CANON1: the stage-1 v values may be >= p (else they are
// products, below p).
template <int V, bool CANON1>
__device__ __forceinline__ void gl_dft8(uint64_t (&x)[8], uint64_t w8, uint64_t w4, uint64_t w83) {
#pragma unroll
  for (int k = 0; k < 8; k += 2) {
    if (CANON1)
      gl_bfly_one(x[k], x[k + 1]);
    else
      gl_butterfly_t(x[k], x[k + 1], x[k + 1]);
  }
  if constexpr (V == 4) {
    gl_bfly_one(x[0], x[2]);
    gl_bfly_w(x[1], x[3], w4);
    gl_bfly_one(x[4], x[6]);
    gl_bfly_w(x[5], x[7], w4);
    gl_bfly_one(x[0], x[4]);
    gl_bfly_w(x[1], x[5], w8);
    gl_bfly_w(x[2], x[6], w4);
    gl_bfly_w(x[3], x[7], w83);
  } else {
    constexpr int E8 = V < 2 ? 24 : 72, E83 = V < 2 ? 72 : 24;
    constexpr bool N8 = V == 1 || V == 3, N4 = V >= 2;
    gl_bfly_one(x[0], x[2]);
    gl_bfly_pow2<48, N4>(x[1], x[3]);
    gl_bfly_one(x[4], x[6]);
    gl_bfly_pow2<48, N4>(x[5], x[7]);
    gl_bfly_one(x[0], x[4]);
    gl_bfly_pow2<E8, N8>(x[1], x[5]);
    gl_bfly_pow2<48, N4>(x[2], x[6]);
    gl_bfly_pow2<E83, N8>(x[3], x[7]);
  }
}

// One radix-8 round (stages l0+1 .. l0+3) of the whole transform: thread t
// holds positions base | k << l0, k < 8, at LDS slots whole_swz(base) ^
// whole_swz(k << l0); P (LDS, two planes of 2^K) holds w^e for e < 2^K.
template <int K, int l0, int V>
__device__ __forceinline__ void gl8_round(uint64_t (&x)[8], const PassArgs &a, uint32_t *lds, const uint32_t *P,
                                          const TileCtx &c, uint64_t w8, uint64_t w4, uint64_t w83) {
  constexpr bool first = l0 == 0, last = l0 + 3 == K;
  const int slot = threadIdx.x;
  const int base = (slot & ((1 << l0) - 1)) | ((slot >> l0) << (l0 + 3));
  const int sb = whole_swz(base);
  if constexpr (first) {
    stage_whole_tile<GlT, 8, true, true>(a, lds, c);
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = GlT::lds_get(lds, 1 << K, sb ^ whole_swz(k << l0));
  if constexpr (first) __syncthreads();  // every staged slot is read before round outputs overwrite LDS
  if constexpr (l0 > 0) {
    const int b = base & ((1 << l0) - 1);
#pragma unroll
    for (int k = 1; k < 8; k++) {
      const int m = ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1);  // brev3(k)
      constexpr int S = l0 + 3 >= K ? (1 << K) : (1 << (l0 + 3));
      x[k] = gl_mul(x[k], GlT::lds_get(P + gl8_table_off(K, l0), S, b * m));
    }
  }
  gl_dft8<V, first>(x, w8, w4, w83);
  if constexpr (last) {
    if (a.has_scale) {
      const uint64_t sc = GlT::from_words(a.scale);
#pragma unroll
      for (int k = 0; k < 8; k++) x[k] = gl_mul(x[k], sc);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) GlT::store_user(a.dst, (size_t)(base | (k << l0)), x[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) GlT::lds_put(lds, 1 << K, sb ^ whole_swz(k << l0), x[k]);
  }
}

template <int K, int l0, int V>
__device__ __forceinline__ void gl8_rounds(uint64_t (&x)[8], const PassArgs &a, uint32_t *lds, const uint32_t *P,
                                           const TileCtx &c, uint64_t w8, uint64_t w4, uint64_t w83) {
  gl8_round<K, l0, V>(x, a, lds, P, c, w8, w4, w83);
  if constexpr (l0 + 3 < K) {
    __syncthreads();
    gl8_rounds<K, l0 + 3, V>(x, a, lds, P, c, w8, w4, w83);
  }
}

// The persistent loop of whole_gl8_kernel for one w8 case (items claimed as
// in whole_batch_kernel).
template <int K, int V>
__device__ __forceinline__ void gl8_items(const BatchXform *__restrict__ xf, const BatchEntry *__restrict__ ent,
                                          uint32_t n, uint32_t *next, uint32_t *lds, const uint32_t *P,
                                          uint32_t &s_item, uint64_t w8, uint64_t w4, uint64_t w83) {
  if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
  __syncthreads();
  uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
  while (it < n) {
    uint32_t claimed = 0;
    if (threadIdx.x == 0) claimed = atomicAdd(next, 1u);
    const BatchXform &X = xf[ent[it].xf];
    PassArgs a;
    a.first = 1;
    a.last = 1;
    a.src = X.src;
    a.dst = X.data;
    a.src_len = X.src_len;
    a.tw = X.tw;
    a.coset = X.coset;
    a.has_scale = X.has_scale;
#pragma unroll
    for (int k = 0; k < 9; k++) a.scale[k] = X.scale[k];
    a.L = K;
    a.s0 = 1;
    a.K = K;
    a.log_g = 0;
    a.tw_one = 1;
    TileCtx c;
    c.K = K;
    c.lg = 0;
    c.G = 1;
    c.tile = 1 << K;
    c.lo_bits = 0;
    c.hi_bits = 0;
    c.hi = c.lo0 = c.c0 = 0;
    uint64_t x[8];
    gl8_rounds<K, 0, V>(x, a, lds, P, c, w8, w4, w83);
    __syncthreads();  // LDS is reused by the next item
    if (threadIdx.x == 0) s_item = claimed;
    __syncthreads();
    it = __builtin_amdgcn_readfirstlane(s_item);
  }
}

// Persistent launch over n Goldilocks transforms of 2^K points (K = 9 or 12)
// that share the table tw_all (the flat per-stage table of one root, 2^K
// entries); next[0..2) as for whole_batch_kernel.
template <int K>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void whole_gl8_kernel(
    const BatchXform *__restrict__ xf, const BatchEntry *__restrict__ ent, uint32_t n, uint32_t *next,
    const uint32_t *__restrict__ tw_all) {
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t s_item;
  uint32_t *P = lds + 2 * (1 << K);
  // P[e] = w^e, e < 2^K: stage K of the flat table holds w^j for j < 2^(K-1)
  // (at 2^(K-1) - 1), and w^(j + 2^(K-1)) = w^j * w^(2^(K-1)).  The factor
  // w^(2^(K-1)), formed as w^(2^(K-1)-1) * w from the table, is -1 for a
  // primitive root; a root of smaller order (the reference accepts any root
  // and runs the same DAG over its powers) gives 1 or another root of unity.
  constexpr int H = 1 << (K - 1);
  const uint64_t wH = gl_mul(GlT::load_dev(tw_all, (size_t)(2 * H - 2)), GlT::load_dev(tw_all, (size_t)H));
  for (int e = threadIdx.x; e < (1 << K); e += blockDim.x) {
    uint64_t v = GlT::load_dev(tw_all, (size_t)(H - 1 + (e & (H - 1))));
    if (e >= H) v = gl_mul(v, wH);
    GlT::lds_put(P, 1 << K, e, v);
  }
  __syncthreads();
  if constexpr (K >= 9) {  // compact tables of the rounds before the last
    for (int j = threadIdx.x; j < 64; j += blockDim.x)
      GlT::lds_put(P + gl8_table_off(K, 3), 64, j, GlT::lds_get(P, 1 << K, j << (K - 6)));
  }
  if constexpr (K >= 12) {
    for (int j = threadIdx.x; j < 512; j += blockDim.x)
      GlT::lds_put(P + gl8_table_off(K, 6), 512, j, GlT::lds_get(P, 1 << K, j << (K - 9)));
  }
  __syncthreads();
  const uint64_t w8 = GlT::lds_get(P, 1 << K, 1 << (K - 3));
  const uint64_t w4 = GlT::lds_get(P, 1 << K, 2 << (K - 3));
  const uint64_t w83 = GlT::lds_get(P, 1 << K, 3 << (K - 3));
  int v = 4;
  if (w8 == (1ull << 24))
    v = 0;
  else if (w8 == kGlP - (1ull << 24))
    v = 1;
  else if (w8 == kGl2p72)
    v = 2;
  else if (w8 == kGlP - kGl2p72)
    v = 3;
  switch (__builtin_amdgcn_readfirstlane(v)) {
    case 0: gl8_items<K, 0>(xf, ent, n, next, lds, P, s_item, w8, w4, w83); break;
    case 1: gl8_items<K, 1>(xf, ent, n, next, lds, P, s_item, w8, w4, w83); break;
    case 2: gl8_items<K, 2>(xf, ent, n, next, lds, P, s_item, w8, w4, w83); break;
    case 3: gl8_items<K, 3>(xf, ent, n, next, lds, P, s_item, w8, w4, w83); break;
    default: gl8_items<K, 4>(xf, ent, n, next, lds, P, s_item, w8, w4, w83); break;
  }
  if (threadIdx.x == 0) retire_workgroup(next);
}

template <int K>
static hipError_t launch_whole_gl8(const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                                   const uint32_t *tw_all, hipStream_t s) {
  const size_t lds = ((size_t)2 * (1 << K) + gl8_tables_words(K)) * 4;  // tile planes, power tables
  const unsigned threads = (1u << K) / 8;
  static std::once_flag once;
  static int grid = 0;
  static hipError_t setup = hipSuccess;
  std::call_once(once, [&] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&whole_gl8_kernel<K>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int per_cu = 0, dev = 0, cus = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, whole_gl8_kernel<K>, (int)threads, lds);
    grid = std::max(1, per_cu) * std::max(1, cus);
    setup = e;
  });
  if (setup != hipSuccess) return setup;
  const unsigned g = (unsigned)std::min<uint32_t>(n, (uint32_t)grid);
  hipLaunchKernelGGL((whole_gl8_kernel<K>), dim3(g), dim3(threads), lds, s, xf, ent, n, counter, tw_all);
  return hipGetLastError();
}

hipError_t launch_whole_gl(int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                           const uint32_t *tw_all, hipStream_t s) {
  // Sizes whose stages split into radix-8 rounds, with the launch's one table.
  if (n > 0 && tw_all && K == 12) return launch_whole_gl8<12>(xf, ent, n, counter, tw_all, s);
  if (n > 0 && tw_all && K == 9) return launch_whole_gl8<9>(xf, ent, n, counter, tw_all, s);
  return launch_whole_t<GlT>(K, xf, ent, n, counter, tw_all, s);
}

hipError_t launch_level_r8_gl(int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                              const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform,
                              uint32_t *counter, int grid, hipStream_t s) {
  return launch_level_t<GlT, 10>(K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform, counter, grid, s);
}

}  // namespace zntt
