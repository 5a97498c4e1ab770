// Persistent kernels for a batch of independent transforms of any mix of
// sizes and directions, replayed from a captured HIP graph.
//
// Every transform is planned into passes exactly as a single transform is
// (plan_passes), and every pass into tiles.  A work item is one
// (transform, pass, tile).  The batch runs as one persistent launch per pass
// level: level p holds pass p of every transform that has one, so items of a
// level are independent, and workgroups that stay resident pull them from an
// atomic counter until the level is drained.  Passes of one transform meet at
// the kernel boundary.  (Inside one launch they could not meet cheaply on
// MI355X: the eight XCDs have private L2 caches, so a device-scope
// release/acquire between workgroups compiles to an L2 write-back and an L2
// invalidate per dependency -- measured at 10x the batch's compute time.)
#include <algorithm>
#include <mutex>

#include "batch.hpp"
#include "pass_kernels.hpp"

namespace zntt {

// One round of a pass with runtime position in the pass (the single-transform
// kernel has these as template parameters).  Threads at or above `nthr` own
// no set but still take part in the barrier.  The `u` operand is normalised at
// every stage, so one body serves both stage parities.
template <class F, int R>
__device__ __forceinline__ void round_dyn(typename F::E (&x)[8], const PassArgs &a, uint32_t *lds,
                                          const TileCtx &c, int l0, bool first_round, bool last_round, int nthr) {
  using E = typename F::E;
  constexpr int SET = 1 << R, NSETS = 8 >> R;
  const bool active = (int)threadIdx.x < nthr;
  int g[NSETS], base[NSETS];
  // A transform that fits one tile (first pass, no outer index bits): its
  // bit-reversed gather would touch a different cache line per lane, so the
  // tile is read in natural order (coalesced), scaled and zero-padded, and
  // written to LDS at the bit-reversed slot; the round then reads LDS.
  const bool staged = first_round && a.first && c.hi_bits == 0;
  if (staged) {
    stage_whole_tile<F, 8, false>(a, lds, c);
    __syncthreads();
  }
  if (active) {
    sfor<NSETS>([&](auto i) {
      const int slot = threadIdx.x + i * nthr;
      g[i] = slot & (c.G - 1);
      const int rest = slot >> c.lg;
      base[i] = (rest & ((1 << l0) - 1)) | ((rest >> l0) << (l0 + R));
    });
    sfor<NSETS>([&](auto i) {
      sfor<SET>([&](auto k) {
        const int mid = base[i] | (k << l0);
        if (first_round && !staged) {
          if (a.first) {
            const size_t src = ((size_t)brev_bits(mid, c.K) << c.hi_bits) + c.c0 + g[i];
            x[i * SET + k] = src < a.src_len ? F::load_user(a.src, src) : F::zero();
          } else {
            x[i * SET + k] = F::load_dev(a.src, (c.hi << (c.lo_bits + c.K)) + ((size_t)mid << c.lo_bits) + c.lo0 + g[i]);
          }
        } else {
          x[i * SET + k] = F::lds_get(lds, c.tile, lds_swz(mid * c.G + g[i]));
        }
      });
    });
    if (first_round && a.first && a.coset && !staged) {  // coset pre-scale at the source index
      // Padding (src >= src_len) was loaded as zero; its factor is zero too.
      auto cos = [&](int e) {
        const int i = e / SET, k = e % SET;
        const int mid = base[i] | (k << l0);
        const size_t src = ((size_t)brev_bits(mid, c.K) << c.hi_bits) + c.c0 + g[i];
        return src < a.src_len ? F::load_dev(a.coset, src) : F::zero();
      };
      sfor<4>([&](auto h) {
        constexpr int e = 2 * decltype(h)::value;
        F::mul2(x[e], cos(e), x[e + 1], cos(e + 1));
      });
    }
  }
  // A pass that is a single round reads and writes the caller's buffer in
  // place: every load must land before any store.  A staged round reads LDS
  // slots that other threads' round outputs overwrite.
  if ((first_round && last_round) || staged) __syncthreads();
  if (!active) return;
  sfor<R>([&](auto t) {
    // This stage's twiddles, loaded just before use to bound register pressure.
    constexpr int T = decltype(t)::value;
    const int s = a.s0 + l0 + T;
    const size_t off = ((size_t)1 << (s - 1)) - 1;
    E tw[NSETS << T];
    sfor<NSETS>([&](auto i) {
      const size_t lo = a.first ? 0 : c.lo0 + g[i];
      const int jbase = base[i] & ((1 << l0) - 1);
      sfor<(1 << T)>([&](auto kk) {
        tw[(i << T) + kk] = F::load_dev(a.tw, off + ((size_t)(jbase | (kk << l0)) << c.lo_bits) + lo);
      });
    });
    using M = StageMap<R, T>;
    sfor<2>([&](auto h) {
      constexpr int j0 = 2 * decltype(h)::value, j1 = j0 + 1;
      F::template bfly2<true>(x[M::u(j0)], x[M::u(j0) + (1 << T)], tw[(M::set(j0) << T) + M::j(j0)],
                              x[M::u(j1)], x[M::u(j1) + (1 << T)], tw[(M::set(j1) << T) + M::j(j1)]);
    });
  });
  if (last_round && a.last && a.has_scale) {
    const E scale = F::from_words(a.scale);
    sfor<4>([&](auto h) {
      constexpr int e = 2 * decltype(h)::value;
      F::mul2(x[e], scale, x[e + 1], scale);
    });
  }
  sfor<NSETS>([&](auto i) {
    sfor<SET>([&](auto k) {
      const int mid = base[i] | (k << l0);
      if (last_round) {
        size_t pos;
        if (a.first)
          pos = ((size_t)brev_bits((uint32_t)(c.c0 + g[i]), c.hi_bits) << c.K) + mid;
        else
          pos = (c.hi << (c.lo_bits + c.K)) + ((size_t)mid << c.lo_bits) + c.lo0 + g[i];
        if (a.last)
          F::store_user(a.dst, pos, x[i * SET + k]);
        else
          F::store_dev(a.dst, pos, x[i * SET + k]);
      } else {
        F::lds_put(lds, c.tile, lds_swz(mid * c.G + g[i]), x[i * SET + k]);
      }
    });
  });
}

template <class F, int BLOCK, int WAVES_PER_SIMD>
__global__ __launch_bounds__(BLOCK, WAVES_PER_SIMD) void batch_kernel(const BatchXform *xf, const BatchEntry *ent,
                                                                      uint32_t nent, uint32_t nitems,
                                                                      uint32_t uniform_tiles, uint32_t *next) {
  using E = typename F::E;
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t s_item, s_pos[2];
  if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
  __syncthreads();
  uint32_t it = s_item;
  for (;;) {
    if (it >= nitems) {
      if (threadIdx.x == 0) retire_workgroup(next);
      return;
    }
    // Claim the following item now, so that the atomic's round trip overlaps
    // this item's work; it is published through s_item at the end.
    uint32_t claimed = 0;
    if (threadIdx.x == 0) claimed = atomicAdd(next, 1u);
    const uint32_t lo = find_entry(ent, nent, it, uniform_tiles, s_pos);
    const BatchEntry e = ent[lo];
    const BatchXform &X = xf[e.xf];
    const int p = (int)e.pass;
    const uint32_t tile_id = it - e.first;
    PassArgs a;
    a.first = p == 0;
    a.last = p == X.npass - 1;
    a.src = a.first ? X.src : X.scratch;
    a.dst = a.last ? X.data : X.scratch;
    a.src_len = X.src_len;
    a.tw_one = 1;  // batches use tables of powers
    a.tw = X.tw;
    a.coset = a.first ? X.coset : nullptr;
    a.has_scale = a.last ? X.has_scale : 0;
#pragma unroll
    for (int k = 0; k < 9; k++) a.scale[k] = X.scale[k];
    a.L = X.L;
    a.s0 = X.s0[p];
    a.K = X.K[p];
    a.log_g = X.lg[p];
    TileCtx c;
    c.K = a.K;
    c.lg = a.log_g;
    c.G = 1 << c.lg;
    c.tile = c.G << c.K;
    c.lo_bits = a.s0 - 1;
    c.hi_bits = a.L - c.lo_bits - c.K;
    c.hi = c.lo0 = c.c0 = 0;
    if (a.first) {
      c.c0 = (size_t)tile_id << c.lg;
    } else {
      const int lbb = c.lo_bits - c.lg;
      c.hi = (size_t)tile_id >> lbb;
      c.lo0 = ((size_t)tile_id & (((size_t)1 << lbb) - 1)) << c.lg;
    }
    const int nthr = c.tile >> 3;
    const int nr = (c.K + 2) / 3;
    const int r0 = c.K - 3 * (nr - 1);
    E x[8];
    int l0 = 0;
    for (int ri = 0; ri < nr; ri++) {
      const int r = ri == 0 ? r0 : 3;
      const bool fr = ri == 0, lr = ri == nr - 1;
      if (r == 1)
        round_dyn<F, 1>(x, a, lds, c, l0, fr, lr, nthr);
      else if (r == 2)
        round_dyn<F, 2>(x, a, lds, c, l0, fr, lr, nthr);
      else
        round_dyn<F, 3>(x, a, lds, c, l0, fr, lr, nthr);
      l0 += r;
      if (!lr) __syncthreads();
    }
    __syncthreads();  // LDS and s_pos are reused by the next item
    if (threadIdx.x == 0) s_item = claimed;
    __syncthreads();
    it = s_item;
  }
}

hipError_t launch_batch_level(int field, const BatchXform *xf, const BatchEntry *ent, uint32_t nent,
                              uint32_t nitems, uint32_t uniform_tiles, uint32_t *counter, int grid, int tile_log,
                              hipStream_t s) {
  if (nitems == 0) return hipSuccess;
  if (uniform_tiles && (uint64_t)uniform_tiles * nent != nitems) return hipErrorInvalidValue;
  const unsigned threads = 1u << (tile_log - 3);
  const int g = (int)(nitems < (uint32_t)grid ? nitems : (uint32_t)grid);
  if (field == kFr) {
    if (tile_log > 11) return hipErrorInvalidValue;
    const size_t lds = (size_t)9 * 4 << tile_log;
    hipLaunchKernelGGL((batch_kernel<FrT, 256, 2>), dim3(g), dim3(threads), lds, s, xf, ent, nent, nitems, uniform_tiles,
                       counter);
  } else if (field == kGl) {
    if (tile_log > 12) return hipErrorInvalidValue;
    const size_t lds = (size_t)2 * 4 << tile_log;
    hipLaunchKernelGGL((batch_kernel<GlT, 512, 4>), dim3(g), dim3(threads), lds, s, xf, ent, nent, nitems, uniform_tiles,
                       counter);
  } else {
    if (tile_log > 12) return hipErrorInvalidValue;
    const size_t lds = (size_t)4 << tile_log;
    if (field == kBb)
      hipLaunchKernelGGL((batch_kernel<BbT, 512, 4>), dim3(g), dim3(threads), lds, s, xf, ent, nent, nitems,
                         uniform_tiles, counter);
    else
      hipLaunchKernelGGL((batch_kernel<KbT, 512, 4>), dim3(g), dim3(threads), lds, s, xf, ent, nent, nitems,
                         uniform_tiles, counter);
  }
  return hipGetLastError();
}

hipError_t launch_batch_level_r8(int field, int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                                 const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform_tiles,
                                 uint32_t *counter, int grid, hipStream_t s) {
  if (!level_r8_ok(field, K, tile_log)) return hipErrorInvalidValue;
  if (uniform_tiles && (uint64_t)uniform_tiles * nent != nitems) return hipErrorInvalidValue;
  if (field == kFr)
    return launch_level_r8_fr(K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform_tiles, counter, grid, s);
  if (field == kGl)
    return launch_level_r8_gl(K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform_tiles, counter, grid, s);
  return launch_level_r8_f31(field, K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform_tiles, counter, grid,
                             s);
}

hipError_t configure_batch_lds() {
  static std::once_flag once;
  static hipError_t result = hipSuccess;
  std::call_once(once, [] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<FrT, 256, 2>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 9 * 4 * (1 << 11));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<GlT, 512, 4>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 2 * 4 * (1 << 12));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<BbT, 512, 4>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 4 * (1 << 12));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<KbT, 512, 4>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 4 * (1 << 12));
    result = e;
  });
  return result;
}

// ---------------------------------------------------------------------------
// Counter-based input stream (synthetic request coefficients).  Element i of
// stream (seed, id) draws words mix64(key + (ctr + 1) * gamma) with
// key = mix64(mix64(seed) ^ id) and ctr = i << 8 | attempt << 2 | word, where
// mix64 is the SplitMix64 finaliser and gamma its increment.  Fr takes four
// words (little-endian), clears the top two bits and accepts a value below r
// (the first of up to 64 attempts; the value is stored as the Montgomery
// words, so the represented element is uniform too); the small fields take
// one word, masked to the modulus' bit length (64 for Goldilocks, 31 for
// BabyBear and KoalaBear), and accept it below p.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

constexpr uint64_t kGamma = 0x9e3779b97f4a7c15ull;

__global__ void fill_counter_fr_kernel(uint32_t *out, size_t n, uint64_t key) {
  const uint64_t r[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                         0x30644e72e131a029ull};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint64_t w[4] = {0, 0, 0, 0};
    for (uint64_t att = 0; att < 64; att++) {
      const uint64_t ctr = (uint64_t)i << 8 | att << 2;
#pragma unroll
      for (int k = 0; k < 4; k++) w[k] = mix64(key + (ctr + k + 1) * kGamma);
      w[3] &= 0x3fffffffffffffffull;
      bool below = false, decided = false;
#pragma unroll
      for (int k = 3; k >= 0; k--) {
        if (!decided && w[k] != r[k]) {
          below = w[k] < r[k];
          decided = true;
        }
      }
      if (below) break;
    }
    uint4 *p = reinterpret_cast<uint4 *>(out + i * 8);
    p[0] = make_uint4((uint32_t)w[0], (uint32_t)(w[0] >> 32), (uint32_t)w[1], (uint32_t)(w[1] >> 32));
    p[1] = make_uint4((uint32_t)w[2], (uint32_t)(w[2] >> 32), (uint32_t)w[3], (uint32_t)(w[3] >> 32));
  }
}

// Small fields: one word, masked to the modulus' bit length, accepted below p.
__global__ void fill_counter_small_kernel(uint32_t *out, size_t n, uint64_t key, uint64_t p, uint64_t mask) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint64_t w = 0;
    for (uint64_t att = 0; att < 64; att++) {
      w = mix64(key + (((uint64_t)i << 8 | att << 2) + 1) * kGamma) & mask;
      if (w < p) break;
    }
    reinterpret_cast<uint2 *>(out)[i] = make_uint2((uint32_t)w, (uint32_t)(w >> 32));
  }
}

uint64_t counter_key(uint64_t seed, uint64_t stream_id) {
  auto m = [](uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  return m(m(seed) ^ stream_id);
}

hipError_t launch_fill_counter(int field, uint32_t *out, size_t n, uint64_t seed, uint64_t stream_id,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t key = counter_key(seed, stream_id);
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  if (field == kFr) {
    hipLaunchKernelGGL(fill_counter_fr_kernel, dim3((unsigned)blocks), dim3(256), 0, s, out, n, key);
  } else {
    const uint64_t p = field == kGl ? 0xffffffff00000001ull : (field == kBb ? kBbP : kKbP);
    const uint64_t mask = field == kGl ? ~0ull : 0x7fffffffull;
    hipLaunchKernelGGL(fill_counter_small_kernel, dim3((unsigned)blocks), dim3(256), 0, s, out, n, key, p, mask);
  }
  return hipGetLastError();
}

}  // namespace zntt
