// The stage-range pass kernels (templates), instantiated per field in
// pass_fr.hip and pass_gl.hip so the two compile in parallel.
//
// A pass covers DIT stages s0..s0+K-1 of a 2^L-point transform; see
// ntt_kernels.hip for the overall schedule.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <utility>

#include "batch.hpp"
#include "field_traits.hpp"
#include "kernels.hpp"

namespace zntt {

// ---------------------------------------------------------------------------
// The stage-range pass.
// ---------------------------------------------------------------------------
template <class F>
__global__ __launch_bounds__(256) void ntt_pass_kernel(const PassArgs a) {
  using E = typename F::E;
  extern __shared__ uint32_t lds[];
  const int K = a.K, lg = a.log_g, G = 1 << lg;
  const int tile = G << K;
  const int lo_bits = a.s0 - 1;
  const int hi_bits = a.L - lo_bits - K;
  const size_t w = blockIdx.x;
  size_t hi = 0, lo0 = 0, c0 = 0;
  if (a.first) {
    c0 = w << lg;
  } else {
    const int lbb = lo_bits - lg;
    hi = w >> lbb;
    lo0 = (w & (((size_t)1 << lbb) - 1)) << lg;
  }

  // Gather the tile: element (g, mid) of the tile sits at LDS index mid*G + g.
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    const int g = e & (G - 1), mid = e >> lg;
    E x;
    if (a.first) {
      const size_t src = ((size_t)brev_bits(mid, K) << hi_bits) + c0 + g;
      x = F::load_user(a.src, src);
      if (a.coset) x = F::mul(x, F::load_dev(a.coset, src));
    } else {
      const size_t pos = (hi << (lo_bits + K)) + ((size_t)mid << lo_bits) + lo0 + g;
      x = F::load_dev(a.src, pos);
    }
    F::lds_put(lds, tile, e, x);
  }
  __syncthreads();

  for (int l = 0; l < K; l++) {
    const int s = a.s0 + l;
    const int half = 1 << l;
    const size_t stage_off = ((size_t)1 << (s - 1)) - 1;
    for (int b = threadIdx.x; b < (tile >> 1); b += blockDim.x) {
      const int g = b & (G - 1), q = b >> lg;
      const int jl = q & (half - 1), blk = q >> l;
      const int mid0 = (blk << (l + 1)) | jl;
      const size_t lo = a.first ? 0 : lo0 + g;
      const E wv = F::load_dev(a.tw, stage_off + ((size_t)jl << lo_bits) + lo);
      const int i0 = mid0 * G + g, i1 = i0 + (half << lg);
      E u = F::lds_get(lds, tile, i0), v = F::lds_get(lds, tile, i1);
      F::template bfly<true>(u, v, wv);
      F::lds_put(lds, tile, i0, u);
      F::lds_put(lds, tile, i1, v);
    }
    __syncthreads();
  }

  E scale;
  if (a.last && a.has_scale) scale = F::from_words(a.scale);
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int g, mid;
    size_t pos;
    if (a.first) {
      g = e >> K;
      mid = e & ((1 << K) - 1);
      pos = ((size_t)brev_bits((uint32_t)(c0 + g), hi_bits) << K) + mid;
    } else {
      g = e & (G - 1);
      mid = e >> lg;
      pos = (hi << (lo_bits + K)) + ((size_t)mid << lo_bits) + lo0 + g;
    }
    E x = F::lds_get(lds, tile, mid * G + g);
    if (a.last) {
      if (a.has_scale) x = F::mul(x, scale);
      F::store_user(a.dst, pos, x);
    } else {
      F::store_dev(a.dst, pos, x);
    }
  }
}

// ---------------------------------------------------------------------------
// The stage-range pass with register rounds (tiles of >= 512 elements).
//
// Same tile and DAG as ntt_pass_kernel, but each thread keeps 8 elements in
// registers and runs up to three stages on them between LDS exchanges:
// a "set" is 2^R elements whose tile positions differ only in mid bits
// [l0, l0+R), which is exactly what stages l0..l0+R-1 couple.  The first
// round reads HBM directly and the last round writes HBM directly, so a pass
// of K stages costs ceil(K/3) - 1 LDS exchanges and barriers instead of K.
// ---------------------------------------------------------------------------

// XOR bits [3,5) of a tile index with bits [6,8): the first round's lanes run
// over 8 consecutive groups and then jump by whole 64-word rows; the swizzle
// sends those runs to distinct LDS banks.  A bijection on any tile >= 256.
__device__ __forceinline__ int lds_swz(int i) { return i ^ (((i >> 6) & 3) << 3); }

// The whole-transform kernel's tile (one transform, G = 1) uses another
// swizzle: bit 5 + j of a tile index, j < 7, XORs column kWholeSwzCol[j] into
// bits 0..5.  Bits 6 and up are kept, bit 5 changes by a function of them
// and bits 0..4 by a function of bits 5 and up, so the map is a bijection on
// every 2^K tile.  The columns were searched so that, for K = 9..12 and four
// or eight elements a thread, the 64 lanes of every staging write and of
// every round's reads and write-backs hit 64 distinct LDS banks (under 32
// banks, each half-wave hits 32, except 2^10 tiles at 1.2-way on average);
// the shift-and-mask swizzle above reaches 4-5-way conflicts on these
// patterns.  The map is linear over GF(2): for indices with disjoint bits,
// whole_swz(a | b) = whole_swz(a) ^ whole_swz(b), so a set's positions are
// its base's image XOR compile-time constants.
constexpr int kWholeSwzCol[7] = {31, 42, 49, 35, 19, 4, 28};
__host__ __device__ constexpr int whole_swz(int i) {
  int m = 0;
  for (int j = 0; j < 7; j++) m ^= (-((i >> (5 + j)) & 1)) & kWholeSwzCol[j];
  return i ^ m;
}

struct TileCtx {
  int K, lg, G, tile, lo_bits, hi_bits;
  size_t hi, lo0, c0;
};

// Compile-time loop: f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>).
// Every index into the register arrays below is a constant expression, so the
// arrays are scalarised into VGPRs however large the unrolled body gets.
template <class Fn, int... Is>
__device__ __forceinline__ void sfor_impl(Fn &&f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class Fn>
__device__ __forceinline__ void sfor(Fn &&f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// ODD: parity of the pass's first stage s0, so that the normalise-at-odd-stage
// rule folds to a constant and each round is one straight-line block.
// Whole-tile staging of a first pass with no outer index bits (the transform
// fits one tile): element e of the transform (natural order, coalesced) goes
// to LDS slot brev(e), multiplied by coset[e] when the pass has
// a coset, zero at and past src_len.  Uses every thread of the workgroup.
// UNROLL: when the workgroup holds exactly N elements a thread and the
// source covers the tile, every thread issues its N loads before the first
// LDS write, so the loads of a tile are in flight together rather than one
// after another.  It costs registers, so kernels whose occupancy depends on a
// few VGPRs keep the loop.  WHOLE: the tile uses whole_swz (the
// whole-transform kernel, where blockDim.x = 2^K / N).
template <class F, int N, bool UNROLL, bool WHOLE = false>
__device__ __forceinline__ void stage_whole_tile(const PassArgs &a, uint32_t *lds, const TileCtx &c) {
  static_assert(N % 2 == 0, "pairs of coset products");
  const int n = c.tile;
  if (UNROLL && (int)blockDim.x * N == n && a.src_len >= (uint64_t)n) {
    typename F::E v[N];
    sfor<N>([&](auto i) { v[i] = F::load_user(a.src, threadIdx.x + i * blockDim.x); });
    if (a.coset) {
      sfor<N / 2>([&](auto h) {
        constexpr int e = 2 * decltype(h)::value;
        F::mul2(v[e], F::load_dev(a.coset, threadIdx.x + e * blockDim.x), v[e + 1],
                F::load_dev(a.coset, threadIdx.x + (e + 1) * blockDim.x));
      });
    }
    if constexpr (WHOLE) {
      // brev(t + i * T) = brev(t) | brev(i * T) for t < T = n / N.
      const int ts = whole_swz(brev_bits(threadIdx.x, c.K));
      sfor<N>([&](auto i) {
        F::lds_put(lds, c.tile, ts ^ whole_swz(brev_bits(decltype(i)::value * (n / N), c.K)), v[i]);
      });
    } else {
      sfor<N>([&](auto i) {
        F::lds_put(lds, c.tile, lds_swz(brev_bits(threadIdx.x + i * blockDim.x, c.K)), v[i]);
      });
    }
    return;
  }
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    typename F::E v = F::zero();
    if ((uint64_t)e < a.src_len) {
      v = F::load_user(a.src, e);
      if (a.coset) v = F::mul(v, F::load_dev(a.coset, e));
    }
    const int r = brev_bits(e, c.K);
    F::lds_put(lds, c.tile, WHOLE ? whole_swz(r) : lds_swz(r), v);
  }
}

// Elements a thread holds in the register rounds of a field's large passes:
// eight, except BN254 Fr, whose nine-limb values make eight a thread cost
// about 240 VGPRs (two waves a SIMD); four halve that, and a 2^11 tile then
// runs as 512 threads, so LDS and registers both allow four waves a SIMD.
template <class F>
constexpr int round_elems() {
  return std::is_same<F, FrT>::value ? 4 : 8;
}

// The butterflies of radix-2 stage T inside a round of R stages over a
// thread's N elements (N >> R sets of 2^R): butterfly j couples elements
// u(j) and u(j) + 2^T of set set(j), with twiddle index j(j) within the stage.
template <int R, int T>
struct StageMap {
  static constexpr int SET = 1 << R, PER_SET = SET / 2;
  static constexpr int set(int j) { return j / PER_SET; }
  static constexpr int kk(int j) {
    const int q = j % PER_SET;
    return ((q >> T) << (T + 1)) | (q & ((1 << T) - 1));
  }
  static constexpr int u(int j) { return set(j) * SET + kk(j); }
  static constexpr int j(int jj) { return kk(jj) & ((1 << T) - 1); }
};

// Zero padding of a first pass: z when the source holds 2^(L - z) values
// (src_len a power of two below 2^L) and the other inputs are zero; else 0.
// Padded kernels test for it in the first kPadStages stages only (blowups up
// to 8; a deeper padding runs its further stages as ordinary butterflies):
// the test in every round costs the register-round kernels registers.
constexpr int kPadStages = 3;
__device__ __forceinline__ int pad_log(const PassArgs &a) {
  const uint64_t len = a.src_len;
  if (!a.first || len == 0 || (len & (len - 1)) != 0 || len >= ((uint64_t)1 << a.L)) return 0;
  return a.L - (63 - __builtin_clzll(len));
}

// TWL: the pass's twiddle table is resident in LDS after the tile's planes
// (a whole transform: 2^K entries, plane stride 2^K), not read from HBM.
// WHOLE: the round runs in the whole-transform batch kernel.
// LB: lane bits.  The buffers hold 2^LB transforms of the same size
// interleaved, element e of transform `lane` at index e * 2^LB + lane, and
// the low LB bits of a tile column select the lane: the tile holds the same
// 2^lg / 2^LB columns of every lane, so consecutive threads touch consecutive
// words.  Positions and twiddle indices use the column without the lane bits
// (c.c0, c.lo0 count such columns); LDS slots use the full column.
// PAD: the first pass reads zero at and past a.src_len, from the source and
// from the coset table (a first pass that zero-pads its input).  ZSKIP (with
// PAD): when the source is 2^(L - z) values (pad_log), the inputs at
// bit-reversed positions whose low z bits are not all zero are the padding,
// so every butterfly of stages 0 .. z-1 has v = 0 and becomes (u, u), and
// coset products of padding are skipped where a whole wave holds only
// padding.
template <class F, int R, int l0, bool first_round, bool last_round, bool ODD, bool TWL = false, int N = 8,
          bool WHOLE = false, int LB = 0, bool PAD = false, bool ZSKIP = false>
__device__ __forceinline__ void pass_round(typename F::E (&x)[N], const PassArgs &a, uint32_t *lds,
                                           const TileCtx &c) {
  using E = typename F::E;
  constexpr int SET = 1 << R, NSETS = N >> R;
  int zlog = 0;  // zero-padding stages of this round's transform (first pass only)
  if constexpr (PAD && ZSKIP) zlog = pad_log(a);
  int g[NSETS], base[NSETS], gr[NSETS], ln[NSETS];
  sfor<NSETS>([&](auto i) {
    const int slot = threadIdx.x + i * blockDim.x;
    g[i] = slot & (c.G - 1);
    gr[i] = g[i] >> LB;
    ln[i] = g[i] & ((1 << LB) - 1);
    const int rest = slot >> c.lg;
    base[i] = (rest & ((1 << l0) - 1)) | ((rest >> l0) << (l0 + R));
  });
  // Buffer index of transform element `e` of set i's lane.
  auto at = [&](auto i, size_t e) { return (e << LB) + ln[i]; };
  // LDS slot of element k of set i.
  int sb[NSETS];
  if constexpr (WHOLE) sfor<NSETS>([&](auto i) { sb[i] = whole_swz(base[i]); });
  auto slot_of = [&](auto i, auto k) {
    if constexpr (WHOLE)
      return sb[i] ^ whole_swz(decltype(k)::value << l0);
    else
      return lds_swz((base[i] | (decltype(k)::value << l0)) * c.G + g[i]);
  };
  // Twiddles: stage l0+t couples (k, k + 2^t) with index jl = mid mod 2^(l0+t),
  // so a set needs 2^t of them at stage t; issue all loads before computing.
  E tw[NSETS * (SET - 1)];
  sfor<R>([&](auto t) {
    const int s = a.s0 + l0 + t;
    const size_t off = ((size_t)1 << (s - 1)) - 1;
    sfor<NSETS>([&](auto i) {
      const size_t lo = a.first ? 0 : c.lo0 + gr[i];
      const int jbase = base[i] & ((1 << l0) - 1);
      sfor<(1 << t)>([&](auto kk) {
        const size_t ti = off + ((size_t)(jbase | (kk << l0)) << c.lo_bits) + lo;
        if constexpr (TWL)
          tw[i * (SET - 1) + (1 << t) - 1 + kk] = F::lds_get(lds + F::W * c.tile, c.tile, (int)ti);
        else
          tw[i * (SET - 1) + (1 << t) - 1 + kk] = F::load_dev(a.tw, ti);
      });
    });
  });
  // Load the sets.  A transform that fits the tile is staged through LDS
  // first (coalesced reads instead of a bit-reversed gather).
  bool staged = false;
  if constexpr (first_round) {
    staged = LB == 0 && a.first && c.hi_bits == 0;
    if (staged) {
      stage_whole_tile<F, N, WHOLE || !std::is_same<F, FrT>::value, WHOLE>(a, lds, c);
      __syncthreads();
    }
  }
  sfor<NSETS>([&](auto i) {
    sfor<SET>([&](auto k) {
      const int mid = base[i] | (k << l0);
      if constexpr (first_round) {
        if (staged) {
          x[i * SET + k] = F::lds_get(lds, c.tile, slot_of(i, k));
        } else if (a.first) {
          const size_t src = ((size_t)brev_bits(mid, c.K) << c.hi_bits) + c.c0 + gr[i];
          if constexpr (PAD)
            x[i * SET + k] = src < a.src_len ? F::load_user(a.src, at(i, src)) : F::zero();
          else
            x[i * SET + k] = F::load_user(a.src, at(i, src));
        } else {
          const size_t pos = (c.hi << (c.lo_bits + c.K)) + ((size_t)mid << c.lo_bits) + c.lo0 + gr[i];
          x[i * SET + k] = F::load_dev(a.src, at(i, pos));
        }
      } else {
        x[i * SET + k] = F::lds_get(lds, c.tile, slot_of(i, k));
      }
    });
  });
  if constexpr (first_round) {
    if (staged) __syncthreads();  // every staged slot is read before round outputs overwrite LDS
    if (a.first && a.coset && !staged) {  // coset pre-scale: data[i] *= shift^i at the source index
      auto cos = [&](int e) {
        const int i = e / SET, k = e % SET;
        const int mid = base[i] | (k << l0);
        const size_t src = ((size_t)brev_bits(mid, c.K) << c.hi_bits) + c.c0 + gr[i];
        if constexpr (PAD)
          return src < a.src_len ? F::load_dev(a.coset, src) : F::zero();
        else
          return F::load_dev(a.coset, src);
      };
      sfor<N / 2>([&](auto h) {
        constexpr int e = 2 * decltype(h)::value;
        if constexpr (PAD && ZSKIP) {
          // Skip the pair's products when both elements are padding in
          // every lane of the wave (their values are zero).
          auto live = [&](int f) {
            const int i = f / SET, k = f % SET;
            const int mid = base[i] | (k << l0);
            return ((size_t)brev_bits(mid, c.K) << c.hi_bits) + c.c0 + gr[i] < a.src_len;
          };
          if (zlog > 0 && !__any(live(e) || live(e + 1))) return;
        }
        F::mul2(x[e], cos(e), x[e + 1], cos(e + 1));
      });
    }
  }
  // R radix-2 stages in registers; carry-normalise u at odd stages.  The
  // N/2 butterflies of a stage go out as interleaved pairs.
  sfor<R>([&](auto t) {
    constexpr int T = decltype(t)::value;
    constexpr bool norm = (((ODD ? 1 : 0) + l0 + T) & 1) != 0;
    using M = StageMap<R, T>;
    if constexpr (PAD && ZSKIP && l0 + T < kPadStages) {
      if (a.first && l0 + T < zlog) {  // v = 0 in every butterfly of this stage
        sfor<N / 2>([&](auto j) {
          constexpr int jj = decltype(j)::value;
          F::template bfly_zero<norm>(x[M::u(jj)], x[M::u(jj) + (1 << T)]);
        });
        return;
      }
    }
    sfor<N / 4>([&](auto h) {
      constexpr int j0 = 2 * decltype(h)::value, j1 = j0 + 1;
      auto both = [&] {
        F::template bfly2<norm>(x[M::u(j0)], x[M::u(j0) + (1 << T)],
                                tw[M::set(j0) * (SET - 1) + (1 << T) - 1 + M::j(j0)], x[M::u(j1)],
                                x[M::u(j1) + (1 << T)], tw[M::set(j1) * (SET - 1) + (1 << T) - 1 + M::j(j1)]);
      };
      // In the first round of a transform's first pass the twiddle index of
      // stage T is M::j, and index 0 is the twiddle 1: no product needed.
      constexpr bool one0 = first_round && M::j(j0) == 0 && F::one_ok(T);
      constexpr bool one1 = first_round && M::j(j1) == 0 && F::one_ok(T);
      if constexpr (one0 || one1) {
        if (a.first && a.tw_one) {
          if constexpr (one0)
            F::template bfly_one<norm, (T > 0)>(x[M::u(j0)], x[M::u(j0) + (1 << T)]);
          else
            F::template bfly<norm>(x[M::u(j0)], x[M::u(j0) + (1 << T)],
                                   tw[M::set(j0) * (SET - 1) + (1 << T) - 1 + M::j(j0)]);
          if constexpr (one1)
            F::template bfly_one<norm, (T > 0)>(x[M::u(j1)], x[M::u(j1) + (1 << T)]);
          else
            F::template bfly<norm>(x[M::u(j1)], x[M::u(j1) + (1 << T)],
                                   tw[M::set(j1) * (SET - 1) + (1 << T) - 1 + M::j(j1)]);
        } else {
          both();
        }
      } else {
        both();
      }
    });
  });
  // Store the sets (to the positions they came from, or to HBM).
  if constexpr (last_round) {
    if (a.last && a.has_scale) {
      const E scale = F::from_words(a.scale);
      sfor<N / 2>([&](auto h) {
        constexpr int e = 2 * decltype(h)::value;
        F::mul2(x[e], scale, x[e + 1], scale);
      });
    }
  }
  sfor<NSETS>([&](auto i) {
    sfor<SET>([&](auto k) {
      const int mid = base[i] | (k << l0);
      if constexpr (last_round) {
        size_t pos;
        if (a.first)
          pos = ((size_t)brev_bits((uint32_t)(c.c0 + gr[i]), c.hi_bits) << c.K) + mid;
        else
          pos = (c.hi << (c.lo_bits + c.K)) + ((size_t)mid << c.lo_bits) + c.lo0 + gr[i];
        if (a.last)
          F::store_user(a.dst, at(i, pos), x[i * SET + k]);
        else
          F::store_dev(a.dst, at(i, pos), x[i * SET + k]);
      } else {
        F::lds_put(lds, c.tile, slot_of(i, k), x[i * SET + k]);
      }
    });
  });
}

// Rounds of a K-stage pass over N = 2^LN elements a thread: the first takes
// K - LN*(ceil(K/LN) - 1) stages, the rest LN each; unrolled at compile time
// so `x` stays in registers.
template <class F, int K, int RI, bool ODD, bool TWL = false, int N = 8, bool WHOLE = false, int LB = 0,
          bool PAD = false, bool ZSKIP = false>
__device__ __forceinline__ void pass_rounds(typename F::E (&x)[N], const PassArgs &a, uint32_t *lds,
                                            const TileCtx &c) {
  constexpr int LN = N == 8 ? 3 : 2;
  constexpr int NR = (K + LN - 1) / LN;
  constexpr int R0 = K - LN * (NR - 1);
  constexpr int R = RI == 0 ? R0 : LN;
  constexpr int L0 = RI == 0 ? 0 : R0 + LN * (RI - 1);
  pass_round<F, R, L0, RI == 0, RI == NR - 1, ODD, TWL, N, WHOLE, LB, PAD, ZSKIP>(x, a, lds, c);
  if constexpr (RI + 1 < NR) {
    __syncthreads();
    pass_rounds<F, K, RI + 1, ODD, TWL, N, WHOLE, LB, PAD, ZSKIP>(x, a, lds, c);
  }
}

// Tile of workgroup b in a grid of g.  Workgroups go to the eight XCDs in
// turn (b to XCD b mod 8), and each XCD has its own L2; with this order the
// workgroups one XCD runs at a time take consecutive tiles, whose rows of
// 36-byte elements share cache lines at their edges.  A bijection on [0, g).
__device__ __forceinline__ size_t xcd_tile(uint32_t b, uint32_t g) {
  return (g & 7) ? b : (size_t)(b & 7) * (g >> 3) + (b >> 3);
}

// LB > 0: 2^LB interleaved transforms (see pass_round); a.log_g counts the
// lane bits, a.L and a.s0 are those of one transform.  PAD: a first pass
// reads zero at and past a.src_len.
template <class F, int K, bool ODD, int N = 8, int LB = 0, bool PAD = false>
__global__ __launch_bounds__(512) void ntt_pass_r8_kernel(const PassArgs a) {
  using E = typename F::E;
  extern __shared__ uint32_t lds[];
  TileCtx c;
  c.K = K;
  c.lg = a.log_g;
  c.G = 1 << c.lg;
  c.tile = c.G << c.K;
  c.lo_bits = a.s0 - 1;
  c.hi_bits = a.L - c.lo_bits - c.K;
  const int lgr = c.lg - LB;  // columns of one transform in the tile
  const size_t w = xcd_tile(blockIdx.x, gridDim.x);
  c.hi = c.lo0 = c.c0 = 0;
  if (a.first) {
    c.c0 = w << lgr;
  } else {
    const int lbb = c.lo_bits - lgr;
    c.hi = w >> lbb;
    c.lo0 = (w & (((size_t)1 << lbb) - 1)) << lgr;
  }
  E x[N];
  pass_rounds<F, K, 0, ODD, false, N, false, LB, PAD>(x, a, lds, c);
}

// Persistent launches claim items from ctr[0] and count finished workgroups
// in ctr[1].  Thread 0 of each workgroup calls this once it claims no more:
// its last claim has returned, so when the count reaches the grid size every
// claim of the launch is done and the last workgroup returns both counters to
// zero for the next launch on the stream.
__device__ __forceinline__ void retire_workgroup(uint32_t *ctr) {
  if (atomicAdd(ctr + 1, 1u) == gridDim.x - 1) {
    atomicExch(ctr, 0u);
    atomicExch(ctr + 1, 0u);
  }
}

// The entry of a persistent launch's work list (entries ordered by `first`,
// strictly increasing from 0) that holds item `it`; every thread of the
// workgroup calls it and gets the same index.  Equal tile counts: a
// division.  Otherwise each step has every thread test one candidate
// boundary, narrowing [lo, hi) by a factor of the workgroup size; two LDS
// slots alternate so that a write never races a read of the step before.
__device__ __forceinline__ uint32_t find_entry(const BatchEntry *ent, uint32_t nent, uint32_t it,
                                               uint32_t uniform_tiles, uint32_t *s_pos) {
  const uint32_t nthreads = blockDim.x;
  uint32_t lo = 0, hi = nent;
  int ping = 0;
  if (uniform_tiles) {
    lo = it / uniform_tiles;
    hi = lo + 1;
  }
  while (hi - lo > 1) {
    const uint32_t stride = (hi - lo + nthreads - 1) / nthreads;
    const uint32_t t = lo + threadIdx.x * stride;
    if (t < hi) {
      const uint32_t nx = t + stride;
      if (ent[t].first <= it && (nx >= hi || ent[nx].first > it)) s_pos[ping] = t;
    }
    __syncthreads();
    lo = s_pos[ping];
    hi = min(lo + stride, hi);
    ping ^= 1;
  }
  return lo;
}

// Persistent kernel over the tiles of one pass level of a batch whose passes
// all have K stages, first stages of parity ODD and tiles of blockDim.x * N
// elements (N = round_elems<F>()): the single-transform pass kernel's
// register rounds, with compile-time round positions, run per claimed tile.
// Items are claimed from next[0] as in the other persistent launches
// (next[0..2) zero at launch and again when it ends).  A first pass reads
// zero at and past its transform's src_len, from the source and from the
// coset table; with ZSKIP it also skips the products of that padding
// (pass_round).
template <class F, int K, bool ODD, bool ZSKIP>
__device__ __forceinline__ void level_r8_body(const BatchXform *__restrict__ xf, const BatchEntry *__restrict__ ent,
                                              uint32_t nent, uint32_t nitems, uint32_t uniform_tiles,
                                              uint32_t *next) {
  using E = typename F::E;
  constexpr int N = round_elems<F>();
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t s_item, s_pos[2];
  if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
  __syncthreads();
  uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
  while (it < nitems) {
    uint32_t claimed = 0;
    if (threadIdx.x == 0) claimed = atomicAdd(next, 1u);
    const uint32_t ei = __builtin_amdgcn_readfirstlane(find_entry(ent, nent, it, uniform_tiles, s_pos));
    const BatchEntry e = ent[ei];
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
    a.K = K;
    a.log_g = X.lg[p];
    TileCtx c;
    c.K = K;
    c.lg = a.log_g;
    c.G = 1 << c.lg;
    c.tile = c.G << K;
    c.lo_bits = a.s0 - 1;
    c.hi_bits = a.L - c.lo_bits - K;
    c.hi = c.lo0 = c.c0 = 0;
    if (a.first) {
      c.c0 = (size_t)tile_id << c.lg;
    } else {
      const int lbb = c.lo_bits - c.lg;
      c.hi = (size_t)tile_id >> lbb;
      c.lo0 = ((size_t)tile_id & (((size_t)1 << lbb) - 1)) << c.lg;
    }
    E x[N];
    pass_rounds<F, K, 0, ODD, false, N, false, 0, true, ZSKIP>(x, a, lds, c);
    __syncthreads();  // LDS and s_pos are reused by the next item
    if (threadIdx.x == 0) s_item = claimed;
    __syncthreads();
    it = __builtin_amdgcn_readfirstlane(s_item);
  }
  if (threadIdx.x == 0) retire_workgroup(next);
}

template <class F, int K, bool ODD>
__global__ __launch_bounds__(512) void level_r8_kernel(const BatchXform *__restrict__ xf,
                                                       const BatchEntry *__restrict__ ent, uint32_t nent,
                                                       uint32_t nitems, uint32_t uniform_tiles, uint32_t *next) {
  level_r8_body<F, K, ODD, false>(xf, ent, nent, nitems, uniform_tiles, next);
}

// The level kernel for launches whose first passes may zero-pad (served LDE
// requests): padding products skipped.  Four waves a SIMD (at most 128
// VGPRs, as the LDS tile allows for every field); the skip's branches would
// otherwise take BN254 Fr's odd-K kernels to 129-130 VGPRs and three waves.
template <class F, int K, bool ODD>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void level_r8_pad_kernel(
    const BatchXform *__restrict__ xf, const BatchEntry *__restrict__ ent, uint32_t nent, uint32_t nitems,
    uint32_t uniform_tiles, uint32_t *next) {
  level_r8_body<F, K, ODD, true>(xf, ent, nent, nitems, uniform_tiles, next);
}

// Waves a SIMD that the whole-transform kernel asks the register allocator
// for: two 512-thread workgroups a CU (the LDS allows two for every field)
// need four, i.e. at most 128 VGPRs.  BN254 Fr's values do not fit that.
template <class F>
constexpr int whole_waves_per_simd() {
  return std::is_same<F, FrT>::value ? 1 : 4;
}

// Persistent kernel over a batch of transforms that each fit one tile: item
// i is transform ent[i].xf, run as a single pass of K = L stages with the
// register rounds above (coalesced staging, compile-time indexing).  Items
// are claimed from next[0] (next[0..2) zero at launch, and zero again when
// the launch ends); each workgroup claims its following item before working
// on the current one.  TWL: every
// transform of the launch uses the table `tw_all` (2^K entries), which each
// workgroup copies into LDS once, behind the tile's planes.
template <class F, int K, bool TWL>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(whole_waves_per_simd<F>()))) void whole_batch_kernel(const BatchXform *__restrict__ xf, const BatchEntry *__restrict__ ent, uint32_t n,
                                                          uint32_t *next, const uint32_t *__restrict__ tw_all) {
  using E = typename F::E;
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t s_item;
  if constexpr (TWL) {
    uint32_t *t = lds + F::W * (1 << K);
    for (int e = threadIdx.x; e < (1 << K); e += blockDim.x) F::lds_put(t, 1 << K, e, F::load_dev(tw_all, e));
  }
  if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
  __syncthreads();
  // The item index is the same in every lane: read it as a scalar, so the
  // descriptor loads below are scalar loads and the buffer pointers stay in
  // SGPRs.
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
    a.tw_one = 1;  // batches use tables of powers
    TileCtx c;
    c.K = K;
    c.lg = 0;
    c.G = 1;
    c.tile = 1 << K;
    c.lo_bits = 0;
    c.hi_bits = 0;
    c.hi = c.lo0 = c.c0 = 0;
    constexpr int N = round_elems<F>();
    E x[N];
    pass_rounds<F, K, 0, true, TWL, N, true>(x, a, lds, c);
    __syncthreads();  // LDS is reused by the next item
    if (threadIdx.x == 0) s_item = claimed;
    __syncthreads();
    it = __builtin_amdgcn_readfirstlane(s_item);
  }
  if (threadIdx.x == 0) retire_workgroup(next);
}

template <class F, int K, bool TWL>
static hipError_t launch_whole_k(const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                                 const uint32_t *tw_all, hipStream_t s) {
  const size_t lds = (size_t)F::W * 4 << K << (TWL ? 1 : 0);
  const unsigned threads = (1u << K) / round_elems<F>();
  // Resident workgroups on the whole device, computed once.
  static std::once_flag once;
  static int grid = 0;
  static hipError_t setup = hipSuccess;
  std::call_once(once, [&] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&whole_batch_kernel<F, K, TWL>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int per_cu = 0, dev = 0, cus = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, whole_batch_kernel<F, K, TWL>, (int)threads, lds);
    grid = std::max(1, per_cu) * std::max(1, cus);
    setup = e;
  });
  if (setup != hipSuccess) return setup;
  const unsigned g = (unsigned)std::min<uint32_t>(n, (uint32_t)grid);
  hipLaunchKernelGGL((whole_batch_kernel<F, K, TWL>), dim3(g), dim3(threads), lds, s, xf, ent, n, counter, tw_all);
  return hipGetLastError();
}

template <class F, int K>
static hipError_t launch_whole_kt(const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                                  const uint32_t *tw_all, hipStream_t s) {
  // The LDS-resident table doubles the tile's LDS; BN254 Fr's 72 KiB tile
  // would then fit once per CU, so it reads its twiddles from HBM.
  if constexpr (F::W <= 2) {
    if (tw_all) return launch_whole_k<F, K, true>(xf, ent, n, counter, tw_all, s);
  }
  return launch_whole_k<F, K, false>(xf, ent, n, counter, nullptr, s);
}

// Transforms of 2^9 up to the field's largest register-round tile (BN254 Fr
// 2^11, the other fields 2^12).
template <class F>
static hipError_t launch_whole_t(int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                                 const uint32_t *tw_all, hipStream_t s) {
  if (n == 0) return hipSuccess;
  switch (K) {
    case 9: return launch_whole_kt<F, 9>(xf, ent, n, counter, tw_all, s);
    case 10: return launch_whole_kt<F, 10>(xf, ent, n, counter, tw_all, s);
    case 11: return launch_whole_kt<F, 11>(xf, ent, n, counter, tw_all, s);
    case 12:
      if constexpr (F::W <= 2) return launch_whole_kt<F, 12>(xf, ent, n, counter, tw_all, s);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

// One level_r8_kernel launch (level_r8_pad_kernel with ZP) over `nitems`
// tiles of 2^tile_log elements (K + each pass's log_g = tile_log), at most
// `grid` workgroups and at most as many as are resident at once.
template <class F, int K, bool ODD, bool ZP>
static hipError_t launch_level_k(int tile_log, const BatchXform *xf, const BatchEntry *ent, uint32_t nent,
                                 uint32_t nitems, uint32_t uniform, uint32_t *counter, int grid, hipStream_t s) {
  constexpr int N = round_elems<F>();
  auto *const kern = ZP ? &level_r8_pad_kernel<F, K, ODD> : &level_r8_kernel<F, K, ODD>;
  const unsigned threads = (1u << tile_log) / N;
  const size_t lds = (size_t)F::W * 4 << tile_log;
  // Resident workgroups per CU for each tile size, computed once.
  static std::once_flag once;
  static int per_cu[13] = {};
  static int cus = 0;
  static hipError_t setup = hipSuccess;
  std::call_once(once, [&] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       F::W * 4 << (std::is_same<F, FrT>::value ? 11 : 12));
    int dev = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    for (int tl = 8; tl <= 12 && e == hipSuccess; tl++) {
      const unsigned t = (1u << tl) / N;
      if (t < 64 || t > 512) continue;
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[tl], kern, (int)t, (size_t)F::W * 4 << tl);
    }
    setup = e;
  });
  if (setup != hipSuccess) return setup;
  if (tile_log < 8 || tile_log > 12 || threads < 64 || threads > 512 || per_cu[tile_log] < 1)
    return hipErrorInvalidValue;
  const int g = std::min<int>(std::min<int>(grid, per_cu[tile_log] * cus), (int)nitems);
  if (g < 1) return hipSuccess;
  hipLaunchKernelGGL(kern, dim3(g), dim3(threads), lds, s, xf, ent, nent, nitems, uniform, counter);
  return hipGetLastError();
}

// Pass lengths 5..KMAX; zero_pad selects the padding-skipping kernel.
template <class F, int KMAX>
static hipError_t launch_level_t(int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                                 const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform,
                                 uint32_t *counter, int grid, hipStream_t s) {
  if (nitems == 0) return hipSuccess;
  switch (K * 4 + (odd ? 2 : 0) + (zero_pad ? 1 : 0)) {
#define ZNTT_LV1(k, o, z)                                                                                  \
  case 4 * k + 2 * o + z:                                                                                \
    if constexpr (k <= KMAX)                                                                             \
      return launch_level_k<F, k, o != 0, z != 0>(tile_log, xf, ent, nent, nitems, uniform, counter, grid, s); \
    break;
#define ZNTT_LV(k) ZNTT_LV1(k, 0, 0) ZNTT_LV1(k, 0, 1) ZNTT_LV1(k, 1, 0) ZNTT_LV1(k, 1, 1)
    ZNTT_LV(5) ZNTT_LV(6) ZNTT_LV(7) ZNTT_LV(8) ZNTT_LV(9) ZNTT_LV(10)
#undef ZNTT_LV
#undef ZNTT_LV1
    default: break;
  }
  return hipErrorInvalidValue;
}

// Pass lengths the planner produces for tiles of >= 512 elements: 5..12.
template <class F>
static hipError_t launch_r8(int K, bool odd, dim3 grid, dim3 block, size_t lds, hipStream_t s, const PassArgs &a) {
  constexpr int N = round_elems<F>();
  switch (K * 2 + (odd ? 1 : 0)) {
#define ZNTT_R8(k)                                                                                        \
  case 2 * k: hipLaunchKernelGGL((ntt_pass_r8_kernel<F, k, false, N>), grid, block, lds, s, a); break; \
  case 2 * k + 1: hipLaunchKernelGGL((ntt_pass_r8_kernel<F, k, true, N>), grid, block, lds, s, a); break;
    ZNTT_R8(5) ZNTT_R8(6) ZNTT_R8(7) ZNTT_R8(8) ZNTT_R8(9) ZNTT_R8(10) ZNTT_R8(11) ZNTT_R8(12)
#undef ZNTT_R8
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class F>
static hipError_t set_r8_lds(size_t bytes) {
  constexpr int N = round_elems<F>();
#define ZNTT_R8A(k, o)                                                                                       \
  {                                                                                                          \
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_r8_kernel<F, k, o, N>),     \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);              \
    if (e != hipSuccess) return e;                                                                           \
  }
#define ZNTT_R8B(k) ZNTT_R8A(k, false) ZNTT_R8A(k, true)
  ZNTT_R8B(5) ZNTT_R8B(6) ZNTT_R8B(7) ZNTT_R8B(8) ZNTT_R8B(9) ZNTT_R8B(10) ZNTT_R8B(11) ZNTT_R8B(12)
#undef ZNTT_R8B
#undef ZNTT_R8A
  return hipSuccess;
}

template <class F>
static hipError_t run_passes_t(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                             const uint32_t *coset, const uint32_t *scale_dev, bool tw_one, hipStream_t stream) {
  const PassPlan plan = plan_passes(field, L);
  int s0 = 1;
  for (int i = 0; i < plan.passes; i++) {
    PassArgs a = {};
    a.first = i == 0;
    a.last = i == plan.passes - 1;
    a.src = a.first ? data : scratch;
    a.dst = a.last ? data : scratch;
    a.tw = tw;
    a.coset = a.first ? coset : nullptr;
    a.has_scale = (a.last && scale_dev) ? 1 : 0;
    if (a.has_scale)
      for (int k = 0; k < F::W; k++) a.scale[k] = scale_dev[k];
    a.L = L;
    a.src_len = (uint64_t)1 << L;
    a.tw_one = tw_one ? 1 : 0;
    a.s0 = s0;
    a.K = plan.K[i];
    a.log_g = plan.log_g[i];
    const size_t tile = (size_t)1 << (a.K + a.log_g);
    const size_t blocks = ((size_t)1 << L) / tile;
    const size_t lds = (size_t)F::W * 4 * tile;
    constexpr int N = round_elems<F>();
    if (tile >= 512 && a.K >= 5 && tile / N <= 512) {  // N elements per thread, register rounds
      hipError_t e =
          launch_r8<F>(a.K, (a.s0 & 1) != 0, dim3((unsigned)blocks), dim3((unsigned)(tile / N)), lds, stream, a);
      if (e != hipSuccess) return e;
    } else {
      const int threads = (int)((tile / 2) < 64 ? 64 : tile / 2);
      hipLaunchKernelGGL(ntt_pass_kernel<F>, dim3((unsigned)blocks), dim3(threads), lds, stream, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    s0 += a.K;
  }
  return hipSuccess;
}

// Raises the dynamic-LDS limit of every pass kernel of field F to `bytes`.
template <class F>
static hipError_t configure_pass_lds(size_t bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_kernel<F>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return set_r8_lds<F>(bytes);
}

}  // namespace zntt
