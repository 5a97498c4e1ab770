// Field traits shared by the generic NTT kernels: element type, the user
// (reference) and device layouts, LDS plane access and the butterfly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_fr.hpp"
#include "field_f31.hpp"
#include "field_gl.hpp"

namespace zntt {

// Transform buffers, tables and scratch are always global memory.  Accessed
// through address-space-1 pointers they compile to global_* instructions; a
// pointer read from a descriptor in memory would otherwise give flat_* ones,
// which also count against the LDS wait counter, so every LDS wait after them
// would wait for HBM as well.
#define ZNTT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const ZNTT_GLOBAL T *gld(const void *p) {
  return (const ZNTT_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ ZNTT_GLOBAL T *gst(void *p) {
  return (ZNTT_GLOBAL T *)p;
}
// Plain vector types for wide accesses through those pointers (HIP's uint4
// class has no copy from another address space).
typedef uint32_t zu32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t zu32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Field traits for the generic kernels.
// ---------------------------------------------------------------------------
struct FrT {
  using E = Fr9;
  static constexpr int W = 9;   // dev words
  static constexpr int UW = 8;  // user words
  __device__ static E load_user(const uint32_t *base, size_t i) {
    const ZNTT_GLOBAL zu32x4 *p = gld<zu32x4>(base + i * 8);
    const zu32x4 a = p[0], b = p[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return fr_unpack8(w);
  }
  __device__ static void store_user(uint32_t *base, size_t i, E x) {
    fr_canon(x);
    uint32_t w[8];
    fr_pack8(x, w);
    ZNTT_GLOBAL zu32x4 *p = gst<zu32x4>(base + i * 8);
    p[0] = zu32x4{w[0], w[1], w[2], w[3]};
    p[1] = zu32x4{w[4], w[5], w[6], w[7]};
  }
  __device__ static E load_dev(const uint32_t *base, size_t i) {
    E x;
    const ZNTT_GLOBAL uint32_t *p = gld<uint32_t>(base + i * 9);
#pragma unroll
    for (int k = 0; k < 9; k++) x.v[k] = p[k];
    return x;
  }
  __device__ static void store_dev(uint32_t *base, size_t i, const E &x) {
    ZNTT_GLOBAL uint32_t *p = gst<uint32_t>(base + i * 9);
#pragma unroll
    for (int k = 0; k < 9; k++) p[k] = x.v[k];
  }
  __device__ static E lds_get(const uint32_t *lds, int plane, int i) {
    E x;
#pragma unroll
    for (int k = 0; k < 9; k++) x.v[k] = lds[k * plane + i];
    return x;
  }
  __device__ static void lds_put(uint32_t *lds, int plane, int i, const E &x) {
#pragma unroll
    for (int k = 0; k < 9; k++) lds[k * plane + i] = x.v[k];
  }
  __device__ static E from_words(const uint32_t *w) {
    E x;
#pragma unroll
    for (int k = 0; k < 9; k++) x.v[k] = w[k];
    return x;
  }
  __device__ static E zero() {
    E x;
#pragma unroll
    for (int k = 0; k < 9; k++) x.v[k] = 0;
    return x;
  }
  __device__ static E mul(const E &a, const E &b) { return fr_mul(a, b); }
  // x0 *= b0 and x1 *= b1 with the two products interleaved.
  __device__ static void mul2(E &x0, const E &b0, E &x1, const E &b1) {
    E t0, t1;
    fr_mul2(x0, b0, x1, b1, t0, t1);
    x0 = t0;
    x1 = t1;
  }
  template <bool NORM = true>
  __device__ static void bfly(E &u, E &v, const E &w) {
    fr_butterfly<NORM>(u, v, w);
  }
  // Two independent butterflies issued together.
  template <bool NORM = true>
  __device__ static void bfly2(E &u0, E &v0, const E &w0, E &u1, E &v1, const E &w1) {
    fr_butterfly2<NORM>(u0, v0, w0, u1, v1, w1);
  }
  // Twiddle 1 in the first two stages of a transform (stages T = 0, 1 of its
  // first round).  At T = 0, v is a loaded or multiplied value; at T = 1 it
  // is a first-stage sum of two of them (LATE).
  template <bool NORM = true, bool LATE = false>
  __device__ static void bfly_one(E &u, E &v) {
    if constexpr (LATE)
      fr_butterfly_one_c8<NORM>(u, v);
    else
      fr_butterfly_one<NORM>(u, v);
  }
  static constexpr bool one_ok(int T) { return T <= 1; }
  // Butterfly whose v is zero (the first stages over a zero-padded input):
  // (u + w 0, u - w 0) = (u, u), with no product.  The copy keeps u's bound,
  // below what a full butterfly would give.
  template <bool NORM = true>
  __device__ static void bfly_zero(E &u, E &v) {
    if (NORM) fr_norm(u);
    v = u;
  }
  // Canonical dev-form constant (for tables).
  __device__ static E canon(E x) {
    fr_canon(x);
    return x;
  }
  __device__ static E neg(const E &x) { return fr_neg_canon(x); }
  __device__ static E add_lazy(const E &a, const E &b) {
    E o;
#pragma unroll
    for (int k = 0; k < 9; k++) o.v[k] = a.v[k] + b.v[k];
    return o;
  }
};

struct GlT {
  using E = uint64_t;
  static constexpr int W = 2;
  static constexpr int UW = 2;
  __device__ static E load_user(const uint32_t *base, size_t i) {
    return gld<uint64_t>(base)[i];
  }
  // Transform values are weakly reduced (any u64); the caller's buffer gets
  // the canonical value.
  __device__ static void store_user(uint32_t *base, size_t i, E x) {
    gst<uint64_t>(base)[i] = gl_canon(x);
  }
  __device__ static E load_dev(const uint32_t *base, size_t i) { return load_user(base, i); }
  __device__ static void store_dev(uint32_t *base, size_t i, const E &x) {
    gst<uint64_t>(base)[i] = x;
  }
  __device__ static E lds_get(const uint32_t *lds, int plane, int i) {
    return (uint64_t)lds[i] | ((uint64_t)lds[plane + i] << 32);
  }
  __device__ static void lds_put(uint32_t *lds, int plane, int i, const E &x) {
    lds[i] = (uint32_t)x;
    lds[plane + i] = (uint32_t)(x >> 32);
  }
  __device__ static E from_words(const uint32_t *w) { return (uint64_t)w[0] | ((uint64_t)w[1] << 32); }
  __device__ static E zero() { return 0; }
  __device__ static E mul(const E &a, const E &b) { return gl_mul(a, b); }
  __device__ static void mul2(E &x0, const E &b0, E &x1, const E &b1) {
    x0 = gl_mul(x0, b0);
    x1 = gl_mul(x1, b1);
  }
  template <bool NORM = true>
  __device__ static void bfly(E &u, E &v, const E &w) {
    gl_butterfly(u, v, w);
  }
  template <bool NORM = true>
  __device__ static void bfly2(E &u0, E &v0, const E &w0, E &u1, E &v1, const E &w1) {
    gl_butterfly(u0, v0, w0);
    gl_butterfly(u1, v1, w1);
  }
  template <bool NORM = true, bool LATE = false>
  __device__ static void bfly_one(E &u, E &v) {
    gl_butterfly_one(u, v);
  }
  static constexpr bool one_ok(int) { return true; }
  template <bool NORM = true>
  __device__ static void bfly_zero(E &u, E &v) {
    v = u;
  }
  __device__ static E canon(E x) { return gl_canon(x); }
  __device__ static E neg(const E &x) { return gl_neg(x); }
};

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) {
  return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

// 31-bit fields: one u32 per element in registers, LDS and scratch (dev
// layout 4 B); the caller's buffer holds u64 (user layout 8 B).
template <uint32_t P, uint32_t NPRIME>
struct F31T {
  using A = F31Arith<P, NPRIME>;
  using E = uint32_t;
  static constexpr int W = 1;
  static constexpr int UW = 2;
  static constexpr uint32_t kP = P;
  __device__ static E load_user(const uint32_t *base, size_t i) { return gld<uint32_t>(base)[2 * i]; }
  __device__ static void store_user(uint32_t *base, size_t i, E x) { gst<zu32x2>(base)[i] = zu32x2{x, 0u}; }
  __device__ static E load_dev(const uint32_t *base, size_t i) { return gld<uint32_t>(base)[i]; }
  __device__ static void store_dev(uint32_t *base, size_t i, const E &x) { gst<uint32_t>(base)[i] = x; }
  __device__ static E lds_get(const uint32_t *lds, int, int i) { return lds[i]; }
  __device__ static void lds_put(uint32_t *lds, int, int i, const E &x) { lds[i] = x; }
  __device__ static E from_words(const uint32_t *w) { return w[0]; }
  __device__ static E zero() { return 0; }
  __device__ static E mul(const E &a, const E &b) { return A::mont(a, b); }
  __device__ static void mul2(E &x0, const E &b0, E &x1, const E &b1) {
    x0 = A::mont(x0, b0);
    x1 = A::mont(x1, b1);
  }
  template <bool NORM = true>
  __device__ static void bfly(E &u, E &v, const E &w) {
    A::butterfly(u, v, w);
  }
  template <bool NORM = true>
  __device__ static void bfly2(E &u0, E &v0, const E &w0, E &u1, E &v1, const E &w1) {
    A::butterfly(u0, v0, w0);
    A::butterfly(u1, v1, w1);
  }
  template <bool NORM = true, bool LATE = false>
  __device__ static void bfly_one(E &u, E &v) {
    A::butterfly_one(u, v);
  }
  static constexpr bool one_ok(int) { return true; }
  template <bool NORM = true>
  __device__ static void bfly_zero(E &u, E &v) {
    v = u;
  }
  __device__ static E canon(E x) { return A::canon(x); }
  __device__ static E neg(const E &x) { return x == 0 ? 0 : P - x; }
};

using BbT = F31T<kBbP, kBbNPrime>;
using KbT = F31T<kKbP, kKbNPrime>;

// Calls fn(F{}) with the traits type of a field id (1 BN254 Fr, 2 Goldilocks,
// 3 BabyBear, 4 KoalaBear).
template <class Fn>
__host__ auto with_field(int field, Fn &&fn) {
  switch (field) {
    case 1: return fn(FrT{});
    case 2: return fn(GlT{});
    case 3: return fn(BbT{});
    default: return fn(KbT{});
  }
}

}  // namespace zntt
