// Host-side field arithmetic for the library's own bookkeeping: root inverses,
// n^-1, primitive roots, root-chain comparisons for the twiddle cache, and the
// conversion of constants into the device's 2^261 Montgomery form.  Never on
// the hot path (a handful of operations per call).
#pragma once
#include <stdint.h>
#include <string.h>

namespace zntt {
namespace host {

typedef unsigned __int128 u128;

struct U256 {
  uint64_t l[4];
  bool operator==(const U256 &o) const { return memcmp(l, o.l, sizeof l) == 0; }
  bool operator!=(const U256 &o) const { return !(*this == o); }
  bool is_zero() const { return (l[0] | l[1] | l[2] | l[3]) == 0; }
};

// ---------------------------------------------------------------------------
// BN254 Fr with R = 2^256 (the reference's storage form, "Zig form").
// ---------------------------------------------------------------------------
struct Fr {
  static constexpr U256 P = {{0x43E1F593F0000001ull, 0x2833E84879B97091ull,
                              0xB85045B68181585Dull, 0x30644E72E131A029ull}};
  static constexpr uint64_t N0 = 0xC2E1F593EFFFFFFFull;  // -P^-1 mod 2^64

  static bool lt(const U256 &a, const U256 &b) {
    for (int i = 3; i >= 0; i--)
      if (a.l[i] != b.l[i]) return a.l[i] < b.l[i];
    return false;
  }
  static U256 sub_raw(const U256 &a, const U256 &b, uint64_t *borrow_out) {
    U256 o;
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) {
      u128 z = (u128)a.l[i] - b.l[i] - br;
      o.l[i] = (uint64_t)z;
      br = (uint64_t)(z >> 64) ? 1 : 0;
    }
    *borrow_out = br;
    return o;
  }
  static U256 add(const U256 &a, const U256 &b) {
    U256 o;
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
      u128 z = (u128)a.l[i] + b.l[i] + c;
      o.l[i] = (uint64_t)z;
      c = (uint64_t)(z >> 64);
    }
    if (c || !lt(o, P)) {
      uint64_t br;
      o = sub_raw(o, P, &br);
    }
    return o;
  }
  static U256 sub(const U256 &a, const U256 &b) {
    uint64_t br;
    U256 o = sub_raw(a, b, &br);
    if (br) {
      uint64_t c = 0;
      for (int i = 0; i < 4; i++) {
        u128 z = (u128)o.l[i] + P.l[i] + c;
        o.l[i] = (uint64_t)z;
        c = (uint64_t)(z >> 64);
      }
    }
    return o;
  }
  static U256 neg(const U256 &a) { return sub(U256{{0, 0, 0, 0}}, a); }
  // a * b * 2^-256 mod P (operand-scanning Montgomery, 64-bit words)
  static U256 mont(const U256 &a, const U256 &b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
      uint64_t c = 0;
      for (int j = 0; j < 4; j++) {
        u128 z = (u128)a.l[i] * b.l[j] + t[j] + c;
        t[j] = (uint64_t)z;
        c = (uint64_t)(z >> 64);
      }
      u128 z = (u128)t[4] + c;
      t[4] = (uint64_t)z;
      t[5] = (uint64_t)(z >> 64);
      const uint64_t m = t[0] * N0;
      c = (uint64_t)(((u128)m * P.l[0] + t[0]) >> 64);
      for (int j = 1; j < 4; j++) {
        u128 y = (u128)m * P.l[j] + t[j] + c;
        t[j - 1] = (uint64_t)y;
        c = (uint64_t)(y >> 64);
      }
      u128 y = (u128)t[4] + c;
      t[3] = (uint64_t)y;
      t[4] = t[5] + (uint64_t)(y >> 64);
    }
    U256 o = {{t[0], t[1], t[2], t[3]}};
    if (t[4] || !lt(o, P)) {
      uint64_t br;
      o = sub_raw(o, P, &br);
    }
    return o;
  }
  static U256 dbl(const U256 &a) { return add(a, a); }
  static U256 pow2k_mod(int k) {  // 2^k mod P, canonical
    U256 x = {{1, 0, 0, 0}};
    for (int i = 0; i < k; i++) x = dbl(x);
    return x;
  }
  static U256 r2() {
    static const U256 v = pow2k_mod(512);
    return v;
  }
  static U256 one() {  // Zig form of 1
    static const U256 v = pow2k_mod(256);
    return v;
  }
  static U256 from_canonical(const U256 &x) { return mont(x, r2()); }
  static U256 to_canonical(const U256 &z) { return mont(z, U256{{1, 0, 0, 0}}); }
  static U256 from_u64(uint64_t v) { return from_canonical(U256{{v, 0, 0, 0}}); }
  // Zig-form power with a 256-bit exponent (square-and-multiply).
  static U256 pow(const U256 &z, const U256 &e) {
    U256 r = one(), b = z;
    for (int i = 0; i < 256; i++) {
      if ((e.l[i >> 6] >> (i & 63)) & 1) r = mont(r, b);
      b = mont(b, b);
    }
    return r;
  }
  static U256 pow_u64(const U256 &z, uint64_t e) { return pow(z, U256{{e, 0, 0, 0}}); }
  // Multiplicative inverse (0 -> 0, as the reference's total inv()).
  static U256 inv(const U256 &z) {
    if (z.is_zero()) return z;
    uint64_t br;
    U256 e = sub_raw(P, U256{{2, 0, 0, 0}}, &br);
    return pow(z, e);
  }
  // Device constant form: nine 29-bit limbs of c * 2^261 mod P, from the Zig
  // form z = c * 2^256 (five modular doublings).
  static void to_dev(const U256 &z, uint32_t out[9]) {
    U256 d = z;
    for (int i = 0; i < 5; i++) d = dbl(d);
    for (int i = 0; i < 9; i++) {
      const int bit = 29 * i, w = bit >> 6, s = bit & 63;
      uint64_t v = d.l[w] >> s;
      if (s > 35 && w + 1 < 4) v |= d.l[w + 1] << (64 - s);
      out[i] = (uint32_t)(v & 0x1FFFFFFFull);
    }
  }
  // Plain integer (< P) as nine 29-bit limbs.
  static void to_limbs29(const U256 &x, uint32_t out[9]) {
    for (int i = 0; i < 9; i++) {
      const int bit = 29 * i, w = bit >> 6, s = bit & 63;
      uint64_t v = x.l[w] >> s;
      if (s > 35 && w + 1 < 4) v |= x.l[w + 1] << (64 - s);
      out[i] = (uint32_t)(v & 0x1FFFFFFFull);
    }
  }
  static U256 load(const void *p) {
    U256 x;
    memcpy(x.l, p, 32);
    return x;
  }
  static void store(void *p, const U256 &x) { memcpy(p, x.l, 32); }
  static constexpr int kTwoAdicity = 28;
  // Primitive 2^log-th root of unity in Zig form: z^(odd << (28 - log)) with
  // z = 5, the smallest quadratic non-residue (roots.zig:149-157).
  static U256 primitive_root(int log_size) {
    uint64_t br;
    U256 q = sub_raw(P, U256{{1, 0, 0, 0}}, &br);
    for (int i = 0; i < kTwoAdicity; i++) {  // odd part of P - 1
      for (int k = 0; k < 3; k++) q.l[k] = (q.l[k] >> 1) | (q.l[k + 1] << 63);
      q.l[3] >>= 1;
    }
    U256 r = pow(from_u64(5), q);
    for (int i = 0; i < kTwoAdicity - log_size; i++) r = mont(r, r);
    return r;
  }
};

// ---------------------------------------------------------------------------
// Goldilocks p = 2^64 - 2^32 + 1, canonical u64 (the reference's storage form).
// ---------------------------------------------------------------------------
// 31-bit prime fields (BabyBear, KoalaBear): canonical values; the device
// keeps constants in Montgomery form c * 2^32 mod p.
struct F31 {
  uint64_t p;
  int two_adicity;
  uint64_t qnr;  // smallest quadratic non-residue (roots.zig:149-157)
  uint64_t mul(uint64_t a, uint64_t b) const { return a * b % p; }
  uint64_t pow(uint64_t a, uint64_t e) const {
    uint64_t r = 1, b = a % p;
    while (e) {
      if (e & 1) r = mul(r, b);
      b = mul(b, b);
      e >>= 1;
    }
    return r;
  }
  uint64_t inv(uint64_t a) const { return a % p == 0 ? 0 : pow(a, p - 2); }
  // z^(odd << (two_adicity - log_size)) = z^((p - 1) >> log_size)
  uint64_t primitive_root(int log_size) const { return pow(qnr, (p - 1) >> log_size); }
  uint32_t to_mont(uint64_t a) const { return (uint32_t)(((a % p) << 32) % p); }
};

constexpr F31 kBabyBear = {0x78000001ull, 27, 11};
constexpr F31 kKoalaBear = {0x7f000001ull, 24, 3};

struct Gl {
  static constexpr uint64_t P = 0xFFFFFFFF00000001ull;
  static uint64_t mul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }
  static uint64_t add(uint64_t a, uint64_t b) {
    u128 s = (u128)a + b;
    return (uint64_t)(s >= P ? s - P : s);
  }
  static uint64_t pow(uint64_t a, uint64_t e) {
    uint64_t r = 1, b = a;
    while (e) {
      if (e & 1) r = mul(r, b);
      b = mul(b, b);
      e >>= 1;
    }
    return r;
  }
  static uint64_t inv(uint64_t a) { return a == 0 ? 0 : pow(a, P - 2); }
  static constexpr int kTwoAdicity = 32;
  // z = 7 is the smallest quadratic non-residue; odd part of p - 1 is 2^32 - 1.
  static uint64_t primitive_root(int log_size) {
    uint64_t r = pow(7, 0xFFFFFFFFull);
    for (int i = 0; i < kTwoAdicity - log_size; i++) r = mul(r, r);
    return r;
  }
};

}  // namespace host
}  // namespace zntt
