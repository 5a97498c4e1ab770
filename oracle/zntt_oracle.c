/*
 * zntt_oracle.c -- CPU restatement of zig-algebra's NTT path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and never as the thing measured or shipped.  The product path
 * (zig-algebra_amd/) does not link, import or call anything here.
 *
 * What it restates (paths relative to the zig-algebra source tree):
 *   - BN254 Fr as four little-endian u64 limbs in Montgomery form, R = 2^256:
 *       constants           libs/field/src/montgomery.zig:36-79
 *       add / sub / neg     libs/field/src/montgomery.zig:197-245
 *       CIOS mul            libs/field/src/montgomery.zig:252-297
 *       binary-GCD inverse  libs/field/src/montgomery.zig:328-395
 *       to/fromMontgomery   libs/field/src/montgomery.zig:398-405
 *       pow (CT ladder)     libs/field/src/field.zig:1265-1287
 *       Fr modulus          libs/curve/src/bn254.zig:67
 *   - small fields (Goldilocks, BabyBear, KoalaBear, and M31 for the
 *     eight-lane transform) as a canonical u64 with u128 `%` arithmetic
 *     (libs/field/src/field.zig:89-424, pow at :534-556); the reference's
 *     Mersenne path for M31 (:298-368) computes the same canonical residues
 *   - the eight-lane M31 transform, Vec8NttM31 (libs/field/src/lib.zig:39-137)
 *   - roots of unity          libs/field/src/roots.zig:22-30, 149-157
 *   - the NTT itself          libs/ntt/src/root.zig:33-212 (all seven
 *     functions, with the same check order and the same butterfly DAG)
 *   - coset-forward NTT and LDE: absent from the reference; defined here as
 *     data[i] *= shift^i followed by ntt (and zero-pad first for the LDE).
 *   - seeded input generation: SplitMix64-seeded xoshiro256++, 32 (Fr) or 8
 *     (small field) little-endian bytes per draw, rejecting values >= modulus
 *     as libs/field/src/field.zig:982-990 and :224-232 do.
 *
 * Parity pins (tests/test_oracle.py): the reference's own DFT known-answer
 * vectors (libs/ntt/src/root.zig:438-449) and small-field arithmetic vectors
 * (libs/field/src/lib.zig:267-276, Goldilocks + BabyBear); Python-int
 * Montgomery/DFT/Horner restatements and textbook NTT digests for BN254 Fr
 * (tests/golden/); invariants (round trip, convolution, twiddled==plain).
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

typedef unsigned __int128 u128;

/* Status codes, numbered as the C ABI numbers them. */
enum {
  OR_OK = 0,
  OR_LOG_TOO_LARGE = 1,
  OR_LENGTH_MISMATCH = 2,
  OR_INVALID_LENGTH = 3,
  OR_INVALID_TWIDDLES = 4,
  OR_OUT_OF_MEMORY = 5,
  OR_ORDER_TOO_LARGE = 10,
  OR_INVALID_ARGUMENT = 8,
  OR_DEGREE_TOO_LARGE = 11,
};

/* ------------------------------------------------------------------------ */
/* BN254 Fr, Montgomery over [4]u64 (montgomery.zig)                          */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t l[4]; } fr_t;

static const uint64_t FR_MOD[4] = {
  0x43E1F593F0000001ull, 0x2833E84879B97091ull,
  0xB85045B68181585Dull, 0x30644E72E131A029ull };

static uint64_t FR_M0_INV;     /* -p^-1 mod 2^64, montgomery.zig:55-60 */
static fr_t FR_R2;             /* R^2 mod p, montgomery.zig:48-52 */
static fr_t FR_ONE_MONT;       /* R mod p, montgomery.zig:63-66 */
static fr_t FR_HALF_P_PLUS_ONE;/* (p+1)/2, montgomery.zig:76 */
static int fr_ready = 0;

/* montgomery.zig:143-150: borrow chain a < b */
static int fr_lt(const uint64_t *a, const uint64_t *b) {
  uint64_t borrow = 0;
  for (int i = 0; i < 4; i++) {
    u128 z = (u128)a[i] - (u128)b[i] - (u128)borrow;
    borrow = (uint64_t)(z >> 64) != 0;
  }
  return borrow != 0;
}

static fr_t fr_select(int on, fr_t a, fr_t b) { return on ? a : b; }

/* montgomery.zig:197-217 */
static inline fr_t fr_add(fr_t a, fr_t b) {
  fr_t out, red;
  uint64_t carry = 0;
  for (int i = 0; i < 4; i++) {
    u128 z = (u128)a.l[i] + b.l[i] + carry;
    out.l[i] = (uint64_t)z;
    carry = (uint64_t)(z >> 64);
  }
  int out_lt_mod = fr_lt(out.l, FR_MOD);
  int need_sub = (carry != 0) == out_lt_mod;
  uint64_t borrow = 0;
  for (int i = 0; i < 4; i++) {
    u128 z = (u128)out.l[i] - (u128)FR_MOD[i] - (u128)borrow;
    red.l[i] = (uint64_t)z;
    borrow = (uint64_t)(z >> 64) != 0;
  }
  return fr_select(need_sub, red, out);
}

/* montgomery.zig:220-240 */
static inline fr_t fr_sub(fr_t a, fr_t b) {
  fr_t out, red;
  uint64_t borrow = 0;
  for (int i = 0; i < 4; i++) {
    uint64_t ai = a.l[i], bi = b.l[i];
    uint64_t diff = ai - bi;
    uint64_t b1 = ai < bi;
    uint64_t b2 = diff < borrow;
    out.l[i] = diff - borrow;
    borrow = b1 | b2;
  }
  uint64_t carry = 0;
  for (int i = 0; i < 4; i++) {
    u128 z = (u128)out.l[i] + FR_MOD[i] + carry;
    red.l[i] = (uint64_t)z;
    carry = (uint64_t)(z >> 64);
  }
  return fr_select(borrow != 0, red, out);
}

static inline fr_t fr_zero(void) { fr_t z = {{0, 0, 0, 0}}; return z; }

/* montgomery.zig:243-245 */
static inline fr_t fr_neg(fr_t a) { return fr_sub(fr_zero(), a); }

/* montgomery.zig:252-297, CIOS with per-iteration shift */
static inline fr_t fr_mul(fr_t a, fr_t b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    uint64_t carry = 0;
    for (int j = 0; j < 4; j++) {
      u128 z = (u128)t[j] + (u128)a.l[i] * b.l[j] + carry;
      t[j] = (uint64_t)z;
      carry = (uint64_t)(z >> 64);
    }
    u128 zc = (u128)t[4] + carry;
    t[4] = (uint64_t)zc;
    t[5] = (uint64_t)(zc >> 64);

    uint64_t m = t[0] * FR_M0_INV;
    uint64_t c = 0;
    for (int j = 0; j < 4; j++) {
      u128 z = (u128)t[j] + (u128)m * FR_MOD[j] + c;
      t[j] = (uint64_t)z;
      c = (uint64_t)(z >> 64);
    }
    u128 zc2 = (u128)t[4] + c;
    t[4] = (uint64_t)zc2;
    t[5] += (uint64_t)(zc2 >> 64);
    for (int j = 0; j < 5; j++) t[j] = t[j + 1];
    t[5] = 0;
  }
  fr_t out = {{t[0], t[1], t[2], t[3]}}, red;
  uint64_t extra = t[4];
  int need_sub = (extra != 0) | !fr_lt(out.l, FR_MOD);
  uint64_t borrow = 0;
  for (int i = 0; i < 4; i++) {
    u128 z = (u128)out.l[i] - (u128)FR_MOD[i] - (u128)borrow;
    red.l[i] = (uint64_t)z;
    borrow = (uint64_t)(z >> 64) != 0;
  }
  return fr_select(need_sub, red, out);
}

static int fr_is_zero(const fr_t *a) { return (a->l[0] | a->l[1] | a->l[2] | a->l[3]) == 0; }
static int fr_eq(const fr_t *a, const fr_t *b) { return memcmp(a, b, sizeof(fr_t)) == 0; }

/* montgomery.zig:398-405 */
static fr_t fr_to_mont(fr_t x) { return fr_mul(x, FR_R2); }
static fr_t fr_from_mont(fr_t x) { fr_t one = {{1, 0, 0, 0}}; return fr_mul(x, one); }

/* montgomery.zig:174-183 */
static void fr_shr1(uint64_t *a) {
  uint64_t carry = 0;
  for (int i = 3; i >= 0; i--) {
    uint64_t low = a[i] << 63;
    a[i] = (a[i] >> 1) | carry;
    carry = low;
  }
}

/* montgomery.zig:328-332 */
static void fr_halve_mod(fr_t *x) {
  uint64_t odd = x->l[0] & 1;
  fr_shr1(x->l);
  if (odd) *x = fr_add(*x, FR_HALF_P_PLUS_ONE);
}

/* montgomery.zig:347-375 (canonical in, canonical out; zero -> zero) */
static fr_t fr_binary_gcd_inverse(fr_t a) {
  if (fr_is_zero(&a)) return fr_zero();
  fr_t u = a, v, x1 = {{1, 0, 0, 0}}, x2 = fr_zero();
  memcpy(v.l, FR_MOD, sizeof v.l);
  while (!fr_is_zero(&v)) {
    if ((u.l[0] & 1) == 0) {
      fr_shr1(u.l);
      fr_halve_mod(&x1);
    } else if ((v.l[0] & 1) == 0) {
      fr_shr1(v.l);
      fr_halve_mod(&x2);
    } else if (fr_lt(v.l, u.l)) {
      u = fr_sub(u, v);
      fr_shr1(u.l);
      x1 = fr_sub(x1, x2);
      fr_halve_mod(&x1);
    } else {
      v = fr_sub(v, u);
      fr_shr1(v.l);
      x2 = fr_sub(x2, x1);
      fr_halve_mod(&x2);
    }
  }
  return x1;
}

/* montgomery.zig:386-395 (total: inv(0) == 0) */
static fr_t fr_inv(fr_t x) {
  if (fr_is_zero(&x)) return fr_zero();
  return fr_to_mont(fr_binary_gcd_inverse(fr_from_mont(x)));
}

static fr_t fr_one(void) { return FR_ONE_MONT; }

/* field.zig:1265-1287: constant-time ladder over BITS = 254 exponent bits.
 * The exponent is four little-endian u64 limbs. */
static fr_t fr_pow_limbs(fr_t base, const uint64_t e[4]) {
  fr_t result = fr_one();
  for (int i = 0; i < 254; i++) {
    int bit = (int)((e[i >> 6] >> (i & 63)) & 1);
    fr_t m = fr_mul(result, base);
    result = fr_select(bit, m, result);
    base = fr_mul(base, base);
  }
  return result;
}

static fr_t fr_pow_u64(fr_t base, uint64_t e) {
  uint64_t ee[4] = {e, 0, 0, 0};
  return fr_pow_limbs(base, ee);
}

/* field.zig:911-921 for a u64 argument (always < p). */
static fr_t fr_from_u64(uint64_t x) { fr_t v = {{x, 0, 0, 0}}; return fr_to_mont(v); }

static void fr_init(void) {
  if (fr_ready) return;
  /* montgomery.zig:55-60: Newton iteration for p0^-1 mod 2^64 */
  uint64_t p0 = FR_MOD[0], inv = 1;
  for (int i = 0; i < 6; i++) inv = inv * (2 - p0 * inv);
  FR_M0_INV = 0 - inv;
  /* R mod p and R^2 mod p by repeated modular doubling of 1 (the reference
   * folds these at comptime from the integer modulus). */
  fr_t x = {{1, 0, 0, 0}};
  for (int i = 0; i < 512; i++) {
    /* x < p < 2^254, so 2x < 2^255 does not overflow the container */
    uint64_t c = 0;
    for (int k = 0; k < 4; k++) {
      uint64_t nc = x.l[k] >> 63;
      x.l[k] = (x.l[k] << 1) | c;
      c = nc;
    }
    if (!fr_lt(x.l, FR_MOD)) {
      uint64_t borrow = 0;
      for (int k = 0; k < 4; k++) {
        u128 z = (u128)x.l[k] - (u128)FR_MOD[k] - (u128)borrow;
        x.l[k] = (uint64_t)z;
        borrow = (uint64_t)(z >> 64) != 0;
      }
    }
    if (i == 255) FR_ONE_MONT = x;
  }
  FR_R2 = x;
  /* (p + 1) / 2 */
  fr_t h;
  memcpy(h.l, FR_MOD, sizeof h.l);
  fr_shr1(h.l);
  h.l[0] += 1; /* p odd: (p >> 1) + 1 == (p + 1) / 2, no carry since p>>1 ends in ...0 */
  FR_HALF_P_PLUS_ONE = h;
  fr_ready = 1;
}

/* ------------------------------------------------------------------------ */
/* Small field, canonical u64 value (field.zig:89-424, non-Mersenne path)      */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t v; } sf_t;

/* The oracle's small field is parameterised by a modulus held in a
 * per-call context, so the same code runs Goldilocks and BabyBear. */
typedef struct { uint64_t p; int bits; int two_adicity; uint64_t odd_part; } sf_ctx;

static sf_ctx sf_make(uint64_t p) {
  sf_ctx c;
  c.p = p;
  c.bits = 64 - __builtin_clzll(p);
  uint64_t q = p - 1;
  int s = 0;
  while ((q & 1) == 0) { q >>= 1; s++; }
  c.two_adicity = s;
  c.odd_part = q;
  return c;
}

/* field.zig:305-306 */
static inline sf_t sf_add(const sf_ctx *c, sf_t a, sf_t b) {
  u128 sum = (u128)a.v + b.v;
  sf_t r = { (uint64_t)(sum >= (u128)c->p ? sum - c->p : sum) };
  return r;
}
/* field.zig:350-355 */
static inline sf_t sf_sub(const sf_ctx *c, sf_t a, sf_t b) {
  sf_t r = { (uint64_t)(a.v >= b.v ? (u128)a.v - b.v : (u128)a.v + c->p - b.v) };
  return r;
}
/* field.zig:367-368 */
static inline sf_t sf_mul(const sf_ctx *c, sf_t a, sf_t b) {
  u128 wide = (u128)a.v * b.v;
  sf_t r = { (uint64_t)(wide % c->p) };
  return r;
}
/* field.zig:371-373 */
static inline sf_t sf_neg(const sf_ctx *c, sf_t a) {
  sf_t r = { a.v == 0 ? 0 : c->p - a.v };
  return r;
}
static inline sf_t sf_one(const sf_ctx *c) { (void)c; sf_t r = {1}; return r; }

/* field.zig:390-424 (total: inv(0) == 0) */
static sf_t sf_inv(const sf_ctx *c, sf_t s) {
  if (s.v == 0) { sf_t z = {0}; return z; }
  u128 a = s.v, b = c->p, x = 1, y = 0, m = c->p;
  while (b != 0) {
    if ((a & 1) == 0) {
      a >>= 1;
      if (x & 1) x += m;
      x >>= 1;
    } else if ((b & 1) == 0) {
      b >>= 1;
      if (y & 1) y += m;
      y >>= 1;
    } else if (a > b) {
      a -= b;
      if (x < y) x += m;
      x -= y;
      a >>= 1;
      if (x & 1) x += m;
      x >>= 1;
    } else {
      b -= a;
      if (y < x) y += m;
      y -= x;
      b >>= 1;
      if (y & 1) y += m;
      y >>= 1;
    }
  }
  sf_t r = {(uint64_t)x};
  return r;
}

/* field.zig:534-556: constant-time ladder over BITS exponent bits */
static sf_t sf_pow(const sf_ctx *c, sf_t base, uint64_t e) {
  sf_t result = sf_one(c);
  for (int i = 0; i < c->bits; i++) {
    int bit = (int)((e >> i) & 1);
    sf_t m = sf_mul(c, result, base);
    result = bit ? m : result;
    base = sf_mul(c, base, base);
  }
  return result;
}

/* ------------------------------------------------------------------------ */
/* roots.zig:22-30, 149-157                                                  */
/* ------------------------------------------------------------------------ */

static int sf_legendre(const sf_ctx *c, sf_t a) {
  uint64_t half = c->odd_part << (c->two_adicity - 1);
  sf_t r = sf_pow(c, a, half);
  if (r.v == 0) return 0;
  return r.v == 1 ? 1 : -1;
}

static int sf_primitive_root(const sf_ctx *c, size_t log_size, sf_t *out) {
  if (log_size > (size_t)c->two_adicity) return OR_ORDER_TOO_LARGE;
  sf_t z = {2 % c->p};
  while (sf_legendre(c, z) != -1) z = sf_add(c, z, sf_one(c));
  uint64_t exponent = c->odd_part << (c->two_adicity - log_size);
  *out = sf_pow(c, z, exponent);
  return OR_OK;
}

static const int FR_TWO_ADICITY = 28;

/* odd part of p - 1, shifted left by `shift` bits, as four limbs */
static void fr_odd_shifted(int shift, uint64_t e[4]) {
  uint64_t q[4];
  memcpy(q, FR_MOD, sizeof q);
  q[0] -= 1; /* p - 1 (p is odd, no borrow) */
  for (int i = 0; i < FR_TWO_ADICITY; i++) fr_shr1(q);
  /* q << shift */
  for (int k = 0; k < 4; k++) e[k] = q[k];
  for (int i = 0; i < shift; i++) {
    uint64_t c = 0;
    for (int k = 0; k < 4; k++) {
      uint64_t nc = e[k] >> 63;
      e[k] = (e[k] << 1) | c;
      c = nc;
    }
  }
}

static int fr_legendre(fr_t a) {
  uint64_t half[4];
  fr_odd_shifted(FR_TWO_ADICITY - 1, half);
  fr_t r = fr_pow_limbs(a, half);
  if (fr_is_zero(&r)) return 0;
  fr_t one = fr_one();
  return fr_eq(&r, &one) ? 1 : -1;
}

static int fr_primitive_root(size_t log_size, fr_t *out) {
  if (log_size > (size_t)FR_TWO_ADICITY) return OR_ORDER_TOO_LARGE;
  fr_t z = fr_from_u64(2);
  while (fr_legendre(z) != -1) z = fr_add(z, fr_one());
  uint64_t e[4];
  fr_odd_shifted(FR_TWO_ADICITY - (int)log_size, e);
  *out = fr_pow_limbs(z, e);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* The NTT (libs/ntt/src/root.zig:33-212), instantiated per field.            */
/* CTX is the field context argument (unused for Fr).                         */
/* ------------------------------------------------------------------------ */

/* root.zig:33-36 */
static int size_from_log(size_t log_n, size_t *n) {
  if (log_n >= 64) return OR_LOG_TOO_LARGE;
  *n = (size_t)1 << log_n;
  return OR_OK;
}

#define DEFINE_NTT(P, T, CTXDECL, ADD, SUB, MUL, NEG, ONE, INV, POW64, FROMU64)         \
  /* root.zig:38-58 */                                                             \
  static int P##_bit_reverse(T *data, size_t n) {                                  \
    if (n == 0 || (n & (n - 1)) != 0) return OR_INVALID_LENGTH;                    \
    int log_n = __builtin_ctzll(n);                                                \
    for (size_t i = 0; i < n; i++) {                                               \
      size_t j = 0;                                                                \
      for (int k = 0; k < log_n; k++) j = (j << 1) | ((i >> k) & 1);               \
      if (j > i) { T tmp = data[i]; data[i] = data[j]; data[j] = tmp; }            \
    }                                                                              \
    return OR_OK;                                                                  \
  }                                                                                \
  /* root.zig:65-90 */                                                             \
  static int P##_ntt(CTXDECL T *data, size_t len, size_t log_n, T root) {          \
    size_t n;                                                                      \
    int st = size_from_log(log_n, &n);                                             \
    if (st) return st;                                                             \
    if (len != n) return OR_LENGTH_MISMATCH;                                       \
    st = P##_bit_reverse(data, len);                                               \
    if (st) return st;                                                             \
    for (size_t s = 1; s <= log_n; s++) {                                          \
      size_t m = (size_t)1 << s, half_m = m >> 1;                                  \
      T wm = POW64(CTXARG root, (uint64_t)1 << (log_n - s));                       \
      for (size_t k = 0; k < n; k += m) {                                          \
        T w = ONE(CTXARG0);                                                        \
        for (size_t j = 0; j < half_m; j++) {                                      \
          T t = MUL(CTXARG w, data[k + j + half_m]);                               \
          T u = data[k + j];                                                       \
          data[k + j] = ADD(CTXARG u, t);                                          \
          data[k + j + half_m] = SUB(CTXARG u, t);                                 \
          w = MUL(CTXARG w, wm);                                                   \
        }                                                                          \
      }                                                                            \
    }                                                                              \
    return OR_OK;                                                                  \
  }                                                                                \
  /* root.zig:95-107 */                                                            \
  static int P##_intt(CTXDECL T *data, size_t len, size_t log_n, T root) {         \
    size_t n;                                                                      \
    int st = size_from_log(log_n, &n);                                             \
    if (st) return st;                                                             \
    if (len != n) return OR_LENGTH_MISMATCH;                                       \
    T root_inv = INV(CTXARG root);                                                 \
    st = P##_ntt(CTXARG data, len, log_n, root_inv);                               \
    if (st) return st;                                                             \
    T n_inv = INV(CTXARG FROMU64(CTXARG (uint64_t)n));                             \
    for (size_t i = 0; i < n; i++) data[i] = MUL(CTXARG data[i], n_inv);           \
    return OR_OK;                                                                  \
  }                                                                                \
  /* root.zig:112-134; stage s of the table is written at flat[2^s - 1] */         \
  static int P##_precompute_twiddles(CTXDECL size_t log_n, T root, T *flat) {      \
    size_t n;                                                                      \
    int st = size_from_log(log_n, &n);                                             \
    if (st) return st;                                                             \
    for (size_t s = 0; s < log_n; s++) {                                           \
      size_t half_m = (size_t)1 << s;                                              \
      T *stage = flat + (half_m - 1);                                              \
      T wm = POW64(CTXARG root, (uint64_t)1 << (log_n - s - 1));                   \
      T w = ONE(CTXARG0);                                                          \
      for (size_t j = 0; j < half_m; j++) { stage[j] = w; w = MUL(CTXARG w, wm); } \
    }                                                                              \
    return OR_OK;                                                                  \
  }                                                                                \
  /* root.zig:148-175 (stages given as a slice of slices) */                       \
  static int P##_ntt_with_twiddles(CTXDECL T *data, size_t len, size_t log_n,      \
                                   const T *const *stages, const size_t *lens,     \
                                   size_t num_stages) {                            \
    size_t n;                                                                      \
    int st = size_from_log(log_n, &n);                                             \
    if (st) return st;                                                             \
    if (len != n) return OR_LENGTH_MISMATCH;                                       \
    if (num_stages != log_n) return OR_INVALID_TWIDDLES;                           \
    st = P##_bit_reverse(data, len);                                               \
    if (st) return st;                                                             \
    for (size_t s = 1; s <= log_n; s++) {                                          \
      size_t m = (size_t)1 << s, half_m = m >> 1;                                  \
      const T *tw = stages[s - 1];                                                 \
      if (lens[s - 1] != half_m) return OR_INVALID_TWIDDLES;                       \
      for (size_t k = 0; k < n; k += m)                                            \
        for (size_t j = 0; j < half_m; j++) {                                      \
          T t = MUL(CTXARG tw[j], data[k + j + half_m]);                           \
          T u = data[k + j];                                                       \
          data[k + j] = ADD(CTXARG u, t);                                          \
          data[k + j + half_m] = SUB(CTXARG u, t);                                 \
        }                                                                          \
    }                                                                              \
    return OR_OK;                                                                  \
  }                                                                                \
  /* root.zig:180-212 */                                                           \
  static int P##_intt_with_twiddles(CTXDECL T *data, size_t len, size_t log_n,     \
                                    const T *const *stages, const size_t *lens,    \
                                    size_t num_stages) {                           \
    size_t n;                                                                      \
    int st = size_from_log(log_n, &n);                                             \
    if (st) return st;                                                             \
    if (len != n) return OR_LENGTH_MISMATCH;                                       \
    if (num_stages != log_n) return OR_INVALID_TWIDDLES;                           \
    st = P##_bit_reverse(data, len);                                               \
    if (st) return st;                                                             \
    for (size_t s = 1; s <= log_n; s++) {                                          \
      size_t m = (size_t)1 << s, half_m = m >> 1;                                  \
      const T *tw = stages[s - 1];                                                 \
      if (lens[s - 1] != half_m) return OR_INVALID_TWIDDLES;                       \
      for (size_t k = 0; k < n; k += m)                                            \
        for (size_t j = 0; j < half_m; j++) {                                      \
          T w = j == 0 ? ONE(CTXARG0) : NEG(CTXARG tw[half_m - j]);                \
          T t = MUL(CTXARG w, data[k + j + half_m]);                               \
          T u = data[k + j];                                                       \
          data[k + j] = ADD(CTXARG u, t);                                          \
          data[k + j + half_m] = SUB(CTXARG u, t);                                 \
        }                                                                          \
    }                                                                              \
    T n_inv = INV(CTXARG FROMU64(CTXARG (uint64_t)n));                             \
    for (size_t i = 0; i < n; i++) data[i] = MUL(CTXARG data[i], n_inv);           \
    return OR_OK;                                                                  \
  }                                                                                \
  /* Coset-forward NTT: data[i] *= shift^i, then ntt (not in the reference). */    \
  static int P##_coset_ntt(CTXDECL T *data, size_t len, size_t log_n, T root,      \
                           T shift) {                                              \
    size_t n;                                                                      \
    int st = size_from_log(log_n, &n);                                             \
    if (st) return st;                                                             \
    if (len != n) return OR_LENGTH_MISMATCH;                                       \
    T w = ONE(CTXARG0);                                                            \
    for (size_t i = 0; i < n; i++) {                                               \
      data[i] = MUL(CTXARG data[i], w);                                            \
      w = MUL(CTXARG w, shift);                                                    \
    }                                                                              \
    return P##_ntt(CTXARG data, len, log_n, root);                                 \
  }

/* Fr instantiation: no context argument. */
#define CTXARG
#define CTXARG0
#define FR_CTXDECL
DEFINE_NTT(fr, fr_t, FR_CTXDECL, fr_add, fr_sub, fr_mul, fr_neg, fr_one, fr_inv, fr_pow_u64, fr_from_u64)
#undef CTXARG
#undef CTXARG0

/* Small-field instantiation: a const sf_ctx *c threads through. */
static inline sf_t sf_from_u64(const sf_ctx *c, uint64_t x) { sf_t r = {x % c->p}; return r; }
#define CTXARG c,
#define CTXARG0 c
#define SF_CTXDECL const sf_ctx *c,
DEFINE_NTT(sf, sf_t, SF_CTXDECL, sf_add, sf_sub, sf_mul, sf_neg, sf_one, sf_inv, sf_pow, sf_from_u64)
#undef CTXARG
#undef CTXARG0

/* ------------------------------------------------------------------------ */
/* Seeded input generation (SplitMix64 -> xoshiro256++)                      */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t s[4]; } xo_t;

static uint64_t splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static void xo_seed(xo_t *x, uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 4; i++) x->s[i] = splitmix64(&st);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t xo_next(xo_t *x) {
  uint64_t *s = x->s;
  uint64_t result = rotl(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

static fr_t fr_draw(xo_t *x) {
  for (;;) {
    fr_t v;
    for (int i = 0; i < 4; i++) v.l[i] = xo_next(x);
    if (fr_lt(v.l, FR_MOD)) return fr_to_mont(v);
  }
}

static uint64_t sf_draw(const sf_ctx *c, xo_t *x) {
  int nbytes = (c->bits + 7) / 8;
  for (;;) {
    uint64_t v = xo_next(x);
    if (nbytes < 8) v &= (((uint64_t)1) << (8 * nbytes)) - 1;
    if (v < c->p) return v;
  }
}

/* ------------------------------------------------------------------------ */
/* Exported entry points (ctypes).  Field ids: 1 = BN254 Fr, 2 = Goldilocks,  */
/* 3 = BabyBear.  Elements are 32-byte Montgomery limbs (Fr) or canonical    */
/* u64 (small fields).                                                        */
/* ------------------------------------------------------------------------ */

#define FIELD_FR 1
#define FIELD_GL 2
#define FIELD_BB 3
#define FIELD_KB 4
#define FIELD_M31 5

static int sf_for(int field, sf_ctx *c) {
  if (field == FIELD_GL) { *c = sf_make(0xFFFFFFFF00000001ull); return 1; }
  if (field == FIELD_BB) { *c = sf_make(0x78000001ull); return 1; }
  if (field == FIELD_KB) { *c = sf_make(0x7F000001ull); return 1; }
  if (field == FIELD_M31) { *c = sf_make(0x7FFFFFFFull); return 1; }
  return 0;
}

int oracle_elem_size(int field) {
  if (field == FIELD_FR) return 32;
  if (field == FIELD_GL || field == FIELD_BB || field == FIELD_KB || field == FIELD_M31) return 8;
  return 0;
}

void oracle_init(void) { fr_init(); }

/* --- field arithmetic (for arithmetic pins) --- */
int oracle_field_op(int field, int op, const void *a, const void *b, void *out) {
  fr_init();
  if (field == FIELD_FR) {
    fr_t x, y, r;
    memcpy(&x, a, 32);
    if (b) memcpy(&y, b, 32); else y = fr_zero();
    switch (op) {
      case 0: r = fr_add(x, y); break;
      case 1: r = fr_sub(x, y); break;
      case 2: r = fr_mul(x, y); break;
      case 3: r = fr_neg(x); break;
      case 4: r = fr_inv(x); break;
      case 5: r = fr_to_mont(x); break;
      case 6: r = fr_from_mont(x); break;
      case 7: r = fr_pow_limbs(x, y.l); break;
      default: return OR_INVALID_ARGUMENT;
    }
    memcpy(out, &r, 32);
    return OR_OK;
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  sf_t x, y, r;
  memcpy(&x.v, a, 8);
  if (b) memcpy(&y.v, b, 8); else y.v = 0;
  switch (op) {
    case 0: r = sf_add(&c, x, y); break;
    case 1: r = sf_sub(&c, x, y); break;
    case 2: r = sf_mul(&c, x, y); break;
    case 3: r = sf_neg(&c, x); break;
    case 4: r = sf_inv(&c, x); break;
    case 5: r = x; break;              /* small fields have no Montgomery form */
    case 6: r = x; break;
    case 7: r = sf_pow(&c, x, y.v); break;
    default: return OR_INVALID_ARGUMENT;
  }
  memcpy(out, &r.v, 8);
  return OR_OK;
}

/* roots.zig:149-157 */
int oracle_primitive_root(int field, size_t log_size, void *out) {
  fr_init();
  if (field == FIELD_FR) {
    fr_t r;
    int st = fr_primitive_root(log_size, &r);
    if (st == OR_OK) memcpy(out, &r, 32);
    return st;
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  sf_t r;
  int st = sf_primitive_root(&c, log_size, &r);
  if (st == OR_OK) memcpy(out, &r.v, 8);
  return st;
}

int oracle_bit_reverse(int field, void *data, size_t len) {
  if (field == FIELD_FR) return fr_bit_reverse((fr_t *)data, len);
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  return sf_bit_reverse((sf_t *)data, len);
}

int oracle_ntt(int field, void *data, size_t len, size_t log_n, const void *root) {
  fr_init();
  if (field == FIELD_FR) { fr_t r; memcpy(&r, root, 32); return fr_ntt((fr_t *)data, len, log_n, r); }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  sf_t r; memcpy(&r.v, root, 8);
  return sf_ntt(&c, (sf_t *)data, len, log_n, r);
}

int oracle_intt(int field, void *data, size_t len, size_t log_n, const void *root) {
  fr_init();
  if (field == FIELD_FR) { fr_t r; memcpy(&r, root, 32); return fr_intt((fr_t *)data, len, log_n, r); }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  sf_t r; memcpy(&r.v, root, 8);
  return sf_intt(&c, (sf_t *)data, len, log_n, r);
}

int oracle_coset_ntt(int field, void *data, size_t len, size_t log_n, const void *root,
                     const void *shift) {
  fr_init();
  if (field == FIELD_FR) {
    fr_t r, g; memcpy(&r, root, 32); memcpy(&g, shift, 32);
    return fr_coset_ntt((fr_t *)data, len, log_n, r, g);
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  sf_t r, g; memcpy(&r.v, root, 8); memcpy(&g.v, shift, 8);
  return sf_coset_ntt(&c, (sf_t *)data, len, log_n, r, g);
}

/* Table layout: stage s (0-based, 2^s entries) at element offset 2^s - 1. */
int oracle_precompute_twiddles(int field, size_t log_n, const void *root, void *flat) {
  fr_init();
  if (field == FIELD_FR) { fr_t r; memcpy(&r, root, 32); return fr_precompute_twiddles(log_n, r, (fr_t *)flat); }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  sf_t r; memcpy(&r.v, root, 8);
  return sf_precompute_twiddles(&c, log_n, r, (sf_t *)flat);
}

int oracle_ntt_with_twiddles(int field, int inverse, void *data, size_t len, size_t log_n,
                             const void *const *stages, const size_t *lens, size_t num_stages) {
  fr_init();
  if (field == FIELD_FR) {
    return inverse ? fr_intt_with_twiddles((fr_t *)data, len, log_n, (const fr_t *const *)stages, lens, num_stages)
                   : fr_ntt_with_twiddles((fr_t *)data, len, log_n, (const fr_t *const *)stages, lens, num_stages);
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  return inverse ? sf_intt_with_twiddles(&c, (sf_t *)data, len, log_n, (const sf_t *const *)stages, lens, num_stages)
                 : sf_ntt_with_twiddles(&c, (sf_t *)data, len, log_n, (const sf_t *const *)stages, lens, num_stages);
}

/* LDE: zero-pad 2^log_k coefficients to 2^(log_k+log_blowup), then
 * coset-forward NTT with the field's primitive 2^(log_k+log_blowup)-th root. */
int oracle_lde(int field, const void *coeffs, size_t log_k, size_t log_blowup,
               const void *shift, void *out) {
  fr_init();
  size_t log_n = log_k + log_blowup;
  if (log_n >= 64) return OR_LOG_TOO_LARGE;
  size_t k = (size_t)1 << log_k, n = (size_t)1 << log_n;
  int es = oracle_elem_size(field);
  if (!es) return OR_INVALID_ARGUMENT;
  unsigned char root[32];
  int st = oracle_primitive_root(field, log_n, root);
  if (st) return st;
  memset(out, 0, n * (size_t)es);
  memcpy(out, coeffs, k * (size_t)es);
  return oracle_coset_ntt(field, out, n, log_n, root, shift);
}

/* Seeded uniform elements; Fr output is in Montgomery form. */
int oracle_gen_uniform(int field, uint64_t seed, void *out, size_t n) {
  fr_init();
  xo_t x;
  xo_seed(&x, seed);
  if (field == FIELD_FR) {
    fr_t *o = (fr_t *)out;
    for (size_t i = 0; i < n; i++) o[i] = fr_draw(&x);
    return OR_OK;
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  uint64_t *o = (uint64_t *)out;
  for (size_t i = 0; i < n; i++) o[i] = sf_draw(&c, &x);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Polynomials over the field (libs/poly/src/poly.zig).  Coefficient i is   */
/* the coefficient of x^i; the degree is that of the array without trailing */
/* zero coefficients, as the reference's normalised Polynomial keeps it.     */
/* ------------------------------------------------------------------------ */

static size_t poly_len(int field, const unsigned char *p, size_t n, int es) {
  while (n > 0) {
    const unsigned char *e = p + (n - 1) * (size_t)es;
    int zero = 1;
    for (int i = 0; i < es; i++) zero &= e[i] == 0;
    if (!zero) break;
    n--;
  }
  (void)field;
  return n;
}

/* Polynomial.mul, poly.zig:197-212: zero if either operand is zero (:198);
 * error.DegreeTooLarge if deg a + deg b > max_degree, here out_cap - 1
 * (:199-200); else r[i + j] += a[i] * b[j] over both degrees (:203-208). */
int oracle_poly_mul(int field, const void *a, size_t a_len, const void *b, size_t b_len, void *out,
                    size_t out_cap, size_t *out_len) {
  fr_init();
  int es = oracle_elem_size(field);
  if (!es) return OR_INVALID_ARGUMENT;
  *out_len = 0;
  size_t na = poly_len(field, (const unsigned char *)a, a_len, es);
  size_t nb = poly_len(field, (const unsigned char *)b, b_len, es);
  if (na == 0 || nb == 0) return OR_OK;
  size_t m = na + nb - 1;
  if (m > out_cap) return OR_DEGREE_TOO_LARGE;
  if (field == FIELD_FR) {
    const fr_t *x = (const fr_t *)a, *y = (const fr_t *)b;
    fr_t *r = (fr_t *)out;
    for (size_t k = 0; k < m; k++) r[k] = fr_zero();
    for (size_t i = 0; i < na; i++)
      for (size_t j = 0; j < nb; j++) r[i + j] = fr_add(r[i + j], fr_mul(x[i], y[j]));
  } else {
    sf_ctx c;
    if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
    const sf_t *x = (const sf_t *)a, *y = (const sf_t *)b;
    sf_t *r = (sf_t *)out;
    for (size_t k = 0; k < m; k++) r[k].v = 0;
    for (size_t i = 0; i < na; i++)
      for (size_t j = 0; j < nb; j++) r[i + j] = sf_add(&c, r[i + j], sf_mul(&c, x[i], y[j]));
  }
  *out_len = m;
  return OR_OK;
}

/* Polynomial.eval, poly.zig:215-224 (Horner from the leading coefficient). */
int oracle_poly_eval(int field, const void *coeffs, size_t n, const void *point, void *out) {
  fr_init();
  if (field == FIELD_FR) {
    const fr_t *c = (const fr_t *)coeffs;
    fr_t z, r = fr_zero();
    memcpy(&z, point, 32);
    for (size_t i = n; i-- > 0;) r = fr_add(fr_mul(r, z), c[i]);
    memcpy(out, &r, 32);
    return OR_OK;
  }
  sf_ctx cx;
  if (!sf_for(field, &cx)) return OR_INVALID_ARGUMENT;
  const sf_t *c = (const sf_t *)coeffs;
  sf_t z, r = {0};
  memcpy(&z.v, point, 8);
  for (size_t i = n; i-- > 0;) r = sf_add(&cx, sf_mul(&cx, r, z), c[i]);
  memcpy(out, &r.v, 8);
  return OR_OK;
}

/* Counter-based stream for synthetic request coefficients (a request's
 * coefficients are drawn independently of every other request's, so they can
 * be generated on the device as requests arrive).  Element i of stream
 * (seed, id) draws words mix64(key + (ctr + 1) * gamma), where mix64 is the
 * SplitMix64 finaliser, gamma its increment, key = mix64(mix64(seed) ^ id)
 * and ctr = i << 8 | attempt << 2 | word.  Fr: four little-endian words, the
 * top two bits cleared, accepted below r (first of up to 64 attempts) and
 * kept as the Montgomery words; small fields: one word masked to the
 * modulus' bit length (64 Goldilocks, 31 BabyBear / KoalaBear), accepted
 * below p. */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int oracle_fill_counter(int field, uint64_t seed, uint64_t stream_id, void *out, size_t n) {
  fr_init();
  const uint64_t key = mix64(mix64(seed) ^ stream_id), gamma = 0x9E3779B97F4A7C15ull;
  if (field == FIELD_FR) {
    fr_t *o = (fr_t *)out;
    for (size_t i = 0; i < n; i++) {
      fr_t v;
      for (uint64_t att = 0; att < 64; att++) {
        uint64_t ctr = (uint64_t)i << 8 | att << 2;
        for (int k = 0; k < 4; k++) v.l[k] = mix64(key + (ctr + (uint64_t)k + 1) * gamma);
        v.l[3] &= 0x3FFFFFFFFFFFFFFFull;
        if (fr_lt(v.l, FR_MOD)) break;
      }
      o[i] = v;
    }
    return OR_OK;
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  const uint64_t mask = c.bits >= 64 ? ~0ull : (((uint64_t)1 << c.bits) - 1);
  uint64_t *o = (uint64_t *)out;
  for (size_t i = 0; i < n; i++) {
    uint64_t w = 0;
    for (uint64_t att = 0; att < 64; att++) {
      w = mix64(key + (((uint64_t)i << 8 | att << 2) + 1) * gamma) & mask;
      if (w < c.p) break;
    }
    o[i] = w;
  }
  return OR_OK;
}

/* Raw xoshiro256++ stream, for pinning the generator itself. */
void oracle_xoshiro_stream(uint64_t seed, uint64_t *out, size_t n) {
  xo_t x;
  xo_seed(&x, seed);
  for (size_t i = 0; i < n; i++) out[i] = xo_next(&x);
}

/* ------------------------------------------------------------------------ */
/* Multi-threaded batch of independent forward transforms (the all-cores CPU */
/* baseline for batched workloads).  Transform b occupies                    */
/* data[offsets[b] .. offsets[b] + 2^log_ns[b]).                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  int field;
  unsigned char *data;
  const size_t *offsets;
  const size_t *log_ns;
  const void *roots;        /* per-log root table indexed by log_n */
  size_t count;
  size_t next;
  pthread_mutex_t mu;
  int status;
} batch_job;

static void *batch_worker(void *arg) {
  batch_job *j = (batch_job *)arg;
  int es = oracle_elem_size(j->field);
  for (;;) {
    pthread_mutex_lock(&j->mu);
    size_t b = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (b >= j->count) break;
    size_t log_n = j->log_ns[b];
    const unsigned char *root = (const unsigned char *)j->roots + log_n * (size_t)es;
    int st = oracle_ntt(j->field, j->data + j->offsets[b] * (size_t)es, (size_t)1 << log_n, log_n, root);
    if (st) j->status = st;
  }
  return NULL;
}

int oracle_ntt_batch_mt(int field, void *data, const size_t *offsets, const size_t *log_ns,
                        const void *roots_by_log, size_t count, int threads) {
  fr_init();
  if (threads < 1) threads = 1;
  batch_job j;
  j.field = field; j.data = (unsigned char *)data; j.offsets = offsets; j.log_ns = log_ns;
  j.roots = roots_by_log; j.count = count; j.next = 0; j.status = OR_OK;
  pthread_mutex_init(&j.mu, NULL);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  if (!th) return OR_OUT_OF_MEMORY;
  for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, batch_worker, &j);
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&j.mu);
  return j.status;
}

/* ------------------------------------------------------------------------ */
/* BLAKE3 and the FRI layer commitment.                                      */
/*                                                                           */
/* The FRI prover commits each layer f (n = 2^k values, natural order) by a  */
/* Merkle tree over n/2 leaves, leaf j = BLAKE3(toBytes(f[j]) ||             */
/* toBytes(f[j + n/2])) (libs/fri/src/root.zig:283-292, :378-384), built as  */
/* MerkleTree.initFromHashes (libs/merkle/src/merkle_tree.zig:147-182): the  */
/* leaf count is rounded up to a power of two with BLAKE3("") leaves,        */
/* nodes[cap + i] = leaf i, nodes[s + i] = BLAKE3(nodes[2(s+i)] ||           */
/* nodes[2(s+i)+1]) for each level start s, root = nodes[1], nodes[0] = 0.   */
/* toBytes is the canonical value, little-endian, in (bits + 7) / 8 bytes    */
/* (libs/field/src/field.zig:244-252 small fields, :1002-1016 Fr).           */
/*                                                                           */
/* The hash is std.crypto.hash.Blake3 of the Zig standard library (the FRI   */
/* and Merkle code import it, fri/src/root.zig:73), i.e. BLAKE3 as its       */
/* specification defines it: 64-byte blocks compressed by 7 rounds of the    */
/* quarter-round G over a 16-word state, 1024-byte chunks, chunk chaining    */
/* values merged pairwise into a binary tree, the last compression flagged   */
/* ROOT.  Pinned by the reference's BLAKE3 known-answer vectors              */
/* (libs/hash/src/root.zig:121-127, :173-247).                               */
/* ------------------------------------------------------------------------ */

enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };
static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
/* Word i of the next round's message is word B3_PERM[i] of this round's. */
static const uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
/* The four column and four diagonal quarter-rounds of one round. */
static const uint8_t B3_QR[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                    {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

static inline uint32_t ror32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void b3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter, uint32_t len,
                        uint32_t flags, uint32_t out[16]) {
  uint32_t v[16], m[16], t[16];
  for (int i = 0; i < 8; i++) { v[i] = cv[i]; v[8 + i] = i < 4 ? B3_IV[i] : 0; }
  v[12] = (uint32_t)counter;
  v[13] = (uint32_t)(counter >> 32);
  v[14] = len;
  v[15] = flags;
  memcpy(m, block, sizeof m);
  for (int r = 0; r < 7; r++) {
    for (int q = 0; q < 8; q++) {
      const int a = B3_QR[q][0], b = B3_QR[q][1], c = B3_QR[q][2], d = B3_QR[q][3];
      v[a] += v[b] + m[2 * q];
      v[d] = ror32(v[d] ^ v[a], 16);
      v[c] += v[d];
      v[b] = ror32(v[b] ^ v[c], 12);
      v[a] += v[b] + m[2 * q + 1];
      v[d] = ror32(v[d] ^ v[a], 8);
      v[c] += v[d];
      v[b] = ror32(v[b] ^ v[c], 7);
    }
    for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
    memcpy(m, t, sizeof m);
  }
  for (int i = 0; i < 8; i++) {
    out[i] = v[i] ^ v[8 + i];
    out[8 + i] = v[8 + i] ^ cv[i];
  }
}

static void b3_words(const uint8_t *p, size_t len, uint32_t w[16]) {
  uint8_t blk[64] = {0};
  memcpy(blk, p, len);
  for (int i = 0; i < 16; i++)
    w[i] = (uint32_t)blk[4 * i] | (uint32_t)blk[4 * i + 1] << 8 | (uint32_t)blk[4 * i + 2] << 16 |
           (uint32_t)blk[4 * i + 3] << 24;
}

static void b3_parent(const uint32_t l[8], const uint32_t r[8], uint32_t flags, uint32_t cv[8]) {
  uint32_t blk[16], o[16];
  memcpy(blk, l, 32);
  memcpy(blk + 8, r, 32);
  b3_compress(B3_IV, blk, 0, 64, B3_PARENT | flags, o);
  memcpy(cv, o, 32);
}

/* One-shot BLAKE3 with a 32-byte output (unkeyed). */
void oracle_blake3(const uint8_t *in, size_t len, uint8_t out[32]) {
  uint32_t stack[64][8];
  int depth = 0;
  uint64_t chunk = 0;
  size_t pos = 0;
  for (;;) {
    const size_t rest = len - pos;
    const int last = rest <= 1024;
    const size_t clen = last ? rest : 1024;
    const size_t nblk = clen == 0 ? 1 : (clen + 63) / 64;
    uint32_t cv[8];
    memcpy(cv, B3_IV, sizeof cv);
    for (size_t b = 0; b < nblk; b++) {
      const size_t blen = clen - b * 64 < 64 ? clen - b * 64 : 64;
      uint32_t w[16], o[16];
      b3_words(in + pos + b * 64, blen, w);
      uint32_t flags = (b == 0 ? B3_CHUNK_START : 0) | (b + 1 == nblk ? B3_CHUNK_END : 0);
      if (last && depth == 0 && b + 1 == nblk) flags |= B3_ROOT;
      b3_compress(cv, w, chunk, (uint32_t)blen, flags, o);
      memcpy(cv, o, sizeof cv);
    }
    if (last) {
      /* The final chunk closes every subtree still open on the stack. */
      while (depth > 0) {
        depth--;
        b3_parent(stack[depth], cv, depth == 0 ? B3_ROOT : 0, cv);
      }
      for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) out[4 * i + k] = (uint8_t)(cv[i] >> (8 * k));
      return;
    }
    /* A chunk followed by more input: merge every completed subtree. */
    chunk++;
    for (uint64_t total = chunk; (total & 1) == 0; total >>= 1) b3_parent(stack[--depth], cv, 0, cv);
    memcpy(stack[depth++], cv, sizeof cv);
    pos += 1024;
  }
}

/* toBytes of element i: the canonical value, little-endian; returns the byte count. */
static int to_bytes(int field, const void *elems, size_t i, uint8_t *out) {
  if (field == FIELD_FR) {
    fr_t x;
    memcpy(&x, (const uint8_t *)elems + 32 * i, 32);
    x = fr_from_mont(x);
    for (int k = 0; k < 4; k++)
      for (int b = 0; b < 8; b++) out[8 * k + b] = (uint8_t)(x.l[k] >> (8 * b));
    return 32;
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return 0;
  const uint64_t v = ((const uint64_t *)elems)[i];
  const int nb = (c.bits + 7) / 8;
  for (int b = 0; b < nb; b++) out[b] = (uint8_t)(v >> (8 * b));
  return nb;
}

/* leaves[j] = BLAKE3(toBytes(f[j]) || toBytes(f[j + n/2])), j < n/2. */
int oracle_fri_leaves(int field, const void *evals, size_t n, uint8_t *leaves) {
  fr_init();
  if (oracle_elem_size(field) == 0) return OR_INVALID_ARGUMENT;
  if (n < 2 || (n & 1)) return OR_INVALID_LENGTH;
  const size_t half = n / 2;
  for (size_t j = 0; j < half; j++) {
    uint8_t buf[64];
    const int a = to_bytes(field, evals, j, buf);
    const int b = to_bytes(field, evals, j + half, buf + a);
    oracle_blake3(buf, (size_t)(a + b), leaves + 32 * j);
  }
  return OR_OK;
}

/* MerkleTree.initFromHashes: nodes holds 2 * cap * 32 bytes, cap the leaf
 * count rounded up to a power of two. */
int oracle_merkle_from_hashes(const uint8_t *hashes, size_t count, uint8_t *nodes, uint8_t root[32]) {
  if (count == 0) return OR_INVALID_LENGTH;
  size_t cap = 1;
  while (cap < count) cap <<= 1;
  memset(nodes, 0, 32);
  memcpy(nodes + 32 * cap, hashes, 32 * count);
  uint8_t empty[32];
  oracle_blake3((const uint8_t *)"", 0, empty);
  for (size_t i = count; i < cap; i++) memcpy(nodes + 32 * (cap + i), empty, 32);
  for (size_t s = cap / 2; s >= 1; s /= 2)
    for (size_t i = 0; i < s; i++) oracle_blake3(nodes + 64 * (s + i), 64, nodes + 32 * (s + i));
  memcpy(root, nodes + 32, 32);
  return OR_OK;
}

/* The commitment of one FRI layer of n = 2^k values: nodes holds n * 32 bytes. */
int oracle_merkle_commit(int field, const void *evals, size_t n, uint8_t *nodes, uint8_t root[32]) {
  if (n < 2 || (n & (n - 1))) return OR_INVALID_LENGTH;
  uint8_t *leaves = (uint8_t *)malloc(16 * n);
  if (!leaves) return OR_OUT_OF_MEMORY;
  int st = oracle_fri_leaves(field, evals, n, leaves);
  if (st == OR_OK) st = oracle_merkle_from_hashes(leaves, n / 2, nodes, root);
  free(leaves);
  return st;
}

/* ------------------------------------------------------------------------ */
/* Fiat-Shamir transcript (libs/transcript/src/root.zig:38-180).             */
/*                                                                           */
/* The transcript is a BLAKE3 hasher fed a byte stream: init writes the      */
/* label's length (u64 little-endian) and the label, absorbBytes the data's  */
/* length and the data.  A squeeze finalises the hash of everything fed since */
/* the last re-key, hands out its first bytes (re-hashing the digest for     */
/* each further 32), and re-keys: the stream restarts as the last digest.    */
/* So the state is exactly the bytes fed since the last re-key, which is how */
/* this restatement keeps it.                                                */
/* ------------------------------------------------------------------------ */

typedef struct { uint8_t *buf; size_t len, cap; } oracle_tr;

static void tr_feed(oracle_tr *t, const void *p, size_t n) {
  if (t->len + n > t->cap) {
    size_t cap = t->cap ? t->cap : 256;
    while (cap < t->len + n) cap *= 2;
    t->buf = (uint8_t *)realloc(t->buf, cap);
    t->cap = cap;
  }
  if (n) memcpy(t->buf + t->len, p, n);
  t->len += n;
}

static void tr_feed_len(oracle_tr *t, uint64_t n) {
  uint8_t b[8];
  for (int i = 0; i < 8; i++) b[i] = (uint8_t)(n >> (8 * i));
  tr_feed(t, b, 8);
}

/* Transcript.init(label) */
oracle_tr *oracle_tr_new(const uint8_t *label, size_t len) {
  oracle_tr *t = (oracle_tr *)calloc(1, sizeof *t);
  tr_feed_len(t, len);
  tr_feed(t, label, len);
  return t;
}

void oracle_tr_free(oracle_tr *t) {
  if (!t) return;
  free(t->buf);
  free(t);
}

/* absorbBytes */
void oracle_tr_absorb(oracle_tr *t, const uint8_t *data, size_t len) {
  tr_feed_len(t, len);
  tr_feed(t, data, len);
}

/* challengeBytes */
void oracle_tr_challenge_bytes(oracle_tr *t, uint8_t *out, size_t n) {
  uint8_t d[32];
  oracle_blake3(t->buf ? t->buf : (const uint8_t *)"", t->len, d);
  size_t off = 0;
  while (off < n) {
    const size_t k = n - off < 32 ? n - off : 32;
    memcpy(out + off, d, k);
    off += k;
    if (off < n) oracle_blake3(d, 32, d);
  }
  t->len = 0;
  tr_feed(t, d, 32);
}

/* challengeU64 */
uint64_t oracle_tr_challenge_u64(oracle_tr *t) {
  uint8_t b[8];
  oracle_tr_challenge_bytes(t, b, 8);
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v |= (uint64_t)b[i] << (8 * i);
  return v;
}

/* challengeFieldChecked: draw NUM_BYTES little-endian bytes until they encode
 * a value below the modulus (fromBytesChecked, field.zig:217-221 and
 * :975-979); out receives the element in the storage layout. */
int oracle_tr_challenge_field(oracle_tr *t, int field, void *out) {
  fr_init();
  if (field == FIELD_FR) {
    for (;;) {
      uint8_t b[32];
      oracle_tr_challenge_bytes(t, b, 32);
      fr_t v;
      for (int k = 0; k < 4; k++) {
        v.l[k] = 0;
        for (int i = 0; i < 8; i++) v.l[k] |= (uint64_t)b[8 * k + i] << (8 * i);
      }
      if (fr_lt(v.l, FR_MOD)) {
        v = fr_to_mont(v);
        memcpy(out, &v, 32);
        return OR_OK;
      }
    }
  }
  sf_ctx c;
  if (!sf_for(field, &c)) return OR_INVALID_ARGUMENT;
  const int nb = (c.bits + 7) / 8;
  for (;;) {
    uint8_t b[8];
    oracle_tr_challenge_bytes(t, b, (size_t)nb);
    uint64_t v = 0;
    for (int i = 0; i < nb; i++) v |= (uint64_t)b[i] << (8 * i);
    if (v < c.p) {
      memcpy(out, &v, 8);
      return OR_OK;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FRI prover (libs/fri/src/root.zig:311-441, buildQueries :506-573).        */
/* ------------------------------------------------------------------------ */

/* One element of any field, with the field's operations. */
typedef union { fr_t fr; sf_t sf; } fe_t;
typedef struct { int field; sf_ctx c; size_t es; int nb; } fctx_t;

static int fctx_make(int field, fctx_t *f) {
  f->field = field;
  if (field == FIELD_FR) { f->es = 32; f->nb = 32; return 1; }
  if (!sf_for(field, &f->c)) return 0;
  f->es = 8;
  f->nb = (f->c.bits + 7) / 8;
  return 1;
}
static fe_t fe_load(const fctx_t *f, const void *p) { fe_t x; memset(&x, 0, sizeof x); memcpy(&x, p, f->es); return x; }
static fe_t fe_add(const fctx_t *f, fe_t a, fe_t b) {
  if (f->field == FIELD_FR) a.fr = fr_add(a.fr, b.fr); else a.sf = sf_add(&f->c, a.sf, b.sf);
  return a;
}
static fe_t fe_sub(const fctx_t *f, fe_t a, fe_t b) {
  if (f->field == FIELD_FR) a.fr = fr_sub(a.fr, b.fr); else a.sf = sf_sub(&f->c, a.sf, b.sf);
  return a;
}
static fe_t fe_mul(const fctx_t *f, fe_t a, fe_t b) {
  if (f->field == FIELD_FR) a.fr = fr_mul(a.fr, b.fr); else a.sf = sf_mul(&f->c, a.sf, b.sf);
  return a;
}
static fe_t fe_inv(const fctx_t *f, fe_t a) {
  if (f->field == FIELD_FR) a.fr = fr_inv(a.fr); else a.sf = sf_inv(&f->c, a.sf);
  return a;
}
static fe_t fe_from_u64(const fctx_t *f, uint64_t v) {
  fe_t x;
  memset(&x, 0, sizeof x);
  if (f->field == FIELD_FR) x.fr = fr_from_u64(v); else x.sf = sf_from_u64(&f->c, v);
  return x;
}
static fe_t fe_one(const fctx_t *f) { return fe_from_u64(f, 1); }
static void fe_bytes(const fctx_t *f, fe_t x, uint8_t *out) { to_bytes(f->field, &x, 0, out); }

/* FriError.InvalidParameters */
#define OR_FRI_INVALID_PARAMETERS 12

/* Config.rounds and Config.validate (root.zig:188-203). */
static int fri_rounds(size_t log_domain, size_t log_initial_degree, size_t log_final, size_t log_residual_degree,
                      size_t num_queries, size_t *rounds) {
  if (log_domain == 0 || log_domain > 63) return OR_FRI_INVALID_PARAMETERS;
  if (log_final >= log_domain) return OR_FRI_INVALID_PARAMETERS;
  if (log_residual_degree >= log_final) return OR_FRI_INVALID_PARAMETERS;
  if (log_final > 12 || log_residual_degree > 12) return OR_FRI_INVALID_PARAMETERS;
  if (num_queries == 0) return OR_FRI_INVALID_PARAMETERS;
  const size_t r = log_domain - log_final;
  if (log_initial_degree > log_domain) return OR_FRI_INVALID_PARAMETERS;
  if (log_initial_degree < r) return OR_FRI_INVALID_PARAMETERS;
  if (log_initial_degree - r != log_residual_degree) return OR_FRI_INVALID_PARAMETERS;
  *rounds = r;
  return OR_OK;
}

/* Domain(F).init(F, log_n).step_gen = omega^(2^(two_adicity - log_n)) with
 * omega = primitiveRootOfUnity(two_adicity) (root.zig:133-138), which is
 * the primitive root of order 2^log_n of roots.zig:149-157. */
static fe_t fri_domain_gen(const fctx_t *f, size_t log_n) {
  fe_t g;
  memset(&g, 0, sizeof g);
  oracle_primitive_root(f->field, log_n, &g);
  return g;
}

/* Coefficients of the polynomial of degree < m through (g^i, v_i), i < m:
 * c_j = m^-1 * sum_i v_i g^(-ij), the unique solution of the Vandermonde
 * system interpolateToCoeffs solves by elimination (root.zig:448-504). */
static void fri_interpolate(const fctx_t *f, const fe_t *v, size_t m, fe_t g, fe_t *c) {
  const fe_t ginv = fe_inv(f, g), minv = fe_inv(f, fe_from_u64(f, m));
  fe_t gj = fe_one(f);  /* g^-j */
  for (size_t j = 0; j < m; j++) {
    fe_t acc, x = fe_one(f);  /* g^-ij */
    memset(&acc, 0, sizeof acc);  /* zero in either representation */
    for (size_t i = 0; i < m; i++) {
      acc = fe_add(f, acc, fe_mul(f, v[i], x));
      x = fe_mul(f, x, gj);
    }
    c[j] = fe_mul(f, acc, minv);
    gj = fe_mul(f, gj, ginv);
  }
}

/* prove(F, allocator, transcript, evals, config).  Outputs, in flat arrays:
 *   roots      rounds x 32          (layers[r].merkle_root; log_size is log_domain - r)
 *   residual   2^log_residual_degree x NUM_BYTES (serializeElem of each coefficient)
 *   pair_index num_queries
 *   values     num_queries x rounds x 2 NUM_BYTES (serializePair)
 *   paths      num_queries x sum_r (log_domain - r - 1) x 32, round by round,
 *              each path from the leaf's sibling up (MerkleTree.prove). */
int oracle_fri_prove(int field, oracle_tr *t, const void *evals, size_t n, size_t log_domain,
                     size_t log_initial_degree, size_t log_final, size_t log_residual_degree, size_t num_queries,
                     uint8_t *roots, uint8_t *residual, uint64_t *pair_index, uint8_t *values, uint8_t *paths) {
  fr_init();
  fctx_t f;
  if (!fctx_make(field, &f)) return OR_INVALID_ARGUMENT;
  size_t rounds;
  int st = fri_rounds(log_domain, log_initial_degree, log_final, log_residual_degree, num_queries, &rounds);
  if (st) return st;
  if (n != ((size_t)1 << log_domain)) return OR_FRI_INVALID_PARAMETERS;
  const int adicity = field == FIELD_FR ? 28 : f.c.two_adicity;
  if (log_domain > (size_t)adicity) return OR_FRI_INVALID_PARAMETERS;

  fe_t **layer = (fe_t **)calloc(rounds + 1, sizeof(fe_t *));
  uint8_t **tree = (uint8_t **)calloc(rounds, sizeof(uint8_t *));
  layer[0] = (fe_t *)calloc(n, sizeof(fe_t));
  for (size_t i = 0; i < n; i++) layer[0][i] = fe_load(&f, (const uint8_t *)evals + i * f.es);
  const fe_t half_inv = fe_inv(&f, fe_from_u64(&f, 2));
  size_t len = n, log_cur = log_domain;
  for (size_t r = 0; r < rounds; r++) {
    const fe_t *cur = layer[r];
    const size_t half = len / 2;
    /* commit: leaves over (j, j + half), initFromHashes, root */
    uint8_t *leaves = (uint8_t *)malloc(32 * half);
    for (size_t j = 0; j < half; j++) {
      uint8_t buf[64];
      fe_bytes(&f, cur[j], buf);
      fe_bytes(&f, cur[j + half], buf + f.nb);
      oracle_blake3(buf, (size_t)(2 * f.nb), leaves + 32 * j);
    }
    tree[r] = (uint8_t *)malloc(64 * half);
    oracle_merkle_from_hashes(leaves, half, tree[r], roots + 32 * r);
    free(leaves);
    /* absorb the root, squeeze alpha */
    oracle_tr_absorb(t, roots + 32 * r, 32);
    fe_t alpha;
    memset(&alpha, 0, sizeof alpha);
    oracle_tr_challenge_field(t, field, &alpha);
    /* fold (root.zig:392-403) */
    const fe_t g = fri_domain_gen(&f, log_cur);
    fe_t *next = (fe_t *)calloc(half, sizeof(fe_t));
    fe_t x = fe_one(&f);
    for (size_t j = 0; j < half; j++) {
      const fe_t fx = cur[j], fnx = cur[j + half];
      const fe_t even = fe_mul(&f, fe_add(&f, fx, fnx), half_inv);
      const fe_t odd = fe_mul(&f, fe_mul(&f, fe_sub(&f, fx, fnx), fe_inv(&f, x)), half_inv);
      next[j] = fe_add(&f, even, fe_mul(&f, alpha, odd));
      x = fe_mul(&f, x, g);
    }
    layer[r + 1] = next;
    len = half;
    log_cur--;
  }
  /* residual: interpolate the final layer, keep 2^log_residual_degree coefficients, absorb them */
  const size_t m = len, rl = (size_t)1 << log_residual_degree;
  fe_t *coef = (fe_t *)calloc(m, sizeof(fe_t));
  fri_interpolate(&f, layer[rounds], m, fri_domain_gen(&f, log_final), coef);
  for (size_t i = 0; i < rl; i++) {
    fe_bytes(&f, coef[i], residual + i * (size_t)f.nb);
    oracle_tr_absorb(t, residual + i * (size_t)f.nb, (size_t)f.nb);
  }
  free(coef);
  /* queries */
  uint8_t *pp = paths;
  for (size_t q = 0; q < num_queries; q++) {
    const uint64_t seed = oracle_tr_challenge_u64(t);
    size_t e = (size_t)(seed % (uint64_t)n);
    pair_index[q] = e;
    size_t lr = n;
    for (size_t r = 0; r < rounds; r++) {
      const size_t half = lr / 2, j = e % half;
      uint8_t *v = values + (q * rounds + r) * 2 * (size_t)f.nb;
      fe_bytes(&f, layer[r][j], v);
      fe_bytes(&f, layer[r][j + half], v + f.nb);
      for (size_t pos = half + j; pos > 1; pos >>= 1) {
        memcpy(pp, tree[r] + 32 * (pos ^ 1), 32);
        pp += 32;
      }
      e = j;
      lr = half;
    }
  }
  for (size_t r = 0; r <= rounds; r++) free(layer[r]);
  for (size_t r = 0; r < rounds; r++) free(tree[r]);
  free(layer);
  free(tree);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Vec8NttM31 (libs/field/src/lib.zig:52-137): eight transforms interleaved  */
/* as data[i * 8 + lane] over M31, restated loop for loop: per-lane bit      */
/* reversal, then stages s = 1..log_n, wm = root^(2^(log_n - s)), w stepped  */
/* by w *= wm, butterflies (u + w v, u - w v) reduced to canonical values.   */
/* ------------------------------------------------------------------------ */

static int vec8_forward(const sf_ctx *c, uint64_t *data, size_t len, size_t log_n, uint64_t root) {
  if (log_n >= 61) return OR_INVALID_LENGTH;  /* transformSize, lib.zig:131-136 */
  const size_t n = (size_t)1 << log_n;
  if (len != 8 * n) return OR_INVALID_LENGTH;
  for (size_t lane = 0; lane < 8; lane++)
    for (size_t i = 0; i < n; i++) {
      size_t j = 0;
      for (size_t k = 0; k < log_n; k++) j = (j << 1) | ((i >> k) & 1);
      if (j > i) {
        const uint64_t t = data[i * 8 + lane];
        data[i * 8 + lane] = data[j * 8 + lane];
        data[j * 8 + lane] = t;
      }
    }
  const sf_t r = {root};
  for (size_t s = 1; s <= log_n; s++) {
    const size_t m = (size_t)1 << s, half_m = m >> 1;
    const sf_t wm = sf_pow(c, r, (uint64_t)1 << (log_n - s));
    for (size_t k = 0; k < n; k += m) {
      sf_t w = sf_one(c);
      for (size_t j = 0; j < half_m; j++) {
        for (size_t l = 0; l < 8; l++) {
          const sf_t u = {data[(k + j) * 8 + l]}, v = {data[(k + j + half_m) * 8 + l]};
          const sf_t t = sf_mul(c, w, v);
          data[(k + j) * 8 + l] = sf_add(c, u, t).v;
          data[(k + j + half_m) * 8 + l] = sf_sub(c, u, t).v;
        }
        w = sf_mul(c, w, wm);
      }
    }
  }
  return OR_OK;
}

int oracle_vec8_m31(int inverse, uint64_t *data, size_t len, size_t log_n, uint64_t root) {
  sf_ctx c;
  sf_for(FIELD_M31, &c);
  if (!inverse) return vec8_forward(&c, data, len, log_n, root);
  const sf_t r = {root};
  int st = vec8_forward(&c, data, len, log_n, sf_inv(&c, r).v);
  if (st) return st;
  const sf_t ninv = sf_inv(&c, sf_from_u64(&c, (uint64_t)1 << log_n));
  for (size_t i = 0; i < len; i++) {
    const sf_t x = {data[i]};
    data[i] = sf_mul(&c, x, ninv).v;
  }
  return OR_OK;
}
