// Host build of the device Goldilocks arithmetic (csrc/field_gl.hpp, compiled
// by g++ as plain functions) checked against unsigned __int128 arithmetic.
//
// gl_reduce128 picks one of three multiples of eps = 2^32 - 1 to subtract
// (carried; not carried with r >= eps; not carried with r < eps).  Random
// products reach the third case about once in 2^32 draws, so inputs are also
// built to land in each case, and the case each input took is counted: the
// run fails if any case was never reached.
#include "field_gl.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

using u128 = unsigned __int128;
using zntt::kGlEps;
using zntt::kGlP;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next64() {  // xorshift64*
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return rng_state * 0x2545F4914F6CDD1Dull;
}

static uint64_t mod128(uint64_t lo, uint64_t hi) { return (uint64_t)((((u128)hi << 64) | lo) % kGlP); }

// Which multiple of eps gl_reduce128 subtracts for (lo, hi), restated from
// the hi-word split: 0 carried, 1 not carried and r >= eps, 2 otherwise.
static int reduce_case(uint64_t lo, uint64_t hi) {
  const uint64_t z = (uint64_t)(uint32_t)hi * kGlEps + (uint32_t)~(uint32_t)(hi >> 32);
  const uint64_t r = lo + z;
  if (r < z) return 0;
  return r < kGlEps ? 2 : 1;
}

static long failures = 0, cases[3] = {0, 0, 0};

static void check_reduce(uint64_t lo, uint64_t hi) {
  const uint64_t got = zntt::gl_reduce128(lo, hi), want = mod128(lo, hi);
  cases[reduce_case(lo, hi)]++;
  if (got != want && failures++ < 10)
    printf("reduce lo=%016llx hi=%016llx got %016llx want %016llx\n", (unsigned long long)lo,
           (unsigned long long)hi, (unsigned long long)got, (unsigned long long)want);
}

static void check_mul(uint64_t a, uint64_t b) {
  const u128 x = (u128)a * b;
  const uint64_t got = zntt::gl_mul(a, b), want = (uint64_t)(x % kGlP);
  if (got != want && failures++ < 10)
    printf("mul a=%016llx b=%016llx got %016llx want %016llx\n", (unsigned long long)a, (unsigned long long)b,
           (unsigned long long)got, (unsigned long long)want);
}

// u, v weakly reduced (any u64), w canonical: outputs congruent to u + wv and
// u - wv.
static void check_butterfly(uint64_t u, uint64_t v, uint64_t w) {
  uint64_t s = u, d = v;
  zntt::gl_butterfly(s, d, w);
  const uint64_t t = (uint64_t)(((u128)v * w) % kGlP), um = u % kGlP;
  const uint64_t ws = (uint64_t)(((u128)um + t) % kGlP), wd = (uint64_t)(((u128)um + kGlP - t) % kGlP);
  bool ok = s % kGlP == ws && d % kGlP == wd;
  uint64_t s1 = u, d1 = v;
  zntt::gl_butterfly_one(s1, d1);
  const uint64_t vm = v % kGlP;
  ok = ok && s1 % kGlP == (uint64_t)(((u128)um + vm) % kGlP) && d1 % kGlP == (uint64_t)(((u128)um + kGlP - vm) % kGlP);
  if (!ok && failures++ < 10)
    printf("butterfly u=%016llx v=%016llx w=%016llx\n", (unsigned long long)u, (unsigned long long)v,
           (unsigned long long)w);
}

// x * 2^e mod p, canonical, for the shift products of the radix-8 rounds
// (e = 24, 48, 72: the powers of two that are 8th roots of unity), any u64 x.
static void check_shifts(uint64_t x) {
  const uint64_t xm = x % kGlP;
  const uint64_t w24 = (uint64_t)(((u128)xm << 24) % kGlP);
  const uint64_t w48 = (uint64_t)(((u128)w24 << 24) % kGlP);
  const uint64_t w72 = (uint64_t)(((u128)w48 << 24) % kGlP);
  const uint64_t g24 = zntt::gl_shl24(x), g48 = zntt::gl_shl48(x), g72 = zntt::gl_shl72(x);
  if ((g24 != w24 || g48 != w48 || g72 != w72) && failures++ < 10)
    printf("shift x=%016llx got %016llx %016llx %016llx want %016llx %016llx %016llx\n", (unsigned long long)x,
           (unsigned long long)g24, (unsigned long long)g48, (unsigned long long)g72, (unsigned long long)w24,
           (unsigned long long)w48, (unsigned long long)w72);
  // 2^72 mod p as the kernel compares it against the table's 8th root.
  if (zntt::kGl2p72 != (uint64_t)(((u128)1 << 72) % kGlP) && failures++ < 10) printf("kGl2p72 wrong\n");
}

// gl_butterfly_t: u any u64, t < p; outputs congruent to u + t and u - t.
static void check_butterfly_t(uint64_t u, uint64_t t) {
  uint64_t s = u, d = 0;
  zntt::gl_butterfly_t(s, d, t);
  const uint64_t um = u % kGlP;
  if ((s % kGlP != (uint64_t)(((u128)um + t) % kGlP) || d % kGlP != (uint64_t)(((u128)um + kGlP - t) % kGlP)) &&
      failures++ < 10)
    printf("butterfly_t u=%016llx t=%016llx\n", (unsigned long long)u, (unsigned long long)t);
}

int main(int argc, char **argv) {
  const long draws = argc > 1 ? atol(argv[1]) : 4000000;
  const uint64_t edges[] = {0,
                            1,
                            2,
                            kGlEps - 1,
                            kGlEps,
                            kGlEps + 1,
                            (uint64_t)1 << 32,
                            kGlP - 2,
                            kGlP - 1,
                            kGlP,
                            kGlP + 1,
                            ~0ull - 1,
                            ~0ull,
                            0x8000000000000000ull,
                            0xFFFFFFFE00000002ull,
                            0x00000001FFFFFFFEull};
  const int ne = sizeof edges / sizeof edges[0];
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      check_reduce(edges[i], edges[j]);
      check_mul(edges[i], edges[j]);
      if (edges[j] < kGlP) check_butterfly_t(edges[i], edges[j]);
      for (int k = 0; k < ne; k++)
        if (edges[k] < kGlP) check_butterfly(edges[i], edges[j], edges[k]);
    }
  for (int i = 0; i < ne; i++) check_shifts(edges[i]);
  for (int e = 0; e < 64; e++) check_shifts((uint64_t)1 << e), check_shifts(~((uint64_t)1 << e));
  for (long it = 0; it < draws; it++) {
    const uint64_t a = next64(), b = next64(), c = next64();
    check_shifts(a);
    check_butterfly_t(a, b % kGlP);
    // Random 128-bit values and random products.
    check_reduce(a, b);
    check_mul(a, b);
    check_butterfly(a, b, c % kGlP);
    // Not carried, r < eps: h0 = 0 and lo + ~h1 < eps.
    const uint32_t h1 = 0xFFFFFFFFu - (uint32_t)(c & 0xFFFF);
    check_reduce(a & 0xFFFF, (uint64_t)h1 << 32);
    // Carried: lo near 2^64 and z large.
    check_reduce(~0ull - (a & 0xFFFFF), b | 0xFFFFFFFFull);
  }
  printf("reduce cases: carried %ld, r >= eps %ld, r < eps %ld\n", cases[0], cases[1], cases[2]);
  if (!cases[0] || !cases[1] || !cases[2]) {
    printf("a reduction case was never reached\n");
    return 1;
  }
  if (failures) {
    printf("%ld failures\n", failures);
    return 1;
  }
  printf("ok gl_reduce\n");
  return 0;
}
