"""The BN254 Fr canonicalisation's quotient estimate (csrc/field_fr.hpp,
fr_canon): for every normalised top limb x8 < 2^29 (values below 2^261), the
estimate floor(x8 * M / 2^40) never exceeds floor(x / r) and falls short of it
by at most one, so one conditional subtraction of r leaves the canonical
value.  Read from the kernel source so the test pins the constants the
kernels use."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "zig-algebra_amd", "csrc", "field_fr.hpp")
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001


def constants():
    text = open(SRC).read()
    m = int(re.search(r"kFrQMagic\s*=\s*(\d+)u", text).group(1))
    limbs = re.search(r"#define ZNTT_FR_P29 \{([^}]*)\}", text).group(1)
    p = [int(v.strip().rstrip("u"), 16) for v in limbs.replace("\\", " ").split(",")]
    return m, p


def test_limbs_are_r():
    _, p = constants()
    assert sum(v << (29 * i) for i, v in enumerate(p)) == R


def test_estimate_never_over():
    m, _ = constants()
    r8 = R >> 232
    assert m * (r8 + 1) < 1 << 40  # q <= x8 / (r8 + 1) <= x / r


def test_estimate_short_by_at_most_one():
    m, _ = constants()
    x8_max = (1 << 29) - 1
    worst, k = 0, 0
    while True:
        # first top limb whose largest value reaches quotient k
        x8 = max(0, -(-(k * R + 1) // (1 << 232)) - 1)
        if x8 > x8_max:
            break
        q_true = (((x8 + 1) << 232) - 1) // R
        worst = max(worst, q_true - ((x8 * m) >> 40))
        k += 1
    q_true = (((x8_max + 1) << 232) - 1) // R
    worst = max(worst, q_true - ((x8_max * m) >> 40))
    assert worst == 1
