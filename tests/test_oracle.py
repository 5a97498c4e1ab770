"""Pins for the CPU oracle (oracle/zntt_oracle.c), run on the CPU.

Each check is chosen so that a plausible mistake in the restatement (a dropped
carry, a w <-> w^-1 flip, a transposed index, a wrong Montgomery constant)
fails at least one of them:

* the reference's own known-answer DFT vectors, libs/ntt/src/root.zig:438-449;
* the reference's own small-field vectors, libs/field/src/lib.zig:257-276;
* Python-int values from definitions (tests/golden/pyref.py): Montgomery
  arithmetic for BN254 Fr, primitive roots, O(n^2) DFT, Horner coset/LDE, and
  digests of a recursive textbook NTT of the seeded workloads;
* the reference's invariant tests restated: bit-reversal (root.zig:290-306),
  round trips (:370-397), convolution (:539-603), twiddle entries (:605-636),
  twiddled == plain (:638-673), error contracts (:675-738).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import pyref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = {"fr": O.FIELD_FR, "goldilocks": O.FIELD_GL, "babybear": O.FIELD_BB}


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def elems(field, ints):
    """Storage-form array from canonical integers."""
    return O.from_ints(field, ints)


def canon(field, arr):
    return [O.value(field, arr[i]) for i in range(arr.shape[0])]


# --------------------------------------------------------------------------
# Field arithmetic
# --------------------------------------------------------------------------

def test_fr_montgomery_constants():
    c = load("fr_constants.json")
    one = O.from_int(O.FIELD_FR, 1)
    assert O.to_int(O.FIELD_FR, one) == int(c["r_mod_p"], 16)
    # to_mont(x) == x * R mod p for a value with every limb populated
    x = (1 << 250) + 12345
    assert O.to_int(O.FIELD_FR, O.from_int(O.FIELD_FR, x)) == pyref.to_mont(x)


def test_fr_arithmetic_matches_python_ints():
    F = O.FIELD_FR
    for v in load("fr_field_vectors.json")["vectors"]:
        a, b = O.from_int(F, int(v["a"], 16)), O.from_int(F, int(v["b"], 16))
        assert O.to_int(F, a) == int(v["a_mont"], 16)
        assert O.value(F, O.field_op(F, "add", a, b)) == int(v["sum"], 16)
        assert O.value(F, O.field_op(F, "sub", a, b)) == int(v["diff"], 16)
        assert O.value(F, O.field_op(F, "mul", a, b)) == int(v["prod"], 16)
        assert O.value(F, O.field_op(F, "neg", b)) == int(v["neg_b"], 16)
        assert O.value(F, O.field_op(F, "inv", a)) == int(v["inv_a"], 16)


@pytest.mark.parametrize("name", ["goldilocks", "babybear"])
def test_small_field_reference_vectors(name):
    F = FIELDS[name]
    for v in load("ref_field_vectors.json")["vectors"][name]:
        a, b = O.from_int(F, int(v["a"], 16)), O.from_int(F, int(v["b"], 16))
        assert O.value(F, O.field_op(F, "add", a, b)) == int(v["sum"], 16)
        assert O.value(F, O.field_op(F, "mul", a, b)) == int(v["prod"], 16)
        assert O.value(F, O.field_op(F, "inv", a)) == int(v["inv_a"], 16)
        assert O.value(F, O.field_op(F, "neg", b)) == int(v["neg_b"], 16)


def test_inverse_of_zero_is_zero():
    for F in FIELDS.values():
        z = O.from_int(F, 0)
        assert O.value(F, O.field_op(F, "inv", z)) == 0


# --------------------------------------------------------------------------
# Roots of unity
# --------------------------------------------------------------------------

def test_primitive_roots_match_python():
    c = load("fr_constants.json")
    for l in range(0, 29):
        st, r = O.primitive_root(O.FIELD_FR, l)
        assert st == O.OK
        assert O.value(O.FIELD_FR, r) == int(c["primitive_roots"][str(l)], 16)
    assert O.primitive_root(O.FIELD_FR, 29)[0] == O.ORDER_TOO_LARGE
    for l in range(0, 33):
        st, r = O.primitive_root(O.FIELD_GL, l)
        assert st == O.OK and O.value(O.FIELD_GL, r) == int(c["goldilocks_primitive_roots"][str(l)], 16)
    assert O.primitive_root(O.FIELD_GL, 33)[0] == O.ORDER_TOO_LARGE


def test_fr_roots_are_primitive():
    p = pyref.FR
    for l in (1, 16, 20, 24, 28):
        w = O.value(O.FIELD_FR, O.primitive_root(O.FIELD_FR, l)[1])
        assert pow(w, 1 << l, p) == 1
        assert pow(w, 1 << (l - 1), p) == p - 1


def test_reference_kat_roots_are_the_library_roots():
    for c in load("ref_ntt_kat.json")["cases"]:
        F = FIELDS[c["field"]]
        assert O.value(F, O.primitive_root(F, c["log_n"])[1]) == int(c["root"], 16)


# --------------------------------------------------------------------------
# Transforms
# --------------------------------------------------------------------------

def test_reference_dft_known_answers():
    """libs/ntt/src/root.zig:405-453, with the reference's own numbers."""
    for c in load("ref_ntt_kat.json")["cases"]:
        F = FIELDS[c["field"]]
        xs = [int(x, 16) for x in c["input"]]
        data = elems(F, xs)
        root = O.from_int(F, int(c["root"], 16))
        assert O.ntt(F, data, c["log_n"], root) == O.OK
        assert canon(F, data) == [int(x, 16) for x in c["expect"]]
        assert O.intt(F, data, c["log_n"], root) == O.OK
        assert canon(F, data) == xs


def test_kat_catches_a_root_convention_flip():
    """A w <-> w^-1 flip in the oracle would fail the known answers."""
    c = load("ref_ntt_kat.json")["cases"][3]  # Goldilocks log_n = 4
    F = FIELDS[c["field"]]
    data = elems(F, [int(x, 16) for x in c["input"]])
    root_inv = O.field_op(F, "inv", O.from_int(F, int(c["root"], 16)))
    O.ntt(F, data, c["log_n"], root_inv)
    assert canon(F, data) != [int(x, 16) for x in c["expect"]]


def test_dft_and_coset_vectors():
    d = load("dft_vectors.json")
    for c in d["dft"]:
        F = FIELDS[c["field"]]
        xs = [int(x, 16) for x in c["input"]]
        root = O.from_int(F, int(c["root"], 16))
        data = elems(F, xs)
        assert O.ntt(F, data, c["log_n"], root) == O.OK
        assert canon(F, data) == [int(x, 16) for x in c["dft"]], (c["field"], c["log_n"])
        data = elems(F, xs)
        shift = O.from_int(F, int(c["coset_shift"], 16))
        assert O.coset_ntt(F, data, c["log_n"], root, shift) == O.OK
        assert canon(F, data) == [int(x, 16) for x in c["coset"]], (c["field"], c["log_n"])
    for c in d["lde"]:
        F = FIELDS[c["field"]]
        coeffs = elems(F, [int(x, 16) for x in c["coeffs"]])
        shift = O.from_int(F, int(c["coset_shift"], 16))
        st, out = O.lde(F, coeffs, c["log_k"], c["log_blowup"], shift)
        assert st == O.OK
        assert canon(F, out) == [int(x, 16) for x in c["evals"]]


def test_bit_reverse_contracts():
    F = O.FIELD_GL
    a = elems(F, [0, 1, 2, 3, 4, 5, 6, 0])
    orig = a.copy()
    assert O.bit_reverse(F, a) == O.OK
    assert O.bit_reverse(F, a) == O.OK
    assert np.array_equal(a, orig)
    b = elems(F, [0, 1, 2, 3])
    O.bit_reverse(F, b)
    assert canon(F, b) == [0, 2, 1, 3]
    assert O.bit_reverse(F, O.empty(F, 0)) == O.INVALID_LENGTH
    assert O.bit_reverse(F, O.empty(F, 3)) == O.INVALID_LENGTH
    assert O.bit_reverse(F, O.empty(F, 4)) == O.OK


def test_error_contracts():
    """libs/ntt/src/root.zig:675-738."""
    for F in FIELDS.values():
        root = O.from_int(F, 3)
        short = O.empty(F, 3)
        assert O.ntt(F, short, 2, root) == O.LENGTH_MISMATCH
        assert O.intt(F, short, 2, root) == O.LENGTH_MISMATCH
        two = O.empty(F, 2)
        assert O.ntt(F, two, 64, O.from_int(F, 1)) == O.LOG_TOO_LARGE
        assert O.intt(F, two, 200, O.from_int(F, 1)) == O.LOG_TOO_LARGE
        assert O.precompute_twiddles(F, 64, O.from_int(F, 1))[0] == O.LOG_TOO_LARGE
    F = O.FIELD_GL
    log_n = 3
    root = O.primitive_root(F, log_n)[1]
    st, tw = O.precompute_twiddles(F, log_n, root)
    assert st == O.OK
    data = elems(F, [1] * 8)
    assert O.ntt_with_twiddles(F, data, log_n, tw[:log_n - 1]) == O.INVALID_TWIDDLES
    assert O.intt_with_twiddles(F, data, log_n, tw[:log_n - 1]) == O.INVALID_TWIDDLES
    bad = [tw[0][:1], tw[1][:1], tw[2]]
    assert O.ntt_with_twiddles(F, data.copy(), log_n, bad) == O.INVALID_TWIDDLES
    copy = data.copy()
    assert O.ntt_with_twiddles(F, copy, log_n, tw) == O.OK
    assert O.intt_with_twiddles(F, copy, log_n, tw) == O.OK
    assert np.array_equal(copy, data)


@pytest.mark.parametrize("name", ["fr", "goldilocks", "babybear"])
def test_twiddle_entries_are_powers(name):
    """libs/ntt/src/root.zig:605-636 against Python pow."""
    F = FIELDS[name]
    p = O.MODULUS[F]
    for log_n in range(1, 7):
        root = O.primitive_root(F, log_n)[1]
        w = O.value(F, root)
        st, tw = O.precompute_twiddles(F, log_n, root)
        assert st == O.OK and len(tw) == log_n
        for s in range(log_n):
            wm = pow(w, 1 << (log_n - s - 1), p)
            assert canon(F, tw[s]) == [pow(wm, j, p) for j in range(1 << s)]


@pytest.mark.parametrize("name", ["fr", "goldilocks", "babybear"])
def test_round_trips_and_twiddled_equals_plain(name):
    F = FIELDS[name]
    for log_n in range(0, 11):
        data = O.gen_uniform(F, 100 + log_n, 1 << log_n)
        orig = data.copy()
        root = O.primitive_root(F, log_n)[1]
        plain = data.copy()
        assert O.ntt(F, plain, log_n, root) == O.OK
        st, tw = O.precompute_twiddles(F, log_n, root)
        tw_out = data.copy()
        assert O.ntt_with_twiddles(F, tw_out, log_n, tw) == O.OK
        assert np.array_equal(plain, tw_out)
        assert O.intt(F, plain, log_n, root) == O.OK
        assert np.array_equal(plain, orig)
        assert O.intt_with_twiddles(F, tw_out, log_n, tw) == O.OK
        assert np.array_equal(tw_out, orig)


@pytest.mark.parametrize("name", ["fr", "goldilocks"])
def test_convolution_property(name):
    """libs/ntt/src/root.zig:539-603: NTT(f (*) g) == NTT(f) . NTT(g)."""
    F = FIELDS[name]
    p = O.MODULUS[F]
    log_n, n = 4, 16
    f = [1, 2, 3] + [0] * 13
    g = [2, 3] + [0] * 14
    conv = [sum(f[j] * g[(i - j) % n] for j in range(n)) % p for i in range(n)]
    root = O.primitive_root(F, log_n)[1]
    fa, ga = elems(F, f), elems(F, g)
    O.ntt(F, fa, log_n, root)
    O.ntt(F, ga, log_n, root)
    prod = O.empty(F, n)
    for i in range(n):
        prod[i] = O.field_op(F, "mul", fa[i], ga[i])
    O.intt(F, prod, log_n, root)
    assert canon(F, prod) == conv


def test_non_primitive_root_follows_the_dit_dag():
    """For a root that is not primitive the reference's DAG is the only
    definition; pin the oracle against a Python restatement of the loop
    (libs/ntt/src/root.zig:65-90) on the bit-reversed input."""
    F = O.FIELD_FR
    p = pyref.FR
    log_n, n = 5, 32
    xs = [(i * 7919 + 3) % p for i in range(n)]
    w = 123456789  # not a root of unity at all
    rev = [int(format(i, f"0{log_n}b")[::-1], 2) for i in range(n)]
    d = [xs[rev[i]] for i in range(n)]
    for s in range(1, log_n + 1):
        m, h = 1 << s, 1 << (s - 1)
        wm = pow(w, 1 << (log_n - s), p)
        for k in range(0, n, m):
            ww = 1
            for j in range(h):
                t = ww * d[k + j + h] % p
                u = d[k + j]
                d[k + j], d[k + j + h] = (u + t) % p, (u - t) % p
                ww = ww * wm % p
    data = elems(F, xs)
    O.ntt(F, data, log_n, O.from_int(F, w))
    assert canon(F, data) == d


# --------------------------------------------------------------------------
# Input generator and large-size digests
# --------------------------------------------------------------------------

def test_generator_matches_python():
    g = pyref.Xoshiro256pp(42)
    assert [int(x) for x in O.xoshiro_stream(42, 64)] == [g.next() for _ in range(64)]
    for name in ("fr", "goldilocks", "babybear"):
        F = FIELDS[name]
        p = O.MODULUS[F]
        ref = pyref.uniform_elements(p, 7, 300)
        got = O.gen_uniform(F, 7, 300)
        assert canon(F, got) == ref


def test_textbook_digests_2p16():
    import hashlib
    c = load("textbook_digests.json")["cases"]["fr_2p16_seed0"]
    F = O.FIELD_FR
    data = O.gen_uniform(F, 0, 1 << 16)
    assert hashlib.sha256(data.tobytes()).hexdigest() == c["input_digest"]
    orig = data.copy()
    root = O.primitive_root(F, 16)[1]
    assert O.ntt(F, data, 16, root) == O.OK
    assert hashlib.sha256(data.tobytes()).hexdigest() == c["forward_digest"]
    assert O.intt(F, data, 16, root) == O.OK
    assert np.array_equal(data, orig)


def test_textbook_digests_2p20():
    import hashlib
    c = load("textbook_digests.json")["cases"]["fr_2p20_seed1"]
    F = O.FIELD_FR
    data = O.gen_uniform(F, 1, 1 << 20)
    assert hashlib.sha256(data.tobytes()).hexdigest() == c["input_digest"]
    root = O.primitive_root(F, 20)[1]
    st, tw = O.precompute_twiddles(F, 20, root)
    assert O.ntt_with_twiddles(F, data, 20, tw) == O.OK
    assert hashlib.sha256(data.tobytes()).hexdigest() == c["forward_digest"]


# ---------------------------------------------------------------------------
# Polynomial product (libs/poly/src/poly.zig:197-224), pinned to Python ints
# ---------------------------------------------------------------------------

def schoolbook(p, a, b):
    """Python-int product of coefficient lists, after removing trailing zeros."""
    while a and a[-1] == 0:
        a = a[:-1]
    while b and b[-1] == 0:
        b = b[:-1]
    if not a or not b:
        return []
    r = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] = (r[i + j] + x * y) % p
    return r


@pytest.mark.parametrize("name", ["fr", "goldilocks", "babybear"])
def test_poly_mul_matches_python_schoolbook(name):
    F = FIELDS[name]
    p = O.MODULUS[F]
    rng = np.random.default_rng(5)
    cases = [(0, 0), (1, 0), (1, 1), (3, 5), (17, 1), (33, 31)]
    for la, lb in cases:
        a = [int(v) for v in rng.integers(0, 1 << 62, la)]
        b = [int(v) for v in rng.integers(0, 1 << 62, lb)]
        if la > 2:
            a[-1] = 0  # a trailing zero is not part of the degree
        st, got = O.poly_mul(F, elems(F, a), elems(F, b))
        assert st == O.OK
        assert canon(F, got) == schoolbook(p, [x % p for x in a], [x % p for x in b]), (la, lb)


def test_poly_mul_degree_too_large_and_zero():
    F = O.FIELD_GL
    a = elems(F, [1, 2, 3])  # degree 2
    b = elems(F, [4, 5])     # degree 1
    st, got = O.poly_mul(F, a, b, cap=3)
    assert st == O.DEGREE_TOO_LARGE and got.shape[0] == 0
    st, got = O.poly_mul(F, a, b, cap=4)
    assert st == O.OK and canon(F, got) == [4, 13, 22, 15]
    st, got = O.poly_mul(F, a, elems(F, [0, 0]), cap=0)
    assert st == O.OK and got.shape[0] == 0  # zero polynomial wins over capacity


def test_poly_eval_matches_python_and_product_identity():
    for F in (O.FIELD_FR, O.FIELD_GL):
        p = O.MODULUS[F]
        a = O.gen_uniform(F, 41, 50)
        b = O.gen_uniform(F, 42, 70)
        z = 123456789123456789 % p
        ai, bi = canon(F, a), canon(F, b)
        assert O.value(F, O.poly_eval(F, a, O.from_int(F, z))) == sum(c * pow(z, i, p) for i, c in enumerate(ai)) % p
        st, r = O.poly_mul(F, a, b)
        assert st == O.OK
        lhs = O.value(F, O.poly_eval(F, r, O.from_int(F, z)))
        rhs = (sum(c * pow(z, i, p) for i, c in enumerate(ai)) * sum(c * pow(z, i, p) for i, c in enumerate(bi))) % p
        assert lhs == rhs


def test_koalabear_small_field_by_definition():
    """KoalaBear (libs/field/src/predef/koalabear.zig) has no reference vectors
    on the ntt path; the generic small field is pinned by Python ints: field
    ops, the primitive roots, and a DFT against its O(n^2) definition."""
    F, p = O.FIELD_KB, O.KB_MODULUS
    assert O.lib().oracle_elem_size(F) == 8
    xs = [0, 1, 2, p - 1, 123456789, (p - 1) // 2]
    for a in xs:
        for b in xs:
            ea, eb = O.from_int(F, a), O.from_int(F, b)
            assert O.value(F, O.field_op(F, "mul", ea, eb)) == a * b % p
            assert O.value(F, O.field_op(F, "add", ea, eb)) == (a + b) % p
            assert O.value(F, O.field_op(F, "sub", ea, eb)) == (a - b) % p
    for l in (1, 5, 24):
        assert O.value(F, O.primitive_root(F, l)[1]) == pyref.primitive_root(p, l)
    x = [int(v) % p for v in np.random.default_rng(3).integers(0, 1 << 62, 16)]
    w = pyref.primitive_root(p, 4)
    data = O.from_ints(F, x)
    assert O.ntt(F, data, 4, O.from_int(F, w)) == O.OK
    assert canon(F, data) == pyref.dft(x, w, p)
