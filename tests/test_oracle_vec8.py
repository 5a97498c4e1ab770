"""Pins for the oracle's eight-lane M31 transform (Vec8NttM31,
libs/field/src/lib.zig:39-137), on the CPU.

* The reference's own test (libs/field/tests/simd_test.zig:155-179): the
  length guard (a short buffer and an absurd log_n are InvalidLength and
  leave the data untouched) and the forward/inverse round trip at
  log_n = 1 with the primitive square root of unity, p - 1.
* Per lane, the transform equals zig-ntt's ntt (libs/ntt/src/root.zig:65-90)
  of that lane, restated separately in the oracle from a different
  reference file, for primitive and non-primitive roots.
* Closed forms: log_n = 1 is (x0 + x1, x0 - x1) for every root; log_n = 2
  written out from the DAG with the root r as a symbol; linearity.
"""
import numpy as np
import pytest

import oracle as O

P = O.M31_MODULUS


def rand(n, seed):
    return np.random.default_rng(seed).integers(0, P, size=n, dtype=np.uint64)


def test_reference_length_guard_and_round_trip():
    log_n, n = 1, 2
    short = np.full(8 * n - 1, 7, dtype=np.uint64)
    assert O.vec8_m31(short, log_n, P - 1) == O.INVALID_LENGTH
    assert O.vec8_m31(short, 62, P - 1) == O.INVALID_LENGTH
    assert O.vec8_m31(short, log_n, P - 1, inverse=True) == O.INVALID_LENGTH
    assert (short == 7).all()
    data = rand(8 * n, 9)
    orig = data.copy()
    assert O.vec8_m31(data, log_n, P - 1) == O.OK
    assert O.vec8_m31(data, log_n, P - 1, inverse=True) == O.OK
    assert np.array_equal(data, orig)


@pytest.mark.parametrize("log_n", [0, 1, 2, 3, 6, 9])
@pytest.mark.parametrize("root", [1, P - 1, 0, 5, 123456789])
def test_each_lane_is_zig_ntt(log_n, root):
    n = 1 << log_n
    data = rand(8 * n, 100 + log_n)
    got = data.copy()
    assert O.vec8_m31(got, log_n, root) == O.OK
    r = np.array([root], dtype=np.uint64)
    for lane in range(8):
        want = np.ascontiguousarray(data[lane::8])
        assert O.ntt(O.FIELD_M31, want, log_n, r) == O.OK
        assert np.array_equal(got[lane::8], want), lane


def test_closed_forms():
    x = rand(8 * 2, 3)
    y = x.copy()
    O.vec8_m31(y, 1, 987654)  # any root: the one stage's twiddle is 1
    for l in range(8):
        a, b = int(x[l]), int(x[8 + l])
        assert int(y[l]) == (a + b) % P and int(y[8 + l]) == (a - b) % P
    r = 31415926
    x = rand(8 * 4, 4)
    y = x.copy()
    O.vec8_m31(y, 2, r)
    for l in range(8):
        x0, x1, x2, x3 = (int(x[8 * i + l]) for i in range(4))
        want = [(x0 + x2 + x1 + x3) % P, (x0 - x2 + r * (x1 - x3)) % P,
                (x0 + x2 - x1 - x3) % P, (x0 - x2 - r * (x1 - x3)) % P]
        assert [int(y[8 * i + l]) for i in range(4)] == want


def test_linearity_and_inverse_scale():
    log_n, n, r = 5, 32, 777
    a, b = rand(8 * n, 5), rand(8 * n, 6)
    c = 424242
    mix = np.array([(c * int(u) + int(v)) % P for u, v in zip(a, b)], dtype=np.uint64)
    for arr in (a, b, mix):
        O.vec8_m31(arr, log_n, r)
    assert all((c * int(u) + int(v)) % P == int(w) for u, v, w in zip(a, b, mix))
    # intt = forward DAG with r^-1, then n^-1 (lib.zig:114-129)
    x = rand(8 * n, 7)
    y, z = x.copy(), x.copy()
    O.vec8_m31(y, log_n, r, inverse=True)
    O.vec8_m31(z, log_n, pow(r, -1, P))
    ninv = pow(n, -1, P)
    assert all(int(u) == int(v) * ninv % P for u, v in zip(y, z))
