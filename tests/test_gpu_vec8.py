"""GPU tests of the eight-lane M31 transform (zntt_ntt_vec8_m31,
zntt_intt_vec8_m31, zntt_vec8_m31_device) against the oracle's restatement
of Vec8NttM31 (libs/field/src/lib.zig:52-137), bit for bit.

Sizes cover one-pass and multi-pass transforms (tiles of 2^12 values per
lane); roots cover the primitive square root (p - 1), 1, 0 (the inverse's
zero-root case) and non-primitive values; the length guards are the
reference's own (libs/field/tests/simd_test.zig:155-168).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
P = O.M31_MODULUS


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


def rand(n, seed):
    return np.random.default_rng(seed).integers(0, P, size=n, dtype=np.uint64)


@pytest.mark.parametrize("log_n", [0, 1, 2, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 17, 20])
@pytest.mark.parametrize("inverse", [False, True], ids=["fwd", "inv"])
def test_host_matches_oracle(Z, ctx, log_n, inverse):
    n = 1 << log_n
    for root in (P - 1, 1, 0, 1234567 + log_n):
        data = rand(8 * n, log_n * 10 + root % 7)
        want = data.copy()
        assert O.vec8_m31(want, log_n, root, inverse=inverse) == O.OK
        got = data.copy()
        ctx.ntt_vec8_m31(got, log_n, root, inverse=inverse)
        assert np.array_equal(got, want), (log_n, root)


@pytest.mark.parametrize("log_n", [3, 14, 21, 22])
def test_device_matches_host_and_inverse_matches_oracle(Z, ctx, log_n):
    import torch
    n = 1 << log_n
    data = rand(8 * n, 77)
    want = data.copy()
    ctx.ntt_vec8_m31(want, log_n, P - 1)
    t = torch.from_numpy(data.view(np.int64).copy()).cuda()
    ctx.vec8_m31_device(t.data_ptr(), log_n, P - 1, stream=torch.cuda.current_stream().cuda_stream)
    assert np.array_equal(t.cpu().numpy().view(np.uint64), want)
    inv = data.copy()
    ctx.ntt_vec8_m31(inv, log_n, 5, inverse=True)
    ref = data.copy()
    O.vec8_m31(ref, log_n, 5, inverse=True)
    assert np.array_equal(inv, ref)


def test_length_guard(Z, ctx):
    short = np.full(8 * 2 - 1, 7, dtype=np.uint64)
    with pytest.raises(Z.InvalidLength):
        ctx.ntt_vec8_m31(short, 1, P - 1)
    with pytest.raises(Z.InvalidLength):
        ctx.ntt_vec8_m31(short, 62, P - 1)
    with pytest.raises(Z.InvalidLength):
        ctx.ntt_vec8_m31(short, 1, P - 1, inverse=True)
    assert (short == 7).all()
    big = np.zeros(8, dtype=np.uint64)
    with pytest.raises(Z.Unsupported):
        ctx.ntt_vec8_m31(big, 26, P - 1, length=8 << 26)
