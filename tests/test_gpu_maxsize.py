"""GPU parity at the largest sizes the device path takes (ZNTT_MAX_LOG_N = 28,
each field's two-adicity bounding its primitive roots): BN254 Fr and
Goldilocks 2^28, BabyBear 2^27, KoalaBear 2^24; and the sizes between those
and the oracle-compared ones whose pass shapes no other test runs.

At these sizes the oracle is too slow to run, so the checks are ones the
mathematics fixes:
* an impulse e_j at a position j with the top index bit set transforms to
  X[k] = w^(j k): sampled outputs (the first, the middle, the last and random
  ones) are compared with powers computed with Python integers, which pins
  the bit reversal and every pass's indexing up to the last element (BN254 Fr
  scratch offsets pass 4 GiB here);
* a random input (the device counter stream) goes forward and back to
  itself bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FR_R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
MOD = {1: FR_R, 2: 0xFFFFFFFF00000001, 3: 0x78000001, 4: 0x7F000001}
CASES = [(1, 28), (2, 28), (3, 27), (4, 24)]


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


def to_int(field, words):
    """Element value from its storage words (int64 view)."""
    w = [int(x) & ((1 << 64) - 1) for x in words]
    if field == 1:
        m = w[0] | w[1] << 64 | w[2] << 128 | w[3] << 192
        return m * pow(1 << 256, -1, FR_R) % FR_R
    return w[0]


@pytest.mark.parametrize("field,L", CASES, ids=[f"f{f}-2^{L}" for f, L in CASES])
def test_max_size_impulse_and_round_trip(Z, ctx, field, L):
    impulse_and_round_trip(Z, ctx, field, L)


# Sizes between the oracle-compared ones and each field's largest whose pass
# shapes no other test runs.  BN254 Fr: 2^21 and 2^23 (2^10 tiles, three
# passes of 7/7/7 and 8/8/7 stages), 2^25..2^27 (2^11 tiles, three passes,
# the 8- and 9-stage kernels as middle and last passes).  The other fields
# (2^12 tiles, passes of at most 10 stages): 2^23 (8/8/7), 2^25 (9/8/8),
# 2^26 (9/9/8).
SHAPES = [(1, 21), (1, 23), (1, 25), (1, 26), (1, 27), (2, 23), (2, 25), (2, 26), (3, 23), (3, 25), (4, 23)]


@pytest.mark.parametrize("field,L", SHAPES, ids=[f"f{f}-2^{L}" for f, L in SHAPES])
def test_pass_shapes_impulse_and_round_trip(Z, ctx, field, L):
    impulse_and_round_trip(Z, ctx, field, L)


def impulse_and_round_trip(Z, ctx, field, L):
    import torch
    n = 1 << L
    W = 4 if field == 1 else 1
    p = MOD[field]
    sp = torch.cuda.current_stream().cuda_stream
    root = Z.primitive_root(field, L)
    w = to_int(field, root.view(np.int64))
    assert pow(w, n // 2, p) == p - 1  # primitive 2^L-th root
    print(f"max size field {field} 2^{L}: impulse", flush=True)
    t = torch.zeros(n * W, dtype=torch.int64, device="cuda")
    j = (1 << (L - 1)) + 12345
    one = torch.from_numpy(Z.from_int(field, 1).view(np.int64).copy())
    t[j * W:(j + 1) * W] = one.cuda()
    ctx.execute_device(field, Z.FORWARD, L, t.data_ptr(), root, stream=sp)
    torch.cuda.synchronize()
    rng = np.random.default_rng(L * 10 + field)
    ks = [0, 1, 2, 3, n // 2 - 1, n // 2, n // 2 + 1, n - 2, n - 1] + [int(k) for k in rng.integers(0, n, 24)]
    for k in ks:
        got = to_int(field, t[k * W:(k + 1) * W].cpu().numpy())
        assert got == pow(w, j * k % n, p), (field, L, k)
    print(f"max size field {field} 2^{L}: round trip", flush=True)
    Z.fill_counter_device(ctx, field, 7, L, t.data_ptr(), n, stream=sp)
    x0 = t.clone()
    ctx.execute_device(field, Z.FORWARD, L, t.data_ptr(), root, stream=sp)
    ctx.execute_device(field, Z.INVERSE, L, t.data_ptr(), root, stream=sp)
    torch.cuda.synchronize()
    assert not torch.equal(t[:W * 64], torch.zeros_like(t[:W * 64]))
    assert torch.equal(t, x0)
    del t, x0
    torch.cuda.empty_cache()
    ctx.trim()
