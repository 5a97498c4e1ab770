"""GPU tests of the NTT polynomial product (zntt_poly_mul*), the widening of
libs/poly's schoolbook Polynomial.mul (poly.zig:197-212).

Small and medium products equal the oracle's schoolbook product bit for bit
(degrees around every power of two, trailing zeros, zero and constant
operands, the DegreeTooLarge contract); large products are checked by the
evaluation identity (a*b)(z) = a(z) b(z) at random points, and the device
entry point against the host one.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_poly_mul_matches_schoolbook(Z, ctx, field):
    sizes = [(1, 1), (1, 7), (2, 2), (3, 6), (16, 17), (33, 32), (64, 65), (100, 29), (255, 258), (513, 511)]
    for k, (la, lb) in enumerate(sizes):
        a = O.gen_uniform(field, 300 + k, la)
        b = O.gen_uniform(field, 400 + k, lb)
        st, want = O.poly_mul(field, a, b)
        assert st == O.OK
        got = ctx.poly_mul(field, a, b)
        assert np.array_equal(got, want), (la, lb)


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_poly_mul_edge_contracts(Z, ctx, field):
    a = O.gen_uniform(field, 1, 10)
    zero = O.empty(field, 5)
    assert ctx.poly_mul(field, a, zero).shape[0] == 0           # poly.zig:198
    assert ctx.poly_mul(field, zero, a, cap=0).shape[0] == 0
    with pytest.raises(Z.DegreeTooLarge):                        # poly.zig:199-200
        ctx.poly_mul(field, a, a, cap=18)
    assert ctx.poly_mul(field, a, a, cap=19).shape[0] == 19
    # Trailing zeros are not part of the degree.
    padded = np.concatenate([a, O.empty(field, 7)])
    st, want = O.poly_mul(field, a, a)
    assert np.array_equal(ctx.poly_mul(field, padded, a, cap=19), want)
    c1, c2 = O.gen_uniform(field, 2, 1), O.gen_uniform(field, 3, 1)
    st, want = O.poly_mul(field, c1, c2)
    assert np.array_equal(ctx.poly_mul(field, c1, c2), want)


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_large_poly_mul_by_evaluation(Z, ctx, field):
    """2^18 x 2^18 coefficients (a 2^19-point product): (a*b)(z) = a(z) b(z)
    at three points, and the device entry point equals the host one."""
    import torch
    n = 1 << 18
    a = O.gen_uniform(field, 61, n)
    b = O.gen_uniform(field, 62, n - 5)
    r = ctx.poly_mul(field, a, b)
    assert r.shape[0] == 2 * n - 6
    p = O.MODULUS[field]
    for z in (3, 123456789, p - 2):
        zz = O.from_int(field, z)
        lhs = O.value(field, O.poly_eval(field, r, zz))
        rhs = O.value(field, O.poly_eval(field, a, zz)) * O.value(field, O.poly_eval(field, b, zz)) % p
        assert lhs == rhs, z
    ta = torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()
    tb = torch.from_numpy(np.ascontiguousarray(b).view(np.int64)).cuda()
    words = 4 if field == O.FIELD_FR else 1
    tout = torch.zeros((2 * n - 6) * words, dtype=torch.int64, device="cuda")
    ctx.poly_mul_device(field, ta.data_ptr(), n, tb.data_ptr(), n - 5, tout.data_ptr(),
                        stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = tout.cpu().numpy().view(np.uint64).reshape(r.shape)
    assert np.array_equal(got, r)


@pytest.mark.parametrize("field", [O.FIELD_FR, O.FIELD_GL], ids=["fr", "gl"])
def test_poly_mul_of_lengths_off_the_copy_slots(Z, ctx, field):
    """Host buffers of lengths that end in a partial 16 MiB slot of the
    page-locked copy ring, both ways (BN254 Fr: 600,001 and 700,003
    coefficients in, 1,300,003 out; Goldilocks four times as many):
    (a*b)(z) = a(z) b(z) at three points."""
    scale = 1 if field == O.FIELD_FR else 4
    na, nb = 600001 * scale, 700003 * scale
    a = O.gen_uniform(field, 71, na)
    b = O.gen_uniform(field, 72, nb)
    r = ctx.poly_mul(field, a, b)
    assert r.shape[0] == na + nb - 1
    p = O.MODULUS[field]
    for z in (5, 987654321, p - 3):
        zz = O.from_int(field, z)
        lhs = O.value(field, O.poly_eval(field, r, zz))
        rhs = O.value(field, O.poly_eval(field, a, zz)) * O.value(field, O.poly_eval(field, b, zz)) % p
        assert lhs == rhs, z

