"""GPU tests of the batched persistent launch (zntt_batch_*).

Each transform of a batch must equal its oracle result bit for bit, across
graph replays and after the autotuner has changed the launch shape; the full
benchmark workloads are checked against committed digests (the Goldilocks
4096 x 2^12 batch) or against one plain device call per transform (the
uneven BN254 Fr batch), which the single-transform tests pin to the oracle.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def W():
    from zntt_amd import workloads
    return workloads


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


def to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x).view(np.int64)).cuda()


def to_host(t, shape):
    return t.cpu().numpy().view(np.uint64).reshape(shape)


def oracle_apply(field, x, log_n, direction, shift):
    root = O.primitive_root(field, log_n)[1]
    y = x.copy()
    if direction == 0:
        st = O.ntt(field, y, log_n, root)
    elif direction == 1:
        st = O.intt(field, y, log_n, root)
    else:
        st = O.coset_ntt(field, y, log_n, root, shift)
    assert st == O.OK
    return y


@pytest.mark.parametrize("F", [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB], ids=["fr", "gl", "bb", "kb"])
def test_uneven_mixed_batch_matches_oracle(Z, W, ctx, F):
    import torch
    spec = W.uneven_batch(count=24, lo=5, hi=14, seed=33)
    shift = O.from_int(F, 5)
    xs = [Z.fill_uniform(F, W.item_seed(33, i), 1 << L) for i, (L, d) in enumerate(spec)]
    ts = [to_dev(x) for x in xs]
    items = [(d, L, t.data_ptr(), Z.primitive_root(F, L), shift if d == 2 else None)
             for (L, d), t in zip(spec, ts)]
    b = Z.Batch(ctx, F, items)
    expect = [x.copy() for x in xs]
    for rnd in range(3):
        if rnd == 2:
            before = [t.clone() for t in ts]
            b.autotune(reps=1)
            torch.cuda.synchronize()
            for t0, t1 in zip(before, ts):  # tuning runs on shadow buffers
                assert torch.equal(t0, t1)
            print("tuned shape", F, b.shape(), flush=True)
        b.execute(stream=torch.cuda.current_stream().cuda_stream)
        b.status(stream=torch.cuda.current_stream().cuda_stream)
        expect = [oracle_apply(F, e, L, d, shift) for e, (L, d) in zip(expect, spec)]
        for i, ((L, d), t) in enumerate(zip(spec, ts)):
            assert np.array_equal(to_host(t, expect[i].shape), expect[i]), (rnd, i, L, d)
    b.close()


@pytest.mark.parametrize("F", [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB], ids=["fr", "gl", "bb", "kb"])
def test_batch_roots_of_any_order_follow_the_dag(Z, ctx, F):
    """The reference runs its DIT DAG over the powers of whatever root it is
    given (root.zig:65-90 does not check the order).  A batch whose roots are
    arbitrary field elements or roots of unity of smaller order than the size
    must give the oracle's result for the same root: whole-transform launches
    (2^9, 2^12, including Goldilocks 2^12 transforms that share one table),
    multi-pass level launches, every direction."""
    import torch
    arb = O.from_int(F, 123456789)
    arb2 = O.from_int(F, 987654321)
    one = O.from_int(F, 1)

    def sq_root(L):  # a root of order 2^(L-1)
        w = O.primitive_root(F, L)[1]
        return O.field_op(F, "mul", w, w)

    shift = O.from_int(F, 5)
    zero = O.from_int(F, 0)
    spec = [(0, 12, arb), (0, 12, arb), (0, 12, arb), (1, 12, arb), (2, 12, arb),
            (0, 9, arb2), (1, 9, sq_root(9)), (2, 9, one), (0, 10, zero), (2, 11, zero),
            (0, 14, arb), (1, 15, sq_root(15)), (2, 17, arb2), (0, 16, one), (0, 13, zero)]
    xs = [O.gen_uniform(F, 4100 + i, 1 << L) for i, (d, L, r) in enumerate(spec)]
    ts = [to_dev(x) for x in xs]
    items = [(d, L, t.data_ptr(), r, shift if d == 2 else None) for (d, L, r), t in zip(spec, ts)]
    b = Z.Batch(ctx, F, items)
    b.execute(stream=torch.cuda.current_stream().cuda_stream)
    b.status(stream=torch.cuda.current_stream().cuda_stream)
    for i, ((d, L, r), x, t) in enumerate(zip(spec, xs, ts)):
        y = x.copy()
        if d == 0:
            assert O.ntt(F, y, L, r) == O.OK
        elif d == 1:
            assert O.intt(F, y, L, r) == O.OK
        else:
            assert O.coset_ntt(F, y, L, r, shift) == O.OK
        assert np.array_equal(to_host(t, y.shape), y), (i, d, L)
    b.close()


def test_goldilocks_equal_batch_matches_oracle(Z, ctx):
    F = Z.FIELD_GL
    count, L = 64, 12
    x = Z.fill_uniform(F, 77, count << L)
    t = to_dev(x)
    root = Z.primitive_root(F, L)
    stride = (1 << L) * 8
    b = Z.Batch(ctx, F, [(Z.FORWARD, L, t.data_ptr() + i * stride, root) for i in range(count)])
    b.execute()
    b.status()
    got = to_host(t, x.shape).reshape(count, 1 << L)
    for i in range(count):
        row = x[i << L:(i + 1) << L].copy()
        O.ntt(F, row, L, root)
        assert np.array_equal(got[i], row), i
    b.close()


@pytest.mark.parametrize("L", [9, 12])
@pytest.mark.parametrize("power", [1, 3, 5, 7, 2, 4, 0])
@pytest.mark.parametrize("direction", [0, 1, 2], ids=["fwd", "inv", "coset"])
def test_goldilocks_radix8_rounds_every_8th_root(Z, ctx, L, power, direction):
    """Goldilocks 2^9 and 2^12 batches whose transforms share one table run
    as radix-8 rounds whose 8-point steps multiply by the root's 8th root of
    unity as a power of two.  The roots w^1, w^3, w^5, w^7 of the standard
    primitive root w give the four primitive 8th roots (-2^24, -2^72, 2^24,
    2^72).  w^2, w^4 and w^0 = 1 are not primitive (their "8th roots" are a
    4th root, -1 and 1, which the kernel multiplies by as general products,
    and their 2^(L-1)-th powers are not -1); the result must still be the
    radix-2 DAG's over their powers, as the reference computes it for any
    root.  Forward, inverse and coset-forward batches, each transform against
    the oracle."""
    F = Z.FIELD_GL
    count = 6
    x = Z.fill_uniform(F, 91 + power + L, count << L)
    t = to_dev(x)
    w = O.primitive_root(F, L)[1]
    root = O.from_int(F, 1)
    for _ in range(power):
        root = O.field_op(F, "mul", root, w)
    shift = O.from_int(F, 7)
    stride = (1 << L) * 8
    if direction == 2:
        spec = [(Z.COSET_FORWARD, L, t.data_ptr() + i * stride, root, shift) for i in range(count)]
    else:
        spec = [(direction, L, t.data_ptr() + i * stride, root) for i in range(count)]
    b = Z.Batch(ctx, F, spec)
    b.execute()
    b.status()
    got = to_host(t, x.shape).reshape(count, 1 << L)
    for i in range(count):
        row = x[i << L:(i + 1) << L].copy()
        if direction == 0:
            assert O.ntt(F, row, L, root) == O.OK
        elif direction == 1:
            assert O.intt(F, row, L, root) == O.OK
        else:
            assert O.coset_ntt(F, row, L, root, shift) == O.OK
        assert np.array_equal(got[i], row), (i, L, power, direction)
    b.close()


def test_goldilocks_4096x2p12_against_oracle_digest(Z, ctx):
    F = Z.FIELD_GL
    d = json.load(open(os.path.join(GOLD, "oracle_digests.json")))["goldilocks_batch"]
    count, L = d["count"], d["log_n"]
    x = Z.fill_uniform(F, d["seed"], count << L)
    assert hashlib.sha256(x.tobytes()).hexdigest() == d["input"]
    t = to_dev(x)
    root = Z.primitive_root(F, L)
    b = Z.Batch(ctx, F, [(Z.FORWARD, L, t.data_ptr() + i * (8 << L), root) for i in range(count)])
    b.execute()
    b.status()
    assert hashlib.sha256(to_host(t, x.shape).tobytes()).hexdigest() == d["forward"]
    b.close()


def test_large_transforms_batch_matches_oracle_across_shapes(Z, ctx):
    """Transforms of three passes (2^19, 2^20) and two (2^17), mixed
    directions: each equals the oracle's, under the default launch plan and
    again after the autotuner has chosen its shape."""
    import torch
    F = Z.FIELD_FR
    spec = [(19, 0), (20, 1), (17, 2), (19, 2), (20, 0), (17, 1)]
    shift = O.from_int(F, 5)
    xs = [Z.fill_uniform(F, 900 + i, 1 << L) for i, (L, d) in enumerate(spec)]
    ts = [to_dev(x) for x in xs]
    items = [(d, L, t.data_ptr(), Z.primitive_root(F, L), shift if d == 2 else None)
             for (L, d), t in zip(spec, ts)]
    b = Z.Batch(ctx, F, items)
    expect = [x.copy() for x in xs]
    for rnd in range(2):
        if rnd == 1:
            b.autotune(reps=1)
            print("tuned shape", b.shape(), flush=True)
        b.execute(stream=torch.cuda.current_stream().cuda_stream)
        b.status(stream=torch.cuda.current_stream().cuda_stream)
        expect = [oracle_apply(F, e, L, d, shift) for e, (L, d) in zip(expect, spec)]
        for i, ((L, d), t) in enumerate(zip(spec, ts)):
            assert np.array_equal(to_host(t, expect[i].shape), expect[i]), (rnd, i, L, d)
    b.close()


def all_shapes(F):
    """Every shape zntt_batch_set_shape accepts (the autotuner's candidates)."""
    base = 11 if F == O.FIELD_FR else 12
    words = 9 if F == O.FIELD_FR else (2 if F == O.FIELD_GL else 1)
    out = []
    for tl in (base - 1, base):
        for kmax in (tl - 2, tl - 3):
            for w in (1, 2, 3, 4):
                if (words * 4 << tl) * w > 160 * 1024:
                    continue
                for order in (0, 1):
                    for r8 in (0, 1):
                        out.append((tl, kmax, w, order, r8))
    return out


@pytest.mark.parametrize("F", [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB], ids=["fr", "gl", "bb", "kb"])
def test_every_batch_shape_matches_oracle(Z, ctx, F):
    """Each shape the autotuner may choose (tile size, pass length, resident
    workgroups, order, level-kernel family) gives every transform the
    oracle's result: one- to three-pass transforms from 2^3, every direction.  A shape
    outside that set is refused and leaves the batch as it was."""
    import torch
    spec = [(9, 0), (10, 1), (11, 2), (12, 0), (13, 1), (15, 2), (16, 0), (17, 1), (13, 0), (15, 1),
            (3, 0), (4, 2), (6, 1), (8, 0)]
    shift = O.from_int(F, 5)
    xs = [Z.fill_uniform(F, 3300 + i, 1 << L) for i, (L, d) in enumerate(spec)]
    expect = [oracle_apply(F, x, L, d, shift) for x, (L, d) in zip(xs, spec)]
    ts = [to_dev(x) for x in xs]
    items = [(d, L, t.data_ptr(), Z.primitive_root(F, L), shift if d == 2 else None)
             for (L, d), t in zip(spec, ts)]
    b = Z.Batch(ctx, F, items)
    sp = torch.cuda.current_stream().cuda_stream
    shapes = all_shapes(F)
    assert len(shapes) >= 16
    for sh in shapes:
        b.set_shape(*sh)
        assert tuple(b.shape().values()) == sh
        for t, x in zip(ts, xs):
            t.copy_(torch.from_numpy(np.ascontiguousarray(x).view(np.int64)))
        b.execute(stream=sp)
        b.status(stream=sp)
        for i, ((L, d), t) in enumerate(zip(spec, ts)):
            assert np.array_equal(to_host(t, expect[i].shape), expect[i]), (sh, i, L, d)
    before = b.shape()
    base = shapes[0][0]
    for bad in [(base - 2, base - 4, 1, 0, 1), (base, base - 1, 1, 0, 1), (base, base - 2, 5, 0, 1),
                (base, base - 2, 1, 2, 1), (base, base - 2, 1, 0, 2)]:
        with pytest.raises(Z.ZnttError):
            b.set_shape(*bad)
        assert b.shape() == before
    b.close()


def test_uneven_512_batch_equals_per_call(Z, W, ctx):
    """The full uneven workload (512 transforms, 2^10..2^20, mixed
    directions): batched results equal one plain device call per transform."""
    import torch
    F = Z.FIELD_FR
    spec = W.uneven_batch()
    shift = O.from_int(F, 5)
    sp = torch.cuda.current_stream().cuda_stream
    batch_t, single_t, items = [], [], []
    for i, (L, d) in enumerate(spec):
        x = Z.fill_uniform(F, W.item_seed(3, i), 1 << L)
        a, s = to_dev(x), to_dev(x)
        batch_t.append(a)
        single_t.append(s)
        root = Z.primitive_root(F, L)
        items.append((d, L, a.data_ptr(), root, shift if d == 2 else None))
        ctx.execute_device(F, d, L, s.data_ptr(), root, shift if d == 2 else None, stream=sp)
    b = Z.Batch(ctx, F, items)
    b.execute(stream=sp)
    b.status(stream=sp)
    for i, (a, s) in enumerate(zip(batch_t, single_t)):
        assert torch.equal(a, s), (i, spec[i])
    b.close()
