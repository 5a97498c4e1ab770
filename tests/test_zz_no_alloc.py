"""GPU test of zntt_execute_device_noalloc, the entry point that allocates
nothing: its results equal zntt_execute_device's bit for bit, on two calls in
a row (the second smaller), for every field and direction, with multi-pass
sizes that use the static device scratch; it allocates no device memory (the
library's and torch's accounting stay put) and refuses oversized and
unprepared transforms without touching the data."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (field, [(direction, log_n), (direction, log_n)]): the second call smaller.
CASES = [
    (1, [(0, 18), (1, 12)]),   # BN254 Fr: 3 then 2 passes
    (1, [(2, 20), (0, 14)]),   # coset-forward at the largest size that fits the scratch
    (2, [(2, 22), (1, 13)]),   # Goldilocks
    (3, [(0, 20), (2, 14)]),   # BabyBear
    (4, [(1, 19), (0, 13)]),   # KoalaBear
]


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


def words(field):
    return 4 if field == 1 else 1


@pytest.mark.parametrize("field,calls", CASES, ids=[f"f{f}-{c[0][1]}-{c[1][1]}" for f, c in CASES])
def test_noalloc_equals_execute_device(Z, field, calls):
    import torch
    ctx = Z.Context(0)
    try:
        sp = torch.cuda.current_stream().cuda_stream
        shift = Z.from_int(field, 7)
        bufs = []
        for i, (d, L) in enumerate(calls):
            x = Z.fill_uniform(field, 100 + i, 1 << L)
            a = torch.from_numpy(x.view(np.int64).reshape(-1)).cuda()
            b = a.clone()
            root = Z.primitive_root(field, L)
            # The reference result, which also puts the tables in the cache.
            ctx.execute_device(field, d, L, a.data_ptr(), root, shift if d == 2 else None, stream=sp)
            bufs.append((d, L, root, a, b))
        torch.cuda.synchronize()
        lib_bytes, torch_bytes = ctx.bytes_in_use(), torch.cuda.memory_allocated()
        for d, L, root, a, b in bufs:  # two calls in a row, the second smaller
            print(f"noalloc field {field} direction {d} log_n {L}", flush=True)
            ctx.execute_device_noalloc(field, d, L, b.data_ptr(), root, shift if d == 2 else None, stream=sp)
        torch.cuda.synchronize()
        for d, L, root, a, b in bufs:
            assert torch.equal(a, b), (field, d, L)
        assert ctx.bytes_in_use() == lib_bytes
        assert torch.cuda.memory_allocated() == torch_bytes
    finally:
        ctx.close()


def test_noalloc_refuses_before_touching_data(Z):
    import torch
    ctx = Z.Context(0)
    try:
        sp = torch.cuda.current_stream().cuda_stream
        F = Z.FIELD_FR
        # 2^21 BN254 Fr needs 72 MiB of scratch, beyond the static 64 MiB.
        t = torch.zeros((1 << 21) * 4, dtype=torch.int64, device="cuda")
        r = Z.primitive_root(F, 21)
        ctx.execute_device(F, Z.FORWARD, 21, t.data_ptr(), r, stream=sp)
        torch.cuda.synchronize()
        before = t.clone()
        with pytest.raises(Z.OutOfMemory):
            ctx.execute_device_noalloc(F, Z.FORWARD, 21, t.data_ptr(), r, stream=sp)
        # A size whose table was never built in this context.
        with pytest.raises(Z.NotPrepared):
            ctx.execute_device_noalloc(F, Z.FORWARD, 13, t.data_ptr(), Z.from_int(F, 3), stream=sp)
        torch.cuda.synchronize()
        assert torch.equal(t, before)
    finally:
        ctx.close()
