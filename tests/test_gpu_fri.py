"""GPU tests of the FRI prover (zntt_fri_prove), the LDE's consumer.

* The device proof equals the oracle's byte for byte in every field: roots,
  residual, query indices, opened pairs and Merkle paths.  That holds with
  the library's own transcript and with the transcript routed back to the
  oracle's through the callbacks (the route a Zig shim takes with its own
  zig-transcript object), and with statement bytes absorbed before proving.
* At 2^22 and 2^24 evaluations (beyond the oracle's reach in a test) the
  device proof of honest, degree-bounded data passes the verifier restated
  from verifyOn (tests/fri_verify.py), random data fails it, and the two
  transcript routes agree.
* Config errors come back as ZNTT_INVALID_PARAMETERS before any device work.
"""
import numpy as np
import pytest

import oracle as O
from fri_verify import verify

pytestmark = pytest.mark.gpu

ALL = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]
CONFIGS = [(8, 5, 6, 3, 8), (10, 7, 8, 5, 20), (9, 5, 4, 0, 3), (7, 6, 2, 1, 5), (14, 11, 6, 3, 30)]


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64).reshape(-1).copy()).cuda()


def stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def honest_evals(field, ld, li, seed):
    coeffs = O.gen_uniform(field, seed, 1 << li)
    st, ev = O.lde(field, coeffs, li, ld - li, O.from_int(field, 1))
    assert st == O.OK
    return ev


def oracle_proof(field, label, ev, cfg, statement=b""):
    t = O.Transcript(label)
    if statement:
        t.absorb_bytes(statement)
    st, proof = O.fri_prove(field, t, ev, *cfg)
    assert st == O.OK
    return proof


@pytest.mark.parametrize("field", ALL, ids=IDS)
@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "d%d_i%d_f%d_r%d_q%d" % c)
def test_proof_equals_oracle(Z, ctx, field, cfg):
    ld = cfg[0]
    ev = honest_evals(field, ld, cfg[1], 60 + ld)
    want = oracle_proof(field, b"fri-v2-test", ev, cfg)
    d = dev(ev)
    got = ctx.fri_prove(field, Z.Transcript(b"fri-v2-test"), d.data_ptr(), 1 << ld, *cfg, stream=stream())
    for k in ("roots", "residual", "pair_index", "values", "paths"):
        assert got[k] == want[k], k
    assert verify(field, b"fri-v2-test", got, *cfg)


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_callback_transcript_and_statement(Z, ctx, field):
    """The library calls back into the caller's transcript: here the oracle's,
    carrying statement bytes absorbed before prove (root.zig:883-900)."""
    cfg = (12, 9, 6, 3, 16)
    ev = honest_evals(field, 12, 9, 5)
    want = oracle_proof(field, b"stmt", ev, cfg, statement=b"public input 42")
    t = O.Transcript(b"stmt")
    t.absorb_bytes(b"public input 42")
    d = dev(ev)
    got = ctx.fri_prove(field, t, d.data_ptr(), 1 << 12, *cfg, stream=stream())
    assert got == want
    # The transcript object was advanced exactly as the oracle's prover
    # advances its own: the next squeezes agree.
    t2 = O.Transcript(b"stmt")
    t2.absorb_bytes(b"public input 42")
    O.fri_prove(field, t2, ev, *cfg)
    assert t.challenge_bytes(16) == t2.challenge_bytes(16)


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_library_transcript_rounds_and_state(Z, ctx, field):
    """With the library's own transcript the rounds after the first draw
    their challenges on the device: the proof equals the oracle's, and the
    transcript is left where the oracle's prover leaves its own (statement
    bytes absorbed before prove; the next squeezes agree)."""
    cfg = (16, 12, 6, 2, 8)  # 10 rounds
    ev = honest_evals(field, 16, 12, 17)
    want = oracle_proof(field, b"lib", ev, cfg, statement=b"statement bytes")
    t = Z.Transcript(b"lib")
    t.absorb_bytes(b"statement bytes")
    d = dev(ev)
    got = ctx.fri_prove(field, t, d.data_ptr(), 1 << 16, *cfg, stream=stream())
    for k in ("roots", "residual", "pair_index", "values", "paths"):
        assert got[k] == want[k], k
    t2 = O.Transcript(b"lib")
    t2.absorb_bytes(b"statement bytes")
    O.fri_prove(field, t2, ev, *cfg)
    assert t.challenge_bytes(40) == t2.challenge_bytes(40)


def test_random_data_is_rejected(Z, ctx):
    F = O.FIELD_GL
    cfg = (10, 7, 8, 5, 20)
    ev = O.gen_uniform(F, 77, 1 << 10)
    d = dev(ev)
    got = ctx.fri_prove(F, Z.Transcript(b"fri-v2-test"), d.data_ptr(), 1 << 10, *cfg, stream=stream())
    assert got == oracle_proof(F, b"fri-v2-test", ev, cfg)
    assert not verify(F, b"fri-v2-test", got, *cfg)


@pytest.mark.parametrize("field,ld", [(O.FIELD_FR, 22), (O.FIELD_GL, 24), (O.FIELD_BB, 24)], ids=["fr22", "gl24", "bb24"])
def test_full_size_honest_proof_verifies(Z, ctx, field, ld):
    import torch
    li, lf, lr, nq = ld - 3, 8, 5, 40
    cfg = (ld, li, lf, lr, nq)
    n = 1 << ld
    words = 4 if field == O.FIELD_FR else 1
    ev = torch.zeros(n * words, dtype=torch.int64, device="cuda")
    Z.fill_counter_device(ctx, field, 3, 9, ev.data_ptr(), 1 << li, stream=stream())
    _, root = O.primitive_root(field, ld)
    ctx.execute_device(field, Z.FORWARD, ld, ev.data_ptr(), root, stream=stream())
    torch.cuda.synchronize()
    got = ctx.fri_prove(field, Z.Transcript(b"big"), ev.data_ptr(), n, *cfg, stream=stream())
    assert verify(field, b"big", got, *cfg)
    again = ctx.fri_prove(field, O.Transcript(b"big"), ev.data_ptr(), n, *cfg, stream=stream())
    assert again == got
    noise = torch.empty(n * words, dtype=torch.int64, device="cuda")
    Z.fill_counter_device(ctx, field, 4, 9, noise.data_ptr(), n, stream=stream())
    bad = ctx.fri_prove(field, Z.Transcript(b"big"), noise.data_ptr(), n, *cfg, stream=stream())
    assert not verify(field, b"big", bad, *cfg)


def test_config_errors(Z, ctx):
    d = dev(O.gen_uniform(O.FIELD_GL, 1, 256))
    for cfg in [(8, 5, 8, 3, 8), (8, 5, 6, 6, 8), (8, 5, 6, 3, 0), (8, 4, 6, 3, 8), (8, 5, 13, 3, 8)]:
        with pytest.raises(Z.InvalidParameters):
            ctx.fri_prove(O.FIELD_GL, Z.Transcript(b"t"), d.data_ptr(), 256, *cfg)
    with pytest.raises(Z.InvalidParameters):
        ctx.fri_prove(O.FIELD_GL, Z.Transcript(b"t"), d.data_ptr(), 128, 8, 5, 6, 3, 8)
    # KoalaBear's two-adicity is 24.
    with pytest.raises(Z.InvalidParameters):
        ctx.fri_prove(O.FIELD_KB, Z.Transcript(b"t"), d.data_ptr(), 1 << 25, 25, 22, 6, 3, 8)
