"""GPU tests of the serving loop (zntt_server_*) and the counter-based
coefficient stream.

Every served LDE must equal the oracle's LDE of the same coefficients bit for
bit, whether the request ran alone or in a mix of concurrent requests of
different sizes, blowups and shifts, under a budget that forces queueing, and
with caller-provided coefficients and outputs as well as generated ones.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


def to_host(t, field, n):
    a = t.cpu().numpy().view(np.uint64)
    return a.reshape(n, 4) if field == O.FIELD_FR else a.reshape(n)


def expected(field, seed, sid, k, b, shift):
    coeffs = O.fill_counter(field, seed, sid, 1 << k)
    st, out = O.lde(field, coeffs, k, b, shift)
    assert st == O.OK
    return out


ALL = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_device_counter_stream_matches_oracle(Z, ctx, field):
    import torch
    words = 4 if field == O.FIELD_FR else 1
    for seed, sid, n in ((4, 0, 1000), (4, 12345, 4096), (7, 1 << 33, 777)):
        t = torch.empty(n * words, dtype=torch.int64, device="cuda")
        Z.fill_counter_device(ctx, field, seed, sid, t.data_ptr(), n,
                              stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(t, field, n), O.fill_counter(field, seed, sid, n)), (seed, sid, n)


def mixed_requests(field, count, seed):
    rng = np.random.default_rng(seed)
    shifts = [O.from_int(field, 5), O.from_int(field, 7), O.from_int(field, 123456789)]
    out = []
    for i in range(count):
        k = int(rng.integers(3, 13))
        b = int(rng.integers(1, 4))
        out.append((k, b, int(rng.integers(0, 3)), 1000 + i))
    return out, shifts


# The Goldilocks case once made a GPU memory fault when it ran after the batch
# tests in one process.  The mistake found since: zntt_ctx_trim and
# zntt_ctx_destroy freed twiddle tables that live batches and servers had
# pinned, and a batch or server destroyed after its context wrote into the
# freed context and cache entries (tests/test_host_lifetime.py shows each
# order under AddressSanitizer).  Children now hold a reference on their
# context and trims skip pinned entries.
MIXED = ALL


@pytest.mark.parametrize("field", MIXED, ids=IDS)
def test_mixed_requests_equal_alone_and_oracle(Z, ctx, field):
    """40 requests of 2^3..2^12 coefficients, blowup 2..8, three shifts, all
    submitted at once (they share steps) and then one at a time: both runs
    equal the oracle bit for bit."""
    reqs, shifts = mixed_requests(field, 40, seed=11 + field)
    srv = Z.Server(ctx, field, hbm_budget=1 << 30, max_log_n=15)
    mix = [srv.submit(k, b, shifts[si], seed=4, stream_id=sid, keep=True) for k, b, si, sid in reqs]
    srv.drain()
    st = srv.stats()
    assert st["completed"] == 40 and st["failed"] == 0
    assert st["steps"] < 40 * 2  # requests shared steps
    alone = []
    for k, b, si, sid in reqs:
        t = srv.submit(k, b, shifts[si], seed=4, stream_id=sid, keep=True)
        srv.wait(t)
        alone.append(t)
    for (k, b, si, sid), tm, ta in zip(reqs, mix, alone):
        want = expected(field, 4, sid, k, b, shifts[si])
        got_mix, got_alone = srv.read(tm), srv.read(ta)
        assert np.array_equal(got_mix, want), (k, b, si, sid)
        assert np.array_equal(got_alone, got_mix), (k, b, si, sid)
        q = srv.query(tm)
        assert q["state"] == Z.REQ_DONE and q["latency_ms"] >= q["queued_ms"] >= 0
        srv.release(tm)
        srv.release(ta)
    assert srv.stats()["bytes_in_use"] == 0
    srv.close()


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_padding_deeper_than_the_skipped_stages(Z, ctx, field):
    """The level kernel skips the products of zero padding in at most the
    first three stages of a request's first pass (blowup up to 8); a
    blowup of 2, 16 or 32 pads 1, 4 or 5 stages.  Multi-pass requests with
    those blowups, served together, equal the oracle's LDE bit for bit."""
    shifts = [O.from_int(field, 5), O.from_int(field, 7)]
    reqs = [(10, 4, 0, 601), (11, 5, 1, 602), (12, 1, 0, 603), (9, 4, 1, 604), (13, 3, 0, 605)]
    srv = Z.Server(ctx, field, hbm_budget=1 << 30, max_log_n=16)
    ts = [srv.submit(k, b, shifts[si], seed=4, stream_id=sid, keep=True) for k, b, si, sid in reqs]
    srv.drain()
    for (k, b, si, sid), t in zip(reqs, ts):
        assert np.array_equal(srv.read(t), expected(field, 4, sid, k, b, shifts[si])), (k, b)
        srv.release(t)
    srv.close()


def test_budget_forces_queueing_results_exact(Z, ctx):
    """A budget that holds two 2^14-evaluation requests at a time."""
    F = O.FIELD_FR
    per = (32 + 36) * (1 << 14) + 32 * (1 << 12)
    srv = Z.Server(ctx, F, hbm_budget=2 * per + 1024, max_log_n=15)
    sh = O.from_int(F, 5)
    ts = [srv.submit(12, 2, sh, seed=4, stream_id=i, keep=True) for i in range(8)]
    got = []
    for t in ts:
        srv.wait(t)
        got.append(srv.read(t))
        srv.release(t)
    st = srv.stats()
    assert st["completed"] == 8 and st["peak_bytes"] <= 2 * per + 1024
    for i, g in enumerate(got):
        assert np.array_equal(g, expected(F, 4, i, 12, 2, sh)), i
    srv.close()


def test_caller_buffers_and_larger_request(Z, ctx):
    """Caller-owned coefficients (2^16, written by the device generator) and
    caller-owned output (2^19), beside a generated small request."""
    import torch
    F = O.FIELD_FR
    k, b = 16, 3
    coeffs = torch.empty((1 << k) * 4, dtype=torch.int64, device="cuda")
    out = torch.zeros((1 << (k + b)) * 4, dtype=torch.int64, device="cuda")
    Z.fill_counter_device(ctx, F, 9, 1, coeffs.data_ptr(), 1 << k,
                          stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    sh = O.from_int(F, 5)
    srv = Z.Server(ctx, F, max_log_n=20)
    t_big = srv.submit(k, b, sh, coeffs_ptr=coeffs.data_ptr(), out_ptr=out.data_ptr())
    t_small = srv.submit(8, 2, sh, seed=9, stream_id=2, keep=True)
    srv.drain()
    assert srv.query(t_big)["data"] == out.data_ptr()
    assert srv.stats()["bytes_in_use"] == (32 << 10)  # only the kept small result
    assert np.array_equal(to_host(out, F, 1 << (k + b)), expected(F, 9, 1, k, b, sh))
    assert np.array_equal(srv.read(t_small), expected(F, 9, 2, 8, 2, sh))
    srv.close()


def test_server_matches_per_call_lde(Z, ctx):
    """The served LDE equals the library's per-call zntt_lde (host buffers)."""
    F = O.FIELD_FR
    sh = O.from_int(F, 5)
    srv = Z.Server(ctx, F, max_log_n=18)
    t = srv.submit(14, 3, sh, seed=4, stream_id=77, keep=True)
    srv.wait(t)
    got = srv.read(t)
    coeffs = O.fill_counter(F, 4, 77, 1 << 14)
    assert np.array_equal(got, ctx.lde(F, coeffs, 14, 3, sh))
    srv.close()


def test_large_requests_run_alone_and_match_oracle(Z, ctx):
    """BN254 Fr requests of 2^20 points and more run each pass as a launch of
    the single-transform pass kernel, whose first pass zero-pads the 2^k
    coefficients; small requests share the same steps' level launches.
    Every result equals the oracle's LDE, in the mix and alone."""
    F = O.FIELD_FR
    shifts = [O.from_int(F, 5), O.from_int(F, 7)]
    reqs = [(18, 2, 0, 501), (17, 3, 1, 502), (10, 2, 0, 503), (20, 1, 1, 504), (12, 3, 1, 505), (19, 2, 0, 506)]
    srv = Z.Server(ctx, F, hbm_budget=1 << 31, max_log_n=21)
    mix = [srv.submit(k, b, shifts[si], seed=4, stream_id=sid, keep=True) for k, b, si, sid in reqs]
    got = []
    for t in mix:
        srv.wait(t)
        got.append(srv.read(t))
    for (k, b, si, sid), g in zip(reqs, got):
        assert np.array_equal(g, expected(F, 4, sid, k, b, shifts[si])), (k, b, si, sid)
    k, b, si, sid = reqs[0]
    t = srv.submit(k, b, shifts[si], seed=4, stream_id=sid, keep=True)
    srv.wait(t)
    assert np.array_equal(srv.read(t), got[0])
    srv.close()


@pytest.mark.parametrize("field", ALL[1:], ids=IDS[1:])
def test_large_requests_of_the_small_fields_match_oracle(Z, ctx, field):
    """Goldilocks, BabyBear and KoalaBear requests of 2^18 .. 2^22
    evaluations (two and three passes, the zero-padding first pass with one
    to three padded stages) served together with small ones: each equals the
    oracle's LDE."""
    shifts = [O.from_int(field, 5), O.from_int(field, 7)]
    reqs = [(18, 3, 0, 701), (19, 2, 1, 702), (17, 1, 0, 703), (20, 2, 1, 704), (11, 3, 0, 705), (9, 2, 1, 706)]
    srv = Z.Server(ctx, field, hbm_budget=1 << 31, max_log_n=22)
    ts = [srv.submit(k, b, shifts[si], seed=4, stream_id=sid, keep=True) for k, b, si, sid in reqs]
    srv.drain()
    for (k, b, si, sid), t in zip(reqs, ts):
        assert np.array_equal(srv.read(t), expected(field, 4, sid, k, b, shifts[si])), (k, b, si, sid)
        srv.release(t)
    srv.close()


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_smallest_requests_and_request_errors(Z, ctx, field):
    """The smallest LDEs the server takes (2^3 evaluations from one, two or
    four coefficients, and one coefficient blown up 16-fold) against the
    oracle, served together; and the request errors, raised at submit:
    fewer than 2^3 evaluations, a coefficient count that is not 2^log_n,
    sizes past the field's two-adicity or 2^64."""
    shift = O.from_int(field, 5)
    reqs = [(0, 3, 801), (1, 2, 802), (2, 1, 803), (0, 4, 804), (3, 0, 805)]
    srv = Z.Server(ctx, field, hbm_budget=1 << 28, max_log_n=12)
    ts = [srv.submit(k, b, shift, seed=4, stream_id=sid, keep=True) for k, b, sid in reqs]
    srv.drain()
    for (k, b, sid), t in zip(reqs, ts):
        assert np.array_equal(srv.read(t), expected(field, 4, sid, k, b, shift)), (k, b)
        srv.release(t)
    with pytest.raises(Z.Unsupported):
        srv.submit(1, 1, shift, seed=4)
    with pytest.raises(Z.LengthMismatch):
        srv.submit(4, 2, shift, seed=4, length=15)
    deep = {O.FIELD_FR: 29, O.FIELD_GL: 33, O.FIELD_BB: 28, O.FIELD_KB: 25}[field]
    with pytest.raises(Z.OrderTooLarge):
        srv.submit(0, deep, shift, seed=4, length=1)
    with pytest.raises(Z.LogTooLarge):
        srv.submit(60, 4, shift, seed=4, length=1)
    assert srv.stats()["bytes_in_use"] == 0
    srv.close()


def test_config5_size_requests_match_oracle_digests(Z, ctx):
    """Requests of the serving benchmark's own schedule at its largest sizes
    (k 21 and 22, blowup 4 and 8: 2^23 .. 2^25 evaluations, the zero-padding
    first pass with K = 8 and 9) and two of its first requests, all submitted
    at once beside small generated requests: each result equals the oracle's
    LDE digest (tests/golden/oracle_digests.json, "serve_lde")."""
    import hashlib
    import json
    import os
    F = O.FIELD_FR
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_digests.json")
    with open(path) as f:
        cases = json.load(f)["serve_lde"]
    assert {(c["log_n"], c["log_blowup"]) for c in cases} >= {(21, 2), (21, 3), (22, 2), (22, 3)}
    sh = O.from_int(F, 5)
    srv = Z.Server(ctx, F, hbm_budget=16 << 30, max_log_n=25)
    tickets = []
    for j, c in enumerate(cases):
        tickets.append(srv.submit(c["log_n"], c["log_blowup"], sh, seed=c["seed"], stream_id=c["stream_id"],
                                  keep=True))
        srv.submit(9 + j % 4, 2 + j % 2, sh, seed=5, stream_id=j)  # small neighbours, results not kept
    for c, t in zip(cases, tickets):
        srv.wait(t)
        got = srv.read(t)
        assert hex(int(got[0][0])) == c["head"][0], c["index"]
        assert hashlib.sha256(memoryview(got)).hexdigest() == c["evaluations"], (c["index"], c["log_n"])
        srv.release(t)
        del got
    srv.drain()
    assert srv.stats()["failed"] == 0
    srv.close()
