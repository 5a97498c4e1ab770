"""Serving loop scheduling on the CPU (no GPU): the server's own scheduling
core driven step by step in simulation mode (zntt_server_sim_pump /
zntt_server_sim_complete), with an injectable allocator.

Covers admission against the HBM budget and the active-request limit, first
come first served order, continuous batching (a request submitted while
others are in flight joins the next step), request validation errors,
allocation failure, oversize requests, cancellation (queued and in flight),
kept results and memory balance.  Also pins the counter-based coefficient
stream of generated requests (oracle and a Python-int restatement).
"""
import numpy as np
import pytest

import oracle as O
import pyref


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


def shift(Z):
    return Z.from_int(Z.FIELD_FR, 5)


def req_bytes(k, b, generated=True, npass=None):
    """Device bytes of one Fr request: output (32 B/elem), scratch (36 B/elem,
    multi-pass only) and generated input (32 B/coefficient)."""
    L = k + b
    n = 1 << L
    if npass is None:
        npass = 1 if L <= 11 else -(-L // 9)
    return 32 * n + (36 * n if npass > 1 else 0) + (32 * (1 << k) if generated else 0)


class CountingAllocator:
    """Fake device allocator: hands out distinct addresses, can be told to fail."""

    def __init__(self):
        self.live = {}
        self.next = 1 << 20
        self.fail_next = 0
        self.fail_with = MemoryError
        self.calls = 0

    def alloc(self, n):
        self.calls += 1
        if self.fail_next:
            self.fail_next -= 1
            raise self.fail_with("injected")
        a = self.next
        self.next += (n + 255) // 256 * 256
        self.live[a] = n
        return a

    def free(self, a, n):
        assert self.live.pop(a) == n


def run_to_completion(srv, max_iter=1000):
    for _ in range(max_iter):
        n = srv.sim_pump()
        st = srv.stats()
        if n == 0 and st["queued"] == 0 and st["active"] == 0:
            return
        srv.sim_complete(n)
    raise AssertionError("server did not drain")


def test_counter_stream_matches_python_restatement():
    for field, p in ((O.FIELD_FR, pyref.FR), (O.FIELD_GL, pyref.GL), (O.FIELD_BB, O.BB_MODULUS),
                     (O.FIELD_KB, O.KB_MODULUS)):
        for seed, sid in ((4, 0), (4, 17), (0, 1 << 40), (123456789, 3)):
            got = O.fill_counter(field, seed, sid, 40)
            want = pyref.counter_stored(p, seed, sid, 40)
            if field == O.FIELD_FR:
                vals = [sum(int(got[i, w]) << (64 * w) for w in range(4)) for i in range(40)]
            else:
                vals = [int(v) for v in got]
            assert vals == want


def test_counter_stream_range_and_independence():
    x = O.fill_counter(O.FIELD_FR, 4, 0, 4096)
    top = x[:, 3]
    assert int(top.max()) <= 0x30644E72E131A029          # below r
    assert len({tuple(r) for r in x.tolist()}) == 4096
    y = O.fill_counter(O.FIELD_FR, 4, 1, 4096)
    assert not np.array_equal(x, y)
    # Prefix property: element i does not depend on n.
    assert np.array_equal(O.fill_counter(O.FIELD_FR, 4, 0, 100), x[:100])
    g = O.fill_counter(O.FIELD_GL, 9, 2, 4096)
    assert int(g.max()) < pyref.GL
    # Rough uniformity of the top bits (a skew would show here).
    hist = np.bincount((g >> np.uint64(61)).astype(np.int64), minlength=8)
    assert hist.min() > 4096 / 8 * 0.8


def test_validation_errors(Z):
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 30, simulate=True)
    sh = shift(Z)
    with pytest.raises(Z.LogTooLarge):
        srv.submit(64, 1, sh)
    with pytest.raises(Z.LogTooLarge):
        srv.submit(40, 30, sh)
    with pytest.raises(Z.LengthMismatch):
        srv.submit(10, 2, sh, length=1000)
    with pytest.raises(Z.OrderTooLarge):
        srv.submit(20, 9, sh)
    with pytest.raises(Z.Unsupported):
        srv.submit(1, 1, sh)
    assert srv.stats()["submitted"] == 0
    with pytest.raises(Z.InvalidArgument):
        srv.drain()  # a simulation has nothing that would run the steps
    srv.close()
    with pytest.raises(Z.InvalidArgument):
        Z.Server(None, Z.FIELD_FR)  # a real server needs a context


def test_budget_admission_fifo_and_peak(Z):
    k, b = 12, 2
    per = req_bytes(k, b)
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=2 * per + per // 2, simulate=True)
    t = [srv.submit(k, b, shift(Z), stream_id=i) for i in range(5)]
    assert srv.sim_pump() == 1
    st = srv.stats()
    assert st["admitted"] == 2 and st["queued"] == 3 and st["bytes_in_use"] == 2 * per
    assert [srv.query(x)["state"] for x in t] == [Z.REQ_RUNNING] * 2 + [Z.REQ_QUEUED] * 3
    # Both admitted requests need two passes: nothing retires after one step.
    assert srv.sim_pump() == 2
    srv.sim_complete(1)
    srv.sim_pump()
    assert srv.stats()["admitted"] == 2
    srv.sim_complete(1)
    srv.sim_pump()
    st = srv.stats()
    assert st["completed"] == 2 and st["admitted"] == 4
    assert [srv.query(x)["state"] for x in t[:4]] == [Z.REQ_DONE] * 2 + [Z.REQ_RUNNING] * 2
    run_to_completion(srv)
    st = srv.stats()
    assert st["completed"] == 5 and st["peak_bytes"] == 2 * per and st["bytes_in_use"] == 0
    assert st["elements_done"] == 5 << (k + b)
    for x in t:
        q = srv.query(x)
        assert q["latency_ms"] >= q["queued_ms"] >= 0
        srv.release(x)
    srv.close()


def test_active_limit(Z):
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 34, max_active=3, simulate=True)
    for i in range(7):
        srv.submit(10, 2, shift(Z), stream_id=i)
    srv.sim_pump()
    assert srv.stats()["active"] == 3
    run_to_completion(srv)
    assert srv.stats()["completed"] == 7
    srv.close()


def test_continuous_batching_joins_next_step(Z):
    """A request arriving while a large one is between passes is admitted and
    runs its first pass in the large one's next step, finishing first."""
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 36, simulate=True)
    big = srv.submit(20, 3, shift(Z))  # 2^23: three passes
    assert srv.sim_pump() == 1
    assert srv.sim_pump() == 2  # second step (big's pass 1) enqueued ahead
    small = srv.submit(6, 2, shift(Z))  # 2^8: one pass
    srv.sim_complete(1)
    srv.sim_pump()  # retires step 0, admits `small`, launches step 2 = big pass 2 + small pass 0
    assert srv.query(small)["state"] == Z.REQ_RUNNING
    st = srv.stats()
    assert st["steps"] == 3 and st["active"] == 2
    srv.sim_complete(2)
    srv.sim_pump()
    assert srv.query(small)["state"] == Z.REQ_DONE
    assert srv.query(big)["state"] == Z.REQ_DONE
    assert srv.stats()["steps"] == 3
    srv.close()


def test_oversize_request_fails_others_proceed(Z):
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 24, simulate=True)
    a = srv.submit(10, 2, shift(Z))
    huge = srv.submit(20, 3, shift(Z))
    c = srv.submit(10, 2, shift(Z))
    run_to_completion(srv)
    assert srv.query(huge)["state"] == Z.REQ_FAILED
    assert srv.query(huge)["status"] == Z.OutOfMemory.status
    assert srv.query(a)["state"] == srv.query(c)["state"] == Z.REQ_DONE
    srv.close()


def test_allocation_failure_waits_for_memory_or_fails(Z):
    al = CountingAllocator()
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 34, simulate=True, allocator=al)
    a = srv.submit(12, 2, shift(Z))
    srv.sim_pump()
    assert srv.query(a)["state"] == Z.REQ_RUNNING
    # With a request in flight, an allocation failure means "wait".
    al.fail_next = 1
    b = srv.submit(12, 2, shift(Z))
    srv.sim_pump()
    assert srv.query(b)["state"] == Z.REQ_QUEUED
    run_to_completion(srv)
    assert srv.query(b)["state"] == Z.REQ_DONE
    # With nothing in flight it is final.
    al.fail_next = 1
    c = srv.submit(12, 2, shift(Z))
    srv.sim_pump()
    q = srv.query(c)
    assert q["state"] == Z.REQ_FAILED and q["status"] == Z.OutOfMemory.status
    # Other allocator errors fail the request at once.
    al.fail_next, al.fail_with = 1, RuntimeError
    d = srv.submit(12, 2, shift(Z))
    srv.sim_pump()
    assert srv.query(d)["status"] == Z.DeviceError.status
    e = srv.submit(12, 2, shift(Z))
    run_to_completion(srv)
    assert srv.query(e)["state"] == Z.REQ_DONE
    srv.close()
    assert al.live == {}  # every buffer was returned


def test_cancellation(Z):
    al = CountingAllocator()
    per = req_bytes(12, 2)
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=per, simulate=True, allocator=al)
    a = srv.submit(12, 2, shift(Z))
    b = srv.submit(12, 2, shift(Z))
    srv.sim_pump()
    srv.cancel(b)  # queued: dropped at once, never admitted
    assert srv.query(b)["state"] == Z.REQ_CANCELLED
    srv.cancel(a)  # in flight: runs out its passes, then frees everything
    assert srv.query(a)["state"] == Z.REQ_RUNNING
    run_to_completion(srv)
    assert srv.query(a)["state"] == Z.REQ_CANCELLED
    st = srv.stats()
    assert st["cancelled"] == 2 and st["admitted"] == 1 and st["bytes_in_use"] == 0
    assert al.live == {}
    srv.close()


def test_kept_results_hold_memory_until_release(Z):
    al = CountingAllocator()
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 30, simulate=True, allocator=al)
    t = srv.submit(12, 2, shift(Z), keep=True)
    with pytest.raises(Z.InvalidArgument):
        srv.release(t)  # not finished
    run_to_completion(srv)
    q = srv.query(t)
    assert q["state"] == Z.REQ_DONE and q["data"] != 0
    assert srv.stats()["bytes_in_use"] == 32 << 14
    assert len(al.live) == 1
    srv.release(t)
    assert srv.stats()["bytes_in_use"] == 0 and al.live == {}
    with pytest.raises(Z.InvalidArgument):
        srv.query(t)
    srv.close()


def test_caller_coefficients_need_no_input_buffer(Z):
    srv = Z.Server(None, Z.FIELD_FR, hbm_budget=1 << 30, simulate=True)
    srv.submit(12, 2, shift(Z), coeffs_ptr=0x1000)
    srv.sim_pump()
    assert srv.stats()["bytes_in_use"] == req_bytes(12, 2, generated=False)
    srv.close()
