"""Replays an arrival schedule of coset-LDE requests against a Server and
reports latency percentiles beside throughput.

Request i of the schedule (t_i, k_i, b_i) is submitted t_i seconds after the
start: a degree 2^k_i - 1 polynomial, generated on the device from the
counter-based stream (seed, i), extended to 2^(k_i + b_i) evaluations on the
coset shift * <w>.  Latency is measured by the server from submission to the
completion of the request's last pass.
"""
from __future__ import annotations

import time

import numpy as np


def replay(srv, arrivals, shift, seed: int = 4, spin_s: float = 2e-4, keep=()) -> dict:
    """Submits every request of `arrivals` on time, waits for all of them and
    returns the latency distribution, the makespan and the server's stats.
    Requests whose index is in `keep` keep their results on the device; their
    tickets are returned in ``kept`` (index -> ticket) for the caller to read
    and release."""
    keep = set(keep)
    tickets = []
    lag = []
    t0 = time.perf_counter()
    for i, (t, k, b) in enumerate(arrivals):
        due = t0 + t
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            if due - now > spin_s:
                time.sleep(due - now - spin_s)
        lag.append(time.perf_counter() - due)
        tickets.append(srv.submit(k, b, shift, seed=seed, stream_id=i, keep=i in keep))
    srv.drain()
    t_end = time.perf_counter()
    lat, queued, elems = [], [], 0
    kept = {}
    for i, (tk, (_, k, b)) in enumerate(zip(tickets, arrivals)):
        q = srv.query(tk)
        if q["state"] != 2:
            raise RuntimeError(f"request {tk} ended in state {q['state']} (status {q['status']})")
        lat.append(q["latency_ms"])
        queued.append(q["queued_ms"])
        elems += 1 << (k + b)
        if i in keep:
            kept[i] = tk
        else:
            srv.release(tk)
    lat = np.array(lat)
    queued = np.array(queued)
    makespan = t_end - t0
    st = srv.stats()
    return {
        "requests": len(arrivals),
        "evaluations": elems,
        "makespan_s": makespan,
        "melem_s": elems / makespan / 1e6,
        "req_s": len(arrivals) / makespan,
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "max_ms": float(lat.max()) if len(lat) else None,
        "queued_p50_ms": float(np.percentile(queued, 50)) if len(queued) else None,
        "queued_p99_ms": float(np.percentile(queued, 99)) if len(queued) else None,
        "submit_lag_p99_ms": float(np.percentile(np.array(lag) * 1e3, 99)) if lag else None,
        "stats": st,
        "kept": kept,
    }
