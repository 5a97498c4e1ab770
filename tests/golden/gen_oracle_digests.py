#!/usr/bin/env python3
"""Digests of the large workloads, computed by the C oracle.

Run in the build container after tests/test_oracle.py passes (the oracle is
then pinned by the reference's own vectors and the Python-int textbook
digests at 2^16 and 2^20):  python tests/golden/gen_oracle_digests.py

Writes ``oracle_digests.json``: SHA-256 of the storage bytes of
  * BN254 Fr, seed 1, 2^20 / 2^22 / 2^24: input, ntt(root), intt(root) of the
    input (root = primitive_root(log_n));
  * Goldilocks, seed 2, 4096 transforms of 2^12 (one stream, transform order):
    input and forward output;
  * serving: the coset LDE (shift 5) of requests of the
    seed-4 Poisson schedule (2000/s for 10 s), coefficients from the
    counter-based stream (seed 4, stream id = request index): the first
    request of each (k, blowup) with k in {21, 22}, blowup in {4, 8}
    (2^23 .. 2^25 evaluations), and the first two requests of the schedule.
The GPU tests compare the device results with these digests, which needs no
oracle run on the GPU host.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(a.tobytes()).hexdigest()


def serve_section() -> list:
    sys.path.insert(0, os.path.join(ROOT, "zig-algebra_amd"))
    from zntt_amd import workloads as W  # noqa: E402
    arrivals = W.serve_arrivals(rate=2000.0, duration=10.0, seed=4)
    pick = [0, 1]
    for k in (21, 22):
        for b in (2, 3):
            pick.append(next(i for i, (_, kk, bb) in enumerate(arrivals) if kk == k and bb == b))
    F = O.FIELD_FR
    shift = O.from_int(F, 5)
    out = []
    for i in pick:
        t0 = time.time()
        _, k, b = arrivals[i]
        coeffs = O.fill_counter(F, 4, i, 1 << k)
        st, ev = O.lde(F, coeffs, k, b, shift)
        assert st == O.OK
        out.append({"index": i, "seed": 4, "stream_id": i, "log_n": k, "log_blowup": b, "shift": 5,
                    "evaluations": sha(ev), "head": [hex(int(v)) for v in ev[0]]})
        print(f"serve request {i}: k={k} b={b}: {time.time() - t0:.1f}s", flush=True)
    return out


def main() -> None:
    if len(sys.argv) > 1 and sys.argv[1] == "--serve-only":
        path = os.path.join(HERE, "oracle_digests.json")
        with open(path) as f:
            out = json.load(f)
        out["serve_lde"] = serve_section()
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
            f.write("\n")
        print("updated serve_lde in oracle_digests.json")
        return
    out = {"fr": {}, "goldilocks_batch": {}}
    for log_n in (20, 22, 24):
        t0 = time.time()
        F = O.FIELD_FR
        x = O.gen_uniform(F, 1, 1 << log_n)
        root = O.primitive_root(F, log_n)[1]
        fwd = x.copy()
        assert O.ntt(F, fwd, log_n, root) == O.OK
        inv = x.copy()
        assert O.intt(F, inv, log_n, root) == O.OK
        out["fr"][str(log_n)] = {"seed": 1, "input": sha(x), "forward": sha(fwd), "inverse": sha(inv),
                                 "forward_head": [hex(int(v)) for v in fwd[0]],
                                 "inverse_head": [hex(int(v)) for v in inv[0]]}
        print(f"fr 2^{log_n}: {time.time() - t0:.1f}s", flush=True)
    F = O.FIELD_GL
    count, log_n = 4096, 12
    x = O.gen_uniform(F, 2, count << log_n)
    root = O.primitive_root(F, log_n)[1]
    y = x.copy().reshape(count, 1 << log_n)
    for b in range(count):
        row = y[b].copy()
        assert O.ntt(F, row, log_n, root) == O.OK
        y[b] = row
    out["goldilocks_batch"] = {"seed": 2, "count": count, "log_n": log_n, "input": sha(x),
                               "forward": sha(y.reshape(-1))}
    out["serve_lde"] = serve_section()
    with open(os.path.join(HERE, "oracle_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote oracle_digests.json")


if __name__ == "__main__":
    main()
