"""Pins for the oracle's Fiat-Shamir transcript and FRI prover (CPU).

* The transcript against the reference's own challenge sequence
  (libs/fri/src/root.zig:1210-1214: label "wire-compat", four Goldilocks
  draws of challengeFieldChecked, produced by an independent build), which
  fixes init's length prefix, the squeeze, the re-key with the full digest
  and the rejection decoder.
* The FRI domain generator against the reference's 256 Goldilocks domain
  elements (root.zig:1272).
* The prover against a verifier restated here in Python integers from
  verifyOn (root.zig:603-718): every Merkle path, every fold along every
  query, and the degree anchor -- the last fold must equal the truncated
  residual's evaluation.  The anchor does not share the fold formula with
  the prover, so a wrong fold (x instead of 1/x, a dropped 1/2, a swapped
  pair) leaves a final layer that is not of low degree and fails it.  The
  verifier also rejects random data and a tampered proof, as the reference's
  own tests require (root.zig:751-880).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from fri_verify import verify

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]


def kat():
    with open(os.path.join(GOLD, "ref_fri_kat.json")) as f:
        return json.load(f)


def test_transcript_reference_sequence():
    k = kat()["transcript"]
    t = O.Transcript(k["label"].encode())
    seq = b"".join(O.value(O.FIELD_GL, t.challenge_field(O.FIELD_GL)).to_bytes(8, "little")
                   for _ in range(k["draws"]))
    assert seq.hex() == k["hex"]


def test_transcript_squeeze_longer_than_a_digest():
    """challengeBytes past 32 bytes re-hashes the digest (root.zig:93-113):
    a 40-byte draw begins with the 32-byte draw of an identical transcript,
    continues with BLAKE3(that digest), and both re-key alike."""
    a, b = O.Transcript(b"x"), O.Transcript(b"x")
    a.absorb_bytes(b"abc")
    b.absorb_bytes(b"abc")
    d32 = a.challenge_bytes(32)
    d40 = b.challenge_bytes(40)
    assert d40[:32] == d32 and d40[32:] == O.blake3(d32)[:8]
    assert a.challenge_bytes(8) != b.challenge_bytes(8)  # re-keyed with different digests


def test_fri_domain_reference_values():
    k = kat()["domain"]
    _, g = O.primitive_root(O.FIELD_GL, k["log_n"])
    g = O.value(O.FIELD_GL, g)
    want = [int(v, 16) for v in k["values"]]
    x = 1
    for i in range(1 << k["log_n"]):
        assert x == want[i], i
        x = x * g % O.GL_MODULUS


def honest_evals(field, ld, li, seed):
    coeffs = O.gen_uniform(field, seed, 1 << li)
    one = O.from_int(field, 1)
    st, ev = O.lde(field, coeffs, li, ld - li, one)
    assert st == O.OK
    return ev


CONFIGS = [(8, 5, 6, 3, 8), (10, 7, 8, 5, 20), (9, 5, 4, 0, 3), (7, 6, 2, 1, 5)]


@pytest.mark.parametrize("field", ALL, ids=IDS)
@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "d%d_i%d_f%d_r%d_q%d" % c)
def test_prove_verifies(field, cfg):
    ld, li, lf, lr, nq = cfg
    ev = honest_evals(field, ld, li, 40 + ld)
    t = O.Transcript(b"fri-v2-test")
    st, proof = O.fri_prove(field, t, ev, ld, li, lf, lr, nq)
    assert st == O.OK
    assert verify(field, b"fri-v2-test", proof, ld, li, lf, lr, nq)
    assert not verify(field, b"another-label", proof, ld, li, lf, lr, nq)


@pytest.mark.parametrize("field", [O.FIELD_GL, O.FIELD_FR], ids=["gl", "fr"])
def test_random_data_and_tampering_rejected(field):
    ld, li, lf, lr, nq = 8, 5, 6, 3, 8
    for trial in range(4):
        ev = O.gen_uniform(field, 1000 + trial, 1 << ld)
        st, proof = O.fri_prove(field, O.Transcript(b"fri-v2-test"), ev, ld, li, lf, lr, nq)
        assert st == O.OK
        assert not verify(field, b"fri-v2-test", proof, ld, li, lf, lr, nq)
    ev = honest_evals(field, ld, li, 7)
    st, proof = O.fri_prove(field, O.Transcript(b"fri-v2-test"), ev, ld, li, lf, lr, nq)
    assert verify(field, b"fri-v2-test", proof, ld, li, lf, lr, nq)
    bad = dict(proof)
    v = bytearray(proof["values"])
    v[0] ^= 1
    bad["values"] = bytes(v)
    assert not verify(field, b"fri-v2-test", bad, ld, li, lf, lr, nq)


def test_config_errors():
    ev = O.gen_uniform(O.FIELD_GL, 1, 256)
    for cfg in [(8, 5, 8, 3, 8), (8, 5, 6, 6, 8), (8, 5, 6, 3, 0), (8, 4, 6, 3, 8), (8, 5, 13, 3, 8)]:
        st, _ = O.fri_prove(O.FIELD_GL, O.Transcript(b"t"), ev, *cfg)
        assert st == O.FRI_INVALID_PARAMETERS, cfg
    st, _ = O.fri_prove(O.FIELD_GL, O.Transcript(b"t"), ev[:128], 8, 5, 6, 3, 8)
    assert st == O.FRI_INVALID_PARAMETERS
