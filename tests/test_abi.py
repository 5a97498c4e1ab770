"""CPU-side checks of the C ABI library (no GPU compute).

* libzntt_gpu.so loads and exports every function include/zntt_gpu.h declares;
* context creation without a GPU fails with ZNTT_NO_DEVICE instead of crashing;
* host-only entry points (status strings, element sizes, primitive roots)
  agree with the Python-int fixtures;
* the C oracle's status codes equal the ABI's;
* the host-side BLAKE3 and transcript (no device work) against the
  reference's known answers and the oracle's transcript, and the FRI proof
  sizes against the oracle's.
"""
import json
import os
import re

import pytest

import oracle as O
import pyref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zntt_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zntt_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


def test_library_exports_every_declared_symbol(Z):
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(Z.lib, name), name
    assert sorted(Z.EXPORTED) == decl


def test_status_codes_match_oracle_and_reference_error_names(Z):
    assert Z.LogTooLarge.status == O.LOG_TOO_LARGE
    assert Z.LengthMismatch.status == O.LENGTH_MISMATCH
    assert Z.InvalidLength.status == O.INVALID_LENGTH
    assert Z.InvalidTwiddles.status == O.INVALID_TWIDDLES
    assert Z.OutOfMemory.status == O.OUT_OF_MEMORY
    assert Z.OrderTooLarge.status == O.ORDER_TOO_LARGE
    assert Z.DegreeTooLarge.status == O.DEGREE_TOO_LARGE
    assert Z.InvalidParameters.status == O.FRI_INVALID_PARAMETERS
    for cls in (Z.LogTooLarge, Z.LengthMismatch, Z.InvalidLength, Z.InvalidTwiddles, Z.OutOfMemory,
                Z.OrderTooLarge, Z.DegreeTooLarge, Z.InvalidParameters):
        assert Z.lib.zntt_status_string(cls.status).decode() == cls.__name__


def test_element_sizes(Z):
    assert Z.element_size(Z.FIELD_FR) == 32
    assert Z.element_size(Z.FIELD_GL) == 8
    assert Z.element_size(Z.FIELD_BB) == 8 and Z.element_size(Z.FIELD_KB) == 8
    assert Z.element_size(99) == 0


def test_primitive_roots_host_side(Z):
    c = json.load(open(os.path.join(ROOT, "tests", "golden", "fr_constants.json")))
    for l in (0, 1, 5, 16, 20, 24, 28):
        r = Z.primitive_root(Z.FIELD_FR, l)
        assert O.value(O.FIELD_FR, r) == int(c["primitive_roots"][str(l)], 16)
    for l in (0, 1, 12, 32):
        r = Z.primitive_root(Z.FIELD_GL, l)
        assert int(r[0]) == int(c["goldilocks_primitive_roots"][str(l)], 16)
    with pytest.raises(Z.OrderTooLarge):
        Z.primitive_root(Z.FIELD_FR, 29)
    with pytest.raises(Z.OrderTooLarge):
        Z.primitive_root(Z.FIELD_GL, 33)
    # 31-bit fields: z^((p-1) >> l) with z the smallest non-residue, checked
    # against the Python-int construction and for primitivity.
    for F, two_adic in ((Z.FIELD_BB, 27), (Z.FIELD_KB, 24)):
        p = Z.MODULUS[F]
        for l in (0, 1, 12, two_adic):
            r = int(Z.primitive_root(F, l)[0])
            assert r == pyref.primitive_root(p, l)
            assert pow(r, 1 << l, p) == 1 and (l == 0 or pow(r, 1 << (l - 1), p) == p - 1)
        with pytest.raises(Z.OrderTooLarge):
            Z.primitive_root(F, two_adic + 1)


def test_context_without_gpu_reports_no_device(Z):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(Z.NoDevice):
        Z.Context(0)


def test_library_generator_matches_oracle(Z):
    """The benchmark's input generator (host code in the library) draws the
    same stream as the oracle's and the Python-int restatement."""
    import numpy as np
    import pyref
    for F in (Z.FIELD_FR, Z.FIELD_GL, Z.FIELD_BB, Z.FIELD_KB):
        got = Z.fill_uniform(F, 3, 500)
        assert np.array_equal(got, O.gen_uniform(F, 3, 500))
    c = json.load(open(os.path.join(ROOT, "tests", "golden", "textbook_digests.json")))["cases"]["fr_2p16_seed0"]
    import hashlib
    assert hashlib.sha256(Z.fill_uniform(Z.FIELD_FR, 0, 1 << 16).tobytes()).hexdigest() == c["input_digest"]
    assert [O.value(O.FIELD_FR, Z.fill_uniform(Z.FIELD_FR, 0, 4)[i]) for i in range(4)] == \
        pyref.uniform_elements(pyref.FR, 0, 4)


def test_workload_metadata_generators():
    """The workloads module's xoshiro256++ equals the oracle's stream; the
    uneven batch and arrival schedules are deterministic and in range."""
    from zntt_amd import workloads as W
    g = W.Xoshiro256pp(3)
    assert [g.next() for _ in range(50)] == [int(v) for v in O.xoshiro_stream(3, 50)]
    spec = W.uneven_batch()
    assert len(spec) == 512 and spec == W.uneven_batch()
    assert all(10 <= L <= 20 and d in (0, 1, 2) for L, d in spec)
    assert {d for _, d in spec} == {0, 1, 2} and {L for L, _ in spec} == set(range(10, 21))
    arr = W.serve_arrivals(duration=1.0)
    assert 1800 < len(arr) < 2200
    assert all(12 <= k <= 22 and b in (2, 3) for _, k, b in arr)
    assert all(t0 < t1 for (t0, _, _), (t1, _, _) in zip(arr, arr[1:]))


def test_host_blake3_reference_known_answers(Z):
    k = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_blake3_kat.json")))
    pat = bytes(k["pattern"])
    data = (pat * (4097 // len(pat) + 1))[:4097]
    assert Z.blake3(b"hello world").hex() == k["hello_world"]
    for v in k["vectors"]:
        assert Z.blake3(data[: v["len"]]).hex() == v["digest"], v["len"]


def test_transcript_reference_sequence_and_oracle(Z):
    k = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_fri_kat.json")))["transcript"]
    t = Z.Transcript(k["label"].encode())
    seq = b""
    for _ in range(k["draws"]):  # challengeFieldChecked over Goldilocks
        while True:
            b = t.challenge_bytes(8)
            if int.from_bytes(b, "little") < O.GL_MODULUS:
                break
        seq += b
    assert seq.hex() == k["hex"]
    # A mixed script of absorbs and squeezes of every length class.
    a, b = Z.Transcript(b"script"), O.Transcript(b"script")
    for i, (absorb, squeeze) in enumerate([(0, 8), (3, 32), (1100, 40), (64, 1), (5000, 100), (0, 0)]):
        data = bytes((i * 37 + j) % 251 for j in range(absorb))
        a.absorb_bytes(data)
        b.absorb_bytes(data)
        assert a.challenge_bytes(squeeze) == b.challenge_bytes(squeeze), i


def test_fri_proof_sizes_match_oracle(Z):
    for field in (Z.FIELD_FR, Z.FIELD_GL, Z.FIELD_BB, Z.FIELD_KB):
        for cfg in [(8, 5, 6, 3, 8), (20, 17, 8, 5, 80), (9, 5, 4, 0, 3)]:
            ld, li, lf, lr, q = cfg
            got = Z.fri_proof_sizes(field, ld, li, lf, lr, q)
            want = O.fri_proof_sizes(field, ld, lf, lr, q)
            assert got == {k: want[k] for k in got}, (field, cfg)
    with pytest.raises(Z.InvalidParameters):
        Z.fri_proof_sizes(Z.FIELD_GL, 8, 4, 6, 3, 8)
