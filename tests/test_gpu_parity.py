"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Bit-exact comparison (the path is integer arithmetic) on seeded inputs at
sizes the oracle finishes in seconds, against committed digests at the full
workload sizes (BN254 Fr 2^20 / 2^22 / 2^24), and through round trips.  Edge
cases follow the reference's tests (libs/ntt/src/root.zig:290-738): sizes 2^0
and 2^1, every pass split, error contracts, non-primitive roots and foreign
(non-power) twiddle tables, which only the exact DIT DAG reproduces.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
import pyref

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]
FNAME = dict(zip(FIELDS, IDS))
# Smallest generator of each multiplicative group: the coset shifts.
GENERATOR = {O.FIELD_FR: 5, O.FIELD_GL: 7, O.FIELD_BB: 31, O.FIELD_KB: 3}


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def Z():
    import zntt_amd
    return zntt_amd


@pytest.fixture(scope="module")
def ctx(Z):
    c = Z.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------------------
# Device field arithmetic
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_device_field_ops_match_oracle(Z, ctx, field):
    say("field ops", FNAME[field])
    n = 4096
    a = O.gen_uniform(field, 11, n)
    b = O.gen_uniform(field, 12, n)
    p = O.MODULUS[field]
    edge = [0, 1, 2, p - 1, p - 2, (p - 1) // 2]
    ea = O.from_ints(field, [x for x in edge for _ in edge])
    eb = O.from_ints(field, [y for _ in edge for y in edge])
    a = np.concatenate([a, ea])
    b = np.concatenate([b, eb])
    for op, name in ((0, "add"), (1, "sub"), (2, "mul")):
        got = ctx.debug_field_op(field, op, a, b)
        exp = np.stack([O.field_op(field, name, a[i], b[i]) for i in range(a.shape[0])])
        assert np.array_equal(got.reshape(exp.shape), exp), name


# ---------------------------------------------------------------------------
# ntt / intt against the oracle at every pass split
# ---------------------------------------------------------------------------

SIZES = list(range(0, 20))


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
@pytest.mark.parametrize("log_n", SIZES)
def test_ntt_intt_match_oracle(Z, ctx, field, log_n):
    say("ntt/intt", FNAME[field], log_n)
    n = 1 << log_n
    x = O.gen_uniform(field, 1000 + log_n, n)
    root = O.primitive_root(field, log_n)[1]
    exp = x.copy()
    assert O.ntt(field, exp, log_n, root) == O.OK
    got = x.copy()
    ctx.ntt(field, got, log_n, root)
    assert np.array_equal(got, exp)
    exp_i = x.copy()
    assert O.intt(field, exp_i, log_n, root) == O.OK
    got_i = x.copy()
    ctx.intt(field, got_i, log_n, root)
    assert np.array_equal(got_i, exp_i)
    ctx.intt(field, got, log_n, root)
    assert np.array_equal(got, x)


# The 32- and 64-bit fields' pass shapes from 2^20 (three passes) that the
# sizes above do not reach: 2^20 (10/10 stages, the second pass last),
# 2^21 and 2^22 (7/7/7 and 8/7/7 stages).
BIG = [(f, L) for f in (O.FIELD_GL, O.FIELD_BB, O.FIELD_KB) for L in (20, 21, 22)]


@pytest.mark.parametrize("field,log_n", BIG, ids=[f"{FNAME[f]}-2^{L}" for f, L in BIG])
def test_three_pass_shapes_match_oracle(Z, ctx, field, log_n):
    test_ntt_intt_match_oracle(Z, ctx, field, log_n)


def test_reference_known_answers_on_device(Z, ctx):
    """libs/ntt/src/root.zig:438-449 (the Goldilocks and BabyBear rows)
    through the GPU."""
    fid = {"goldilocks": Z.FIELD_GL, "babybear": Z.FIELD_BB}
    seen = set()
    for c in load("ref_ntt_kat.json")["cases"]:
        F = fid[c["field"]]
        seen.add(F)
        xs = np.array([int(v, 16) for v in c["input"]], dtype=np.uint64)
        root = np.array([int(c["root"], 16)], dtype=np.uint64)
        data = xs.copy()
        ctx.ntt(F, data, c["log_n"], root)
        assert [int(v) for v in data] == [int(v, 16) for v in c["expect"]], (c["field"], c["log_n"])
        ctx.intt(F, data, c["log_n"], root)
        assert np.array_equal(data, xs)
    assert seen == {Z.FIELD_GL, Z.FIELD_BB}


def test_fr_dft_vectors_on_device(Z, ctx):
    for c in load("dft_vectors.json")["dft"]:
        F = {"fr": O.FIELD_FR, "goldilocks": O.FIELD_GL}.get(c["field"])
        if F is None:
            continue
        data = O.from_ints(F, [int(v, 16) for v in c["input"]])
        root = O.from_int(F, int(c["root"], 16))
        ctx.ntt(F, data, c["log_n"], root)
        assert [O.value(F, data[i]) for i in range(data.shape[0])] == [int(v, 16) for v in c["dft"]]


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_non_primitive_root_follows_the_dag(Z, ctx, field):
    """For a root that is not a primitive root the reference's DIT DAG is the
    only definition; the device must reproduce it exactly."""
    for log_n in (5, 12, 15, 20):
        x = O.gen_uniform(field, 77 + log_n, 1 << log_n)
        for r in (123456789, 0, 1):  # any element; 0 and 1 make every twiddle past j = 0 zero or one
            root = O.from_int(field, r)
            exp = x.copy()
            O.ntt(field, exp, log_n, root)
            got = x.copy()
            ctx.ntt(field, got, log_n, root)
            assert np.array_equal(got, exp), (log_n, r)
            exp = x.copy()
            O.intt(field, exp, log_n, root)  # inv(0) = 0, as the reference returns (field.zig:384-386, montgomery.zig:386-388)
            got = x.copy()
            ctx.intt(field, got, log_n, root)
            assert np.array_equal(got, exp), (log_n, r, "intt")


# ---------------------------------------------------------------------------
# Twiddled variants
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("field", FIELDS, ids=IDS)
@pytest.mark.parametrize("log_n", [1, 3, 10, 13, 16])
def test_precompute_and_twiddled_transforms(Z, ctx, field, log_n):
    say("twiddled", FNAME[field], log_n)
    root = O.primitive_root(field, log_n)[1]
    tw = ctx.precompute_twiddles(field, log_n, root)
    st, otw = O.precompute_twiddles(field, log_n, root)
    assert st == O.OK and len(tw) == log_n
    for s in range(log_n):
        assert np.array_equal(tw[s], otw[s]), s
    x = O.gen_uniform(field, 5000 + log_n, 1 << log_n)
    exp = x.copy()
    O.ntt_with_twiddles(field, exp, log_n, otw)
    got = x.copy()
    ctx.ntt_with_twiddles(field, got, log_n, tw)
    assert np.array_equal(got, exp)
    exp_i = exp.copy()
    O.intt_with_twiddles(field, exp_i, log_n, otw)
    ctx.intt_with_twiddles(field, got, log_n, tw)
    assert np.array_equal(got, exp_i)
    assert np.array_equal(got, x)
    tw.free()


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_foreign_non_power_tables(Z, ctx, field):
    """A caller's table that is not powers of anything (root.zig:148-212 uses
    it as given): forward and the -tw[h-j] inverse must match the oracle."""
    for log_n in (4, 12, 14):
        stages = [O.gen_uniform(field, 900 + s, 1 << s) for s in range(log_n)]
        x = O.gen_uniform(field, 800 + log_n, 1 << log_n)
        for inverse in (False, True):
            exp = x.copy()
            O.ntt_with_twiddles(field, exp, log_n, stages, inverse=inverse)
            got = x.copy()
            (ctx.intt_with_twiddles if inverse else ctx.ntt_with_twiddles)(field, got, log_n, stages)
            assert np.array_equal(got, exp), (log_n, inverse)
    # Entry 0 of every stage is 1 but the rest is arbitrary: the device may
    # then skip the products by entry 0 of the first stages, and only those.
    one = O.from_int(field, 1)
    for log_n in (4, 11, 12, 14):
        stages = [O.gen_uniform(field, 700 + s, 1 << s) for s in range(log_n)]
        for st in stages:
            st[0] = one if field == O.FIELD_FR else one[0]
        x = O.gen_uniform(field, 600 + log_n, 1 << log_n)
        exp = x.copy()
        O.ntt_with_twiddles(field, exp, log_n, stages)
        got = x.copy()
        ctx.ntt_with_twiddles(field, got, log_n, stages)
        assert np.array_equal(got, exp), ("entry0=1", log_n)


# ---------------------------------------------------------------------------
# Error contracts (libs/ntt/src/root.zig:675-738)
# ---------------------------------------------------------------------------

def test_error_contracts(Z, ctx):
    for F in FIELDS:
        root = O.from_int(F, 3)
        short = O.empty(F, 3)
        with pytest.raises(Z.LengthMismatch):
            ctx.ntt(F, short, 2, root)
        with pytest.raises(Z.LengthMismatch):
            ctx.intt(F, short, 2, root)
        two = O.empty(F, 2)
        with pytest.raises(Z.LogTooLarge):
            ctx.ntt(F, two, 64, root)
        with pytest.raises(Z.LogTooLarge):
            ctx.intt(F, two, 200, root)
        with pytest.raises(Z.LogTooLarge):
            ctx.precompute_twiddles(F, 64, root)
        with pytest.raises(Z.InvalidLength):
            ctx.bit_reverse(F, O.empty(F, 0))
        with pytest.raises(Z.InvalidLength):
            ctx.bit_reverse(F, O.empty(F, 3))
    F = O.FIELD_GL
    log_n = 3
    root = O.primitive_root(F, log_n)[1]
    tw = ctx.precompute_twiddles(F, log_n, root)
    data = O.from_ints(F, [1] * 8)
    before = data.copy()
    with pytest.raises(Z.InvalidTwiddles):
        ctx.ntt_with_twiddles(F, data, log_n, tw.stages[:2])
    with pytest.raises(Z.InvalidTwiddles):
        ctx.intt_with_twiddles(F, data, log_n, tw.stages[:2])
    with pytest.raises(Z.InvalidTwiddles):
        ctx.ntt_with_twiddles(F, data, log_n, [tw[0][:1], tw[1][:1], tw[2]])
    assert np.array_equal(data, before)  # rejected before the data is touched
    copy = data.copy()
    ctx.ntt_with_twiddles(F, copy, log_n, tw)
    ctx.intt_with_twiddles(F, copy, log_n, tw)
    assert np.array_equal(copy, data)
    tw.free()


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_bit_reverse_matches_oracle(Z, ctx, field):
    for log_n in (0, 1, 2, 7, 13):
        x = O.gen_uniform(field, 60 + log_n, 1 << log_n)
        exp = x.copy()
        O.bit_reverse(field, exp)
        got = x.copy()
        ctx.bit_reverse(field, got)
        assert np.array_equal(got, exp)


# ---------------------------------------------------------------------------
# Coset and LDE
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_coset_and_lde_match_oracle_and_horner(Z, ctx, field):
    shift = O.from_int(field, GENERATOR[field])
    for log_n in (0, 1, 6, 12, 15, 17, 19, 20):
        x = O.gen_uniform(field, 300 + log_n, 1 << log_n)
        root = O.primitive_root(field, log_n)[1]
        exp = x.copy()
        O.coset_ntt(field, exp, log_n, root, shift)
        got = x.copy()
        ctx.coset_ntt(field, got, log_n, root, shift)
        assert np.array_equal(got, exp), log_n
    fname = {O.FIELD_FR: "fr", O.FIELD_GL: "goldilocks"}.get(field)
    for c in load("dft_vectors.json")["lde"]:
        if c["field"] != fname:
            continue
        coeffs = O.from_ints(field, [int(v, 16) for v in c["coeffs"]])
        out = ctx.lde(field, coeffs, c["log_k"], c["log_blowup"], O.from_int(field, int(c["coset_shift"], 16)))
        assert [O.value(field, out[i]) for i in range(out.shape[0])] == [int(v, 16) for v in c["evals"]]
    for log_k, log_b in ((10, 2), (12, 3), (15, 2), (17, 3)):
        coeffs = O.gen_uniform(field, 40 + log_k, 1 << log_k)
        st, exp = O.lde(field, coeffs, log_k, log_b, shift)
        assert st == O.OK
        got = ctx.lde(field, coeffs, log_k, log_b, shift)
        assert np.array_equal(got, exp)


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_lde_and_coset_edge_cases(Z, ctx, field):
    """Degenerate LDE and coset shapes against the oracle: one coefficient,
    no blowup, blowup of a single coefficient, shifts 0 and 1; and the size
    errors of zntt_lde raised before its buffers are touched (sizes whose
    buffers could not be allocated, so the C ABI is called directly)."""
    import ctypes
    for log_k, log_b in ((0, 0), (0, 3), (4, 0), (1, 1), (3, 2)):
        coeffs = O.gen_uniform(field, 500 + 8 * log_k + log_b, 1 << log_k)
        for sh in (GENERATOR[field], 0, 1):
            shift = O.from_int(field, sh)
            st, exp = O.lde(field, coeffs, log_k, log_b, shift)
            assert st == O.OK
            got = ctx.lde(field, coeffs, log_k, log_b, shift)
            assert np.array_equal(got, exp), (log_k, log_b, sh)
    for log_n in (0, 1, 3, 9):
        x = O.gen_uniform(field, 520 + log_n, 1 << log_n)
        root = O.primitive_root(field, log_n)[1]
        for sh in (0, 1):
            shift = O.from_int(field, sh)
            exp = x.copy()
            assert O.coset_ntt(field, exp, log_n, root, shift) == O.OK
            got = x.copy()
            ctx.coset_ntt(field, got, log_n, root, shift)
            assert np.array_equal(got, exp), (log_n, sh)
    coeffs = O.gen_uniform(field, 7, 1)
    out = O.empty(field, 1)
    shift = O.from_int(field, 1)
    too_deep = {O.FIELD_FR: 29, O.FIELD_GL: 33, O.FIELD_BB: 28, O.FIELD_KB: 25}[field]  # two-adicity + 1
    for log_k, log_b, err in ((0, too_deep, Z.OrderTooLarge), (60, 10, Z.LogTooLarge), (70, 0, Z.LogTooLarge)):
        st = Z.lib.zntt_lde(ctx._h, field, coeffs.ctypes.data, log_k, log_b, shift.ctypes.data, out.ctypes.data)
        with pytest.raises(err):
            Z.check(st)


# ---------------------------------------------------------------------------
# Full-size workloads (BN254 Fr 2^16 .. 2^24) against committed digests# ---------------------------------------------------------------------------
# Full-size workloads (BN254 Fr 2^16 .. 2^24) against committed digests
# ---------------------------------------------------------------------------

def test_c1_roundtrip_2p16_seed0(Z, ctx):
    say("C1 2^16")
    c = load("textbook_digests.json")["cases"]["fr_2p16_seed0"]
    F = O.FIELD_FR
    x = O.gen_uniform(F, 0, 1 << 16)
    assert sha(x) == c["input_digest"]
    root = Z.primitive_root(Z.FIELD_FR, 16)
    data = x.copy()
    ctx.ntt(F, data, 16, root)
    assert sha(data) == c["forward_digest"]
    ctx.intt(F, data, 16, root)
    assert np.array_equal(data, x)


@pytest.mark.parametrize("log_n", [20, 22, 24])
def test_c2_large_fr_against_oracle_digests(Z, ctx, log_n):
    say("C2", log_n)
    d = load("oracle_digests.json")["fr"][str(log_n)]
    F = O.FIELD_FR
    x = O.gen_uniform(F, 1, 1 << log_n)
    assert sha(x) == d["input"]
    root = Z.primitive_root(Z.FIELD_FR, log_n)
    fwd = x.copy()
    ctx.ntt(F, fwd, log_n, root)
    assert sha(fwd) == d["forward"]
    if log_n == 20:
        assert sha(fwd) == load("textbook_digests.json")["cases"]["fr_2p20_seed1"]["forward_digest"]
    inv = x.copy()
    ctx.intt(F, inv, log_n, root)
    assert sha(inv) == d["inverse"]
    ctx.intt(F, fwd, log_n, root)  # round trip
    assert np.array_equal(fwd, x)


def test_single_hip_runtime_in_process(Z, ctx):
    """The library and torch must share one HIP runtime (torch bundles its own)."""
    import torch
    assert torch.cuda.is_available()
    maps = open("/proc/self/maps").read().splitlines()
    hips = sorted({ln.split()[-1] for ln in maps if "libamdhip64" in ln})
    say("HIP runtimes mapped:", hips)
    assert len(hips) == 1, hips


def test_device_resident_execute(Z, ctx):
    """zntt_execute_device on torch-allocated HBM, on torch's stream."""
    import torch
    F = O.FIELD_FR
    log_n = 18
    x = O.gen_uniform(F, 9, 1 << log_n)
    root = O.primitive_root(F, log_n)[1]
    shift = O.from_int(F, 5)
    t = torch.from_numpy(x.view(np.int64).copy()).cuda()
    s = torch.cuda.current_stream().cuda_stream
    for direction, fn in ((Z.FORWARD, O.ntt), (Z.INVERSE, O.intt), (Z.COSET_FORWARD, None)):
        d = t.clone()
        ctx.execute_device(F, direction, log_n, d.data_ptr(), root, shift if fn is None else None, stream=s)
        torch.cuda.synchronize()
        exp = x.copy()
        if fn is None:
            O.coset_ntt(F, exp, log_n, root, shift)
        else:
            fn(F, exp, log_n, root)
        assert np.array_equal(d.cpu().numpy().view(np.uint64).reshape(exp.shape), exp), direction


def test_context_budget_exhaustion(Z):
    """A context whose HBM budget cannot hold a transform's tables and
    staging reports OutOfMemory (the reference's error.OutOfMemory,
    root.zig:112) and stays usable for transforms that fit."""
    small = Z.Context(0, hbm_budget_bytes=1 << 20)
    try:
        F = O.FIELD_FR
        x = O.gen_uniform(F, 5, 1 << 16)  # 2 MiB of data alone
        with pytest.raises(Z.OutOfMemory):
            small.ntt(F, x.copy(), 16, O.primitive_root(F, 16)[1])
        y = O.gen_uniform(F, 6, 1 << 8)
        exp = y.copy()
        O.ntt(F, exp, 8, O.primitive_root(F, 8)[1])
        small.ntt(F, y, 8, O.primitive_root(F, 8)[1])
        assert np.array_equal(y, exp)
        assert small.bytes_in_use() <= 1 << 20
    finally:
        small.close()


@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_reserved_context_matches_oracle_and_holds_no_new_memory(Z, field):
    """After Context.reserve(14) (zntt_ctx_reserve) the seven zig-ntt
    functions at sizes up to 2^14 -- primitive and other roots, bit reversal,
    foreign and registered twiddle tables in both directions -- match the
    oracle bit for bit and leave the context's device bytes unchanged: the
    tables they need are built into the reservation's slots.  The budget
    holds the reservation and one registered table, nothing more."""
    say("reserved", FNAME[field])
    c = Z.Context(0, hbm_budget_bytes=8 << 20)
    try:
        c.reserve(14)
        root14 = O.primitive_root(field, 14)[1]
        reg = c.precompute_twiddles(field, 14, root14)
        st, otw = O.precompute_twiddles(field, 14, root14)
        assert st == O.OK
        held = c.bytes_in_use()
        other = O.from_int(field, 123456789)
        for log_n in (1, 7, 12, 14):
            n = 1 << log_n
            x = O.gen_uniform(field, 4000 + log_n, n)
            for root in (O.primitive_root(field, log_n)[1], other):
                for fwd in (True, False):
                    exp = x.copy()
                    assert (O.ntt if fwd else O.intt)(field, exp, log_n, root) == O.OK
                    got = x.copy()
                    (c.ntt if fwd else c.intt)(field, got, log_n, root)
                    assert np.array_equal(got, exp), (log_n, fwd)
            exp = x.copy()
            O.bit_reverse(field, exp)
            got = x.copy()
            c.bit_reverse(field, got)
            assert np.array_equal(got, exp)
            stages = [O.gen_uniform(field, 4100 + s, 1 << s) for s in range(log_n)]
            for inverse in (False, True):
                exp = x.copy()
                O.ntt_with_twiddles(field, exp, log_n, stages, inverse=inverse)
                got = x.copy()
                (c.intt_with_twiddles if inverse else c.ntt_with_twiddles)(field, got, log_n, stages)
                assert np.array_equal(got, exp), ("foreign", log_n, inverse)
        x = O.gen_uniform(field, 4200, 1 << 14)
        for inverse in (False, True):
            exp = x.copy()
            O.ntt_with_twiddles(field, exp, 14, otw, inverse=inverse)
            got = x.copy()
            (c.intt_with_twiddles if inverse else c.ntt_with_twiddles)(field, got, 14, reg)
            assert np.array_equal(got, exp), ("registered", inverse)
        assert c.bytes_in_use() == held
        reg.free()
    finally:
        c.close()


# ---------------------------------------------------------------------------
# Boundary: device-resident twiddled calls, stale registrations, streams
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("field", FIELDS, ids=IDS)
def test_twiddled_device_matches_oracle(Z, ctx, field):
    """nttWithTwiddles / inttWithTwiddles on device data, over a registered
    table and over a foreign (non-power) one."""
    import torch
    sp = torch.cuda.current_stream().cuda_stream
    for log_n in (9, 13, 17, 20):
        say("twiddled device", FNAME[field], log_n)
        root = O.primitive_root(field, log_n)[1]
        reg = ctx.precompute_twiddles(field, log_n, root)
        st, otw = O.precompute_twiddles(field, log_n, root)
        foreign = [O.gen_uniform(field, 300 + s, 1 << s) for s in range(log_n)]
        x = O.gen_uniform(field, 400 + log_n, 1 << log_n)
        for table, otable in ((reg, otw), (foreign, foreign)):
            for inverse in (False, True):
                exp = x.copy()
                O.ntt_with_twiddles(field, exp, log_n, otable, inverse=inverse)
                t = torch.from_numpy(x.view(np.int64).copy()).cuda()
                ctx.with_twiddles_device(field, t.data_ptr(), log_n, table, inverse=inverse, stream=sp)
                torch.cuda.synchronize()
                got = t.cpu().numpy().view(np.uint64).reshape(exp.shape)
                assert np.array_equal(got, exp), (log_n, inverse, table is reg)
        reg.free()


@pytest.mark.parametrize("field", [O.FIELD_FR, O.FIELD_GL], ids=["fr", "gl"])
def test_stale_registration_is_not_used(Z, ctx, field):
    """A registered table's host buffers reused for another table without
    zntt_free_twiddles (same addresses, other content): the call uses the
    buffers' current content, as the reference would."""
    log_n = 12
    root = O.primitive_root(field, log_n)[1]
    tw = ctx.precompute_twiddles(field, log_n, root)
    other = O.primitive_root(field, log_n + 1)[1]  # a root of another order: another table
    st, otw = O.precompute_twiddles(field, log_n, other)
    for s in range(log_n):
        tw.stages[s][...] = otw[s]  # in place: the registered pointers now hold another table
    x = O.gen_uniform(field, 77, 1 << log_n)
    for inverse in (False, True):
        exp = x.copy()
        O.ntt_with_twiddles(field, exp, log_n, otw, inverse=inverse)
        got = x.copy()
        (ctx.intt_with_twiddles if inverse else ctx.ntt_with_twiddles)(field, got, log_n, tw)
        assert np.array_equal(got, exp), inverse
    tw.free()


def test_multi_pass_transforms_on_two_streams(Z, ctx):
    """Multi-pass device transforms of one context enqueued on two streams
    without a host synchronisation between them: both results exact (the
    shared scratch is ordered across the streams)."""
    import torch
    F = O.FIELD_FR
    log_n = 17
    root = O.primitive_root(F, log_n)[1]
    xs = [O.gen_uniform(F, 1200 + i, 1 << log_n) for i in range(4)]
    exps = []
    for x in xs:
        e = x.copy()
        O.ntt(F, e, log_n, root)
        exps.append(e)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    ts = [torch.from_numpy(x.view(np.int64).copy()).cuda() for x in xs]
    torch.cuda.synchronize()
    for i, t in enumerate(ts):
        ctx.execute_device(F, Z.FORWARD, log_n, t.data_ptr(), root, stream=streams[i % 2].cuda_stream)
    torch.cuda.synchronize()
    for t, e in zip(ts, exps):
        assert np.array_equal(t.cpu().numpy().view(np.uint64).reshape(e.shape), e)
