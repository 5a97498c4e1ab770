"""GPU tests of the Merkle commitment (zntt_merkle_commit,
zntt_merkle_from_hashes): BLAKE3 leaves over antipodal pairs of a FRI layer
and MerkleTree.initFromHashes above them.

Roots and whole node arrays equal the oracle's byte for byte at sizes the
oracle finishes in seconds, across every launch split of the device tree
(one launch below 2^9 leaves, two up to 2^18, three above).  At 2^24
evaluations the device tree is checked by properties: every sampled inner
node hashes its two children, sampled leaves equal the oracle's leaf of the
same pair, and the root-only call (scratch heap) returns the full tree's root.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ALL = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]


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


def nodes_host(t, count):
    return t.cpu().numpy().view(np.uint8).reshape(count, 32)


@pytest.mark.parametrize("field", ALL, ids=IDS)
@pytest.mark.parametrize("log_n", [1, 2, 3, 9, 10, 11, 14, 19, 20])
def test_commit_matches_oracle(Z, ctx, field, log_n):
    import torch
    n = 1 << log_n
    ev = O.gen_uniform(field, 100 + log_n, n)
    want_nodes, want_root = O.merkle_commit(field, ev)
    d = dev(ev)
    nodes = torch.full((n * 4,), -1, dtype=torch.int64, device="cuda")
    root = ctx.merkle_commit_device(field, d.data_ptr(), log_n, nodes.data_ptr(), stream=stream())
    assert root == want_root
    assert np.array_equal(nodes_host(nodes, n), want_nodes)
    # root only: the context's scratch heap
    assert ctx.merkle_commit_device(field, d.data_ptr(), log_n, stream=stream()) == want_root


@pytest.mark.parametrize("count", [1, 2, 3, 5, 511, 512, 513, 1000, 1 << 14, (1 << 18) + 3])
def test_from_hashes_matches_oracle(Z, ctx, count):
    import torch
    rng = np.random.default_rng(count)
    hs = rng.integers(0, 256, size=(count, 32), dtype=np.uint8)
    want_nodes, want_root = O.merkle_from_hashes(hs)
    cap = want_nodes.shape[0] // 2
    d = torch.from_numpy(hs.reshape(-1).copy()).cuda()
    nodes = torch.full((2 * cap * 4,), -1, dtype=torch.int64, device="cuda")
    root = ctx.merkle_from_hashes_device(d.data_ptr(), count, nodes.data_ptr(), stream=stream())
    assert root == want_root
    assert np.array_equal(nodes_host(nodes, 2 * cap), want_nodes)
    assert ctx.merkle_from_hashes_device(d.data_ptr(), count, stream=stream()) == want_root


def test_commit_of_lde_output(Z, ctx):
    """The LDE's consumer: commit the device LDE of 2^12 coefficients, blowup
    8, and compare with the oracle's commitment of the oracle's LDE."""
    import torch
    F = O.FIELD_GL
    coeffs = O.gen_uniform(F, 3, 1 << 12)
    sh = O.from_int(F, 7)
    st, lde = O.lde(F, coeffs, 12, 3, sh)
    assert st == O.OK
    got = ctx.lde(F, coeffs, 12, 3, sh)
    assert np.array_equal(got, lde)
    _, want_root = O.merkle_commit(F, lde)
    d = dev(got)
    assert ctx.merkle_commit_device(F, d.data_ptr(), 15, stream=stream()) == want_root


@pytest.mark.parametrize("field", [O.FIELD_FR, O.FIELD_BB], ids=["fr", "bb"])
def test_full_size_tree_properties(Z, ctx, field):
    """2^24 evaluations: tree consistency on samples, leaves against the oracle."""
    import torch
    log_n = 24
    n = 1 << log_n
    half = n // 2
    words = 4 if field == O.FIELD_FR else 1
    ev = torch.empty(n * words, dtype=torch.int64, device="cuda")
    Z.fill_counter_device(ctx, field, 11, 5, ev.data_ptr(), n, stream=stream())
    nodes = torch.empty(n * 4, dtype=torch.int64, device="cuda")
    root = ctx.merkle_commit_device(field, ev.data_ptr(), log_n, nodes.data_ptr(), stream=stream())
    assert root == ctx.merkle_commit_device(field, ev.data_ptr(), log_n, stream=stream())
    nh = nodes.cpu().numpy().view(np.uint8).reshape(n, 32)
    assert not nh[0].any() and nh[1].tobytes() == root
    rng = np.random.default_rng(1)
    for i in list(rng.integers(1, half, size=300)) + [1, 2, 3, half - 1]:
        assert nh[i].tobytes() == O.blake3(nh[2 * i].tobytes() + nh[2 * i + 1].tobytes()), i
    evh = ev.cpu().numpy().view(np.uint64)
    evh = evh.reshape(n, 4) if field == O.FIELD_FR else evh
    for j in list(rng.integers(0, half, size=200)) + [0, half - 1]:
        pair = np.ascontiguousarray(np.stack([evh[j], evh[j + half]]))
        assert nh[half + j].tobytes() == O.fri_leaves(field, pair)[0].tobytes(), j


def test_merkle_errors(Z, ctx):
    import torch
    d = torch.zeros(16, dtype=torch.int64, device="cuda")
    with pytest.raises(Z.InvalidLength):
        ctx.merkle_commit_device(O.FIELD_GL, d.data_ptr(), 0)
    with pytest.raises(Z.Unsupported):
        ctx.merkle_commit_device(O.FIELD_GL, d.data_ptr(), 29)
    with pytest.raises(Z.LogTooLarge):
        ctx.merkle_commit_device(O.FIELD_GL, d.data_ptr(), 64)
    with pytest.raises(Z.InvalidArgument):
        ctx.merkle_commit_device(99, d.data_ptr(), 3)
    with pytest.raises(Z.InvalidLength):
        ctx.merkle_from_hashes_device(d.data_ptr(), 0)
