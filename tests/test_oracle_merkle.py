"""Pins for the oracle's BLAKE3, FRI leaf hashing and Merkle tree (CPU).

* BLAKE3 against the reference's own known answers
  (libs/hash/src/root.zig:121-127, :173-247): 45 lengths from 0 to 4097
  bytes of a repeating pattern -- every block and chunk edge, the chunk
  counter and the parent nodes -- and "hello world".
* The FRI leaf (libs/fri/src/root.zig:283-292) against bytes built from
  Python integers: the canonical value (out of Montgomery form for BN254 Fr,
  tests/golden/pyref.py), little-endian, 32 / 8 / 4 / 4 bytes, paired with
  the value half a layer away.  A Montgomery-form byte string, a wrong byte
  count or a (j, j + 1) pairing fails it.
* MerkleTree.initFromHashes (libs/merkle/src/merkle_tree.zig:147-182)
  against a recursive restatement (split the padded leaf list in halves) and
  the closed forms: one leaf is its own root, two leaves hash once, padding
  leaves are BLAKE3("") (the empty-input known answer), node 0 stays zero.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import pyref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NUM_BYTES = {O.FIELD_FR: 32, O.FIELD_GL: 8, O.FIELD_BB: 4, O.FIELD_KB: 4}
ALL = [O.FIELD_FR, O.FIELD_GL, O.FIELD_BB, O.FIELD_KB]
IDS = ["fr", "gl", "bb", "kb"]


def kat():
    with open(os.path.join(GOLD, "ref_blake3_kat.json")) as f:
        return json.load(f)


def test_blake3_reference_known_answers():
    k = kat()
    assert O.blake3(b"hello world").hex() == k["hello_world"]
    pat = bytes(k["pattern"])
    data = (pat * (4097 // len(pat) + 1))[:4097]
    assert len(k["vectors"]) == 45
    for v in k["vectors"]:
        assert O.blake3(data[: v["len"]]).hex() == v["digest"], v["len"]


def canonical_bytes(field, e):
    if field == O.FIELD_FR:
        m = sum(int(e[k]) << (64 * k) for k in range(4))
        return pyref.from_mont(m).to_bytes(32, "little")
    return int(e).to_bytes(NUM_BYTES[field], "little")


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_fri_leaves_from_integer_bytes(field):
    n = 64
    ev = O.gen_uniform(field, 17 + field, n)
    got = O.fri_leaves(field, ev)
    assert got.shape == (n // 2, 32)
    for j in range(n // 2):
        msg = canonical_bytes(field, ev[j]) + canonical_bytes(field, ev[j + n // 2])
        assert len(msg) == 2 * NUM_BYTES[field]
        assert got[j].tobytes() == O.blake3(msg), j


def test_fr_leaf_is_not_montgomery_bytes():
    """The leaf of the element 1 hashes the integer 1, not R mod r."""
    one = O.from_int(O.FIELD_FR, 1).reshape(1, 4)
    ev = np.concatenate([one, one])
    leaf = O.fri_leaves(O.FIELD_FR, ev)[0].tobytes()
    assert leaf == O.blake3((1).to_bytes(32, "little") * 2)


def recursive_root(leaves):
    if len(leaves) == 1:
        return leaves[0]
    h = len(leaves) // 2
    return O.blake3(recursive_root(leaves[:h]) + recursive_root(leaves[h:]))


@pytest.mark.parametrize("count", [1, 2, 3, 4, 5, 7, 8, 9, 33, 100])
def test_merkle_from_hashes(count):
    rng = np.random.default_rng(count)
    hs = rng.integers(0, 256, size=(count, 32), dtype=np.uint8)
    nodes, root = O.merkle_from_hashes(hs)
    cap = 1 << (count - 1).bit_length()
    empty = O.blake3(b"")
    assert empty.hex() == kat()["vectors"][0]["digest"]
    padded = [hs[i].tobytes() for i in range(count)] + [empty] * (cap - count)
    assert nodes.shape == (2 * cap, 32)
    assert not nodes[0].any()
    assert root == nodes[1].tobytes() == recursive_root(padded)
    for i in range(cap):
        assert nodes[cap + i].tobytes() == padded[i]
    if count == 1:
        assert root == hs[0].tobytes()
    if count == 2:
        assert root == O.blake3(hs[0].tobytes() + hs[1].tobytes())


@pytest.mark.parametrize("field", ALL, ids=IDS)
def test_merkle_commit_is_tree_of_leaves(field):
    for n in (2, 4, 32):
        ev = O.gen_uniform(field, 5, n)
        nodes, root = O.merkle_commit(field, ev)
        n2, r2 = O.merkle_from_hashes(O.fri_leaves(field, ev))
        assert root == r2 and np.array_equal(nodes, n2)
        if n == 2:
            assert root == O.fri_leaves(field, ev)[0].tobytes()
