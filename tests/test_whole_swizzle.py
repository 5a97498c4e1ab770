"""LDS layout of the whole-transform batch kernel (csrc/pass_kernels.hpp).

The kernel keeps one 2^K transform in LDS under the XOR swizzle whole_swz.
These checks read the swizzle's columns from the source and verify, on the
CPU, that it is a bijection on every tile size the kernel is built for.
They also check that the 64 lanes of each LDS access the kernel makes hit 64
distinct banks: the bit-reversed staging writes, and each register round's
reads and write-backs.
"""
import pathlib
import re

import numpy as np
import pytest

SRC = pathlib.Path(__file__).resolve().parents[1] / "zig-algebra_amd" / "csrc" / "pass_kernels.hpp"


def columns():
    m = re.search(r"kWholeSwzCol\[7\]\s*=\s*\{([^}]*)\}", SRC.read_text())
    assert m, "kWholeSwzCol not found"
    return [int(v) for v in m.group(1).split(",")]


def whole_swz(i, cols):
    m = np.zeros_like(i)
    for j, c in enumerate(cols):
        m ^= np.where((i >> (5 + j)) & 1, c, 0)
    return i ^ m


def brev(x, bits):
    out = np.zeros_like(x)
    for b in range(bits):
        out |= ((x >> b) & 1) << (bits - 1 - b)
    return out


def accesses(K, N):
    """Tile indices of every wave-wide LDS access, one row of 64 lanes each."""
    LN = 3 if N == 8 else 2
    NR = (K + LN - 1) // LN
    R0 = K - LN * (NR - 1)
    T = (1 << K) // N
    lanes = np.arange(T).reshape(-1, 64)
    rows = [brev(lanes + i * T, K) for i in range(N)]  # staging writes
    for ri in range(NR):
        R = R0 if ri == 0 else LN
        l0 = 0 if ri == 0 else R0 + LN * (ri - 1)
        for i in range(N >> R):
            rest = lanes + i * T
            base = (rest & ((1 << l0) - 1)) | ((rest >> l0) << (l0 + R))
            rows += [base | (k << l0) for k in range(1 << R)]
    return np.concatenate(rows)


CASES = [(12, 8), (11, 8), (10, 8), (9, 8), (11, 4), (10, 4), (9, 4)]


@pytest.mark.parametrize("K", [9, 10, 11, 12])
def test_bijection(K):
    i = np.arange(1 << K)
    f = whole_swz(i, columns())
    assert f.min() == 0 and f.max() == (1 << K) - 1
    assert len(np.unique(f)) == 1 << K


@pytest.mark.parametrize("K,N", CASES)
def test_every_access_hits_distinct_banks(K, N):
    rows = whole_swz(accesses(K, N), columns())
    banks = rows % 64
    assert all(len(set(r)) == 64 for r in banks.tolist())


def test_linear_split_used_by_the_kernel():
    # slot(base | k << l0) = slot(base) ^ slot(k << l0) for disjoint bits, and
    # brev(t + i*T) = brev(t) | brev(i*T) for t < T.
    cols = columns()
    rng = np.random.default_rng(5)
    a = rng.integers(0, 1 << 12, 4096)
    for l0 in range(0, 10):
        hi = ((1 << 3) - 1) << l0
        base = a & ~hi
        k = (a & hi)
        assert np.array_equal(whole_swz(base | k, cols), whole_swz(base, cols) ^ whole_swz(k, cols))


def test_old_swizzle_conflicts_on_these_patterns():
    # The shift-and-mask swizzle of the multi-column tiles, for comparison.
    rows = accesses(12, 8)
    old = rows ^ (((rows >> 6) & 3) << 3)
    worst = max(int(np.bincount(r % 64).max()) for r in old)
    assert worst >= 4
