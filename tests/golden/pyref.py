"""Python-int restatements used to pin the C oracle (test infrastructure).

Nothing here is derived from the oracle's code: every function is written from
the mathematics it names, with Python's arbitrary-precision integers.

* Montgomery constants for R = 2^256 (the reference derives the same at
  comptime, ``libs/field/src/montgomery.zig:36-79``).
* ``primitive_root`` follows the documented construction of
  ``libs/field/src/roots.zig:149-157``: z = smallest quadratic non-residue >= 2,
  root = z^(odd_part << (two_adicity - log)).
* ``dft`` is the O(n^2) definition X[k] = sum_j x[j] * w^(j*k), the check the
  reference itself uses (``libs/ntt/src/main.zig:39-47``,
  ``libs/ntt/src/root.zig:399-404``).
* ``textbook_ntt`` is a recursive even/odd radix-2 transform (a different
  algorithm from the reference's iterative DIT; for a primitive root both give
  the unique DFT).
* ``horner_coset`` evaluates f(shift * w^k) directly, pinning the coset/LDE
  definition.
* ``Xoshiro256pp`` is the published xoshiro256++ generator seeded by SplitMix64.
* ``counter_stored`` restates the counter-based stream of synthetic request
  coefficients from its definition (SplitMix64's finaliser over a counter).
"""
from __future__ import annotations

import hashlib

FR = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
GL = 0xFFFFFFFF00000001
BB = 0x78000001
R_BITS = 256
MASK64 = (1 << 64) - 1


def two_adicity(p: int) -> tuple[int, int]:
    q, s = p - 1, 0
    while q % 2 == 0:
        q //= 2
        s += 1
    return s, q


def legendre(a: int, p: int) -> int:
    r = pow(a, (p - 1) // 2, p)
    return 0 if r == 0 else (1 if r == 1 else -1)


def smallest_qnr(p: int) -> int:
    z = 2
    while legendre(z, p) != -1:
        z += 1
    return z


def primitive_root(p: int, log_size: int) -> int:
    s, q = two_adicity(p)
    if log_size > s:
        raise ValueError("OrderTooLarge")
    return pow(smallest_qnr(p), q << (s - log_size), p)


def mont_constants(p: int = FR) -> dict:
    R = 1 << R_BITS
    return {
        "r_mod_p": R % p,
        "r2_mod_p": (R * R) % p,
        "m0_inv64": (-pow(p, -1, 1 << 64)) % (1 << 64),
        "m0_inv32": (-pow(p, -1, 1 << 32)) % (1 << 32),
    }


def to_mont(x: int, p: int = FR) -> int:
    return (x << R_BITS) % p


def from_mont(x: int, p: int = FR) -> int:
    return (x * pow(1 << R_BITS, -1, p)) % p


def dft(xs: list[int], w: int, p: int) -> list[int]:
    n = len(xs)
    return [sum(x * pow(w, j * k, p) for j, x in enumerate(xs)) % p for k in range(n)]


def textbook_ntt(xs: list[int], w: int, p: int) -> list[int]:
    n = len(xs)
    if n == 1:
        return [xs[0] % p]
    even = textbook_ntt(xs[0::2], w * w % p, p)
    odd = textbook_ntt(xs[1::2], w * w % p, p)
    out = [0] * n
    h = n // 2
    wk = 1
    for k in range(h):
        t = wk * odd[k] % p
        out[k] = (even[k] + t) % p
        out[k + h] = (even[k] - t) % p
        wk = wk * w % p
    return out


def horner_coset(coeffs: list[int], n: int, w: int, shift: int, p: int) -> list[int]:
    """Evaluate f(x) = sum_i coeffs[i] x^i at shift * w^k for k < n."""
    out = []
    pt = shift % p
    for _ in range(n):
        acc = 0
        for c in reversed(coeffs):
            acc = (acc * pt + c) % p
        out.append(acc)
        pt = pt * w % p
    return out


class Xoshiro256pp:
    """xoshiro256++ (Blackman & Vigna), state seeded by four SplitMix64 outputs."""

    def __init__(self, seed: int):
        st = seed & MASK64
        self.s = []
        for _ in range(4):
            st = (st + 0x9E3779B97F4A7C15) & MASK64
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def _rotl(x: int, k: int) -> int:
        return ((x << k) | (x >> (64 - k))) & MASK64

    def next(self) -> int:
        s = self.s
        result = (self._rotl((s[0] + s[3]) & MASK64, 23) + s[0]) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result


def uniform_elements(p: int, seed: int, n: int) -> list[int]:
    """Canonical values: draw ceil(bits/8) little-endian bytes, reject >= p."""
    g = Xoshiro256pp(seed)
    nbytes = (p.bit_length() + 7) // 8
    words = (nbytes + 7) // 8
    out = []
    while len(out) < n:
        v = 0
        for i in range(words):
            v |= g.next() << (64 * i)
        v &= (1 << (8 * nbytes)) - 1
        if v < p:
            out.append(v)
    return out


def storage_bytes(p: int, canon: list[int]) -> bytes:
    """Bytes as the field type stores them (Montgomery limbs for Fr)."""
    if p == FR:
        return b"".join(to_mont(x).to_bytes(32, "little") for x in canon)
    return b"".join(x.to_bytes(8, "little") for x in canon)


def digest(p: int, canon: list[int]) -> str:
    return hashlib.sha256(storage_bytes(p, canon)).hexdigest()


def _mix64(z: int) -> int:
    """SplitMix64's output finaliser."""
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def counter_stored(p: int, seed: int, stream_id: int, n: int) -> list[int]:
    """Stored integers of elements 0..n-1 of stream (seed, stream_id): for Fr
    the 256-bit Montgomery word (four draws, top two bits cleared, first below
    p), for the small fields the canonical value (one draw masked to p's bit
    length, first below p).  Draw
    (i, attempt, word) is mix64(key + (i*256 + attempt*4 + word + 1) * gamma)
    with key = mix64(mix64(seed) ^ stream_id)."""
    gamma = 0x9E3779B97F4A7C15
    key = _mix64(_mix64(seed) ^ stream_id)
    words = 4 if p == FR else 1
    out = []
    for i in range(n):
        for att in range(64):
            v = 0
            for w in range(words):
                v |= _mix64((key + (i * 256 + att * 4 + w + 1) * gamma) & MASK64) << (64 * w)
            if p == FR:
                v &= (1 << 254) - 1
            else:
                v &= (1 << p.bit_length()) - 1
            if v < p:
                break
        out.append(v)
    return out
