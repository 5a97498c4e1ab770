"""ctypes loader for the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as the
thing measured or shipped.  The oracle restates
``libs/ntt/src/root.zig:33-212`` over ``libs/field/src/montgomery.zig`` (BN254
Fr) and ``libs/field/src/field.zig:89-424`` (small fields); see the header of
``zntt_oracle.c`` for the line-by-line map.

Element layout (identical to the reference's in-memory Zig types):
  * BN254 Fr: numpy ``uint64`` array of shape ``(n, 4)`` -- four little-endian
    limbs of the Montgomery form ``x * 2^256 mod r``;
  * Goldilocks / BabyBear: numpy ``uint64`` array of shape ``(n,)`` holding the
    canonical value.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

FIELD_FR = 1
FIELD_GL = 2
FIELD_BB = 3
FIELD_KB = 4
FIELD_M31 = 5
M31_MODULUS = 0x7FFFFFFF

FR_MODULUS = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
GL_MODULUS = 0xFFFFFFFF00000001
BB_MODULUS = 0x78000001
KB_MODULUS = 0x7F000001
MODULUS = {FIELD_FR: FR_MODULUS, FIELD_GL: GL_MODULUS, FIELD_BB: BB_MODULUS, FIELD_KB: KB_MODULUS}

OK = 0
LOG_TOO_LARGE = 1
LENGTH_MISMATCH = 2
INVALID_LENGTH = 3
INVALID_TWIDDLES = 4
OUT_OF_MEMORY = 5
INVALID_ARGUMENT = 8
ORDER_TOO_LARGE = 10
DEGREE_TOO_LARGE = 11
FRI_INVALID_PARAMETERS = 12
NUM_BYTES = {FIELD_FR: 32, FIELD_GL: 8, FIELD_BB: 4, FIELD_KB: 4}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, sz, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int
        L.oracle_init.restype = None
        L.oracle_elem_size.argtypes = [i32]
        L.oracle_field_op.argtypes = [i32, i32, vp, vp, vp]
        L.oracle_primitive_root.argtypes = [i32, sz, vp]
        L.oracle_bit_reverse.argtypes = [i32, vp, sz]
        L.oracle_ntt.argtypes = [i32, vp, sz, sz, vp]
        L.oracle_intt.argtypes = [i32, vp, sz, sz, vp]
        L.oracle_coset_ntt.argtypes = [i32, vp, sz, sz, vp, vp]
        L.oracle_precompute_twiddles.argtypes = [i32, sz, vp, vp]
        L.oracle_ntt_with_twiddles.argtypes = [i32, i32, vp, sz, sz, vp, vp, sz]
        L.oracle_lde.argtypes = [i32, vp, sz, sz, vp, vp]
        L.oracle_gen_uniform.argtypes = [i32, u64, vp, sz]
        L.oracle_fill_counter.argtypes = [i32, u64, u64, vp, sz]
        L.oracle_poly_mul.argtypes = [i32, vp, sz, vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.oracle_poly_eval.argtypes = [i32, vp, sz, vp, vp]
        L.oracle_xoshiro_stream.argtypes = [u64, vp, sz]
        L.oracle_xoshiro_stream.restype = None
        L.oracle_ntt_batch_mt.argtypes = [i32, vp, vp, vp, vp, sz, i32]
        L.oracle_blake3.argtypes = [ctypes.c_char_p, sz, vp]
        L.oracle_blake3.restype = None
        L.oracle_fri_leaves.argtypes = [i32, vp, sz, vp]
        L.oracle_merkle_from_hashes.argtypes = [vp, sz, vp, vp]
        L.oracle_merkle_commit.argtypes = [i32, vp, sz, vp, vp]
        L.oracle_tr_new.argtypes = [ctypes.c_char_p, sz]
        L.oracle_tr_new.restype = vp
        L.oracle_tr_free.argtypes = [vp]
        L.oracle_tr_free.restype = None
        L.oracle_tr_absorb.argtypes = [vp, ctypes.c_char_p, sz]
        L.oracle_tr_absorb.restype = None
        L.oracle_tr_challenge_bytes.argtypes = [vp, vp, sz]
        L.oracle_tr_challenge_bytes.restype = None
        L.oracle_tr_challenge_u64.argtypes = [vp]
        L.oracle_tr_challenge_u64.restype = u64
        L.oracle_tr_challenge_field.argtypes = [vp, i32, vp]
        L.oracle_fri_prove.argtypes = [i32, vp, vp, sz, sz, sz, sz, sz, sz, vp, vp, vp, vp, vp]
        L.oracle_vec8_m31.argtypes = [i32, vp, sz, sz, u64]
        L.oracle_init()
        _lib = L
    return _lib


def limbs(field: int) -> int:
    return 4 if field == FIELD_FR else 1


def empty(field: int, n: int) -> np.ndarray:
    return np.zeros((n, 4) if field == FIELD_FR else (n,), dtype=np.uint64)


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"] and a.dtype == np.uint64
    return a.ctypes.data_as(ctypes.c_void_p)


def _as_elem(field: int, x) -> np.ndarray:
    """One element as a contiguous 1-D uint64 array of its limbs."""
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64).reshape(-1))


def elem(field: int, value: int) -> np.ndarray:
    """Raw element bytes from an integer already in storage form."""
    if field == FIELD_FR:
        return np.array([(value >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)
    return np.array([value], dtype=np.uint64)


def to_int(field: int, a: np.ndarray) -> int:
    a = np.asarray(a, dtype=np.uint64).reshape(-1)
    return sum(int(x) << (64 * i) for i, x in enumerate(a))


# --- field arithmetic ------------------------------------------------------
_OPS = {"add": 0, "sub": 1, "mul": 2, "neg": 3, "inv": 4, "to_mont": 5, "from_mont": 6, "pow": 7}


def field_op(field: int, op: str, a, b=None):
    """One field operation; returns the same shape as ``a`` (a scalar for a
    small-field scalar, four limbs for Fr)."""
    scalar = np.ndim(a) == 0
    ea = _as_elem(field, a)
    eb = None if b is None else _as_elem(field, b)
    out = np.zeros_like(ea)
    st = lib().oracle_field_op(field, _OPS[op], _ptr(ea), _ptr(eb), _ptr(out))
    assert st == OK, st
    return out[0] if scalar else out


def from_int(field: int, x: int) -> np.ndarray:
    """Element for integer x (reduced mod p); Montgomery form for Fr."""
    x %= MODULUS[field]
    e = elem(field, x)
    return field_op(field, "to_mont", e) if field == FIELD_FR else e


def value(field: int, a) -> int:
    """Canonical integer of one element."""
    return to_int(field, field_op(field, "from_mont", _as_elem(field, a)) if field == FIELD_FR else a)


def from_ints(field: int, ints) -> np.ndarray:
    """Storage-form array from canonical integers."""
    a = empty(field, len(ints))
    for i, x in enumerate(ints):
        e = from_int(field, x)
        a[i] = e if field == FIELD_FR else e[0]
    return a


def primitive_root(field: int, log_size: int) -> tuple[int, np.ndarray]:
    out = empty(field, 1).reshape(-1)
    st = lib().oracle_primitive_root(field, log_size, _ptr(out))
    return st, out


# --- the seven functions + coset/LDE ---------------------------------------
def bit_reverse(field: int, data: np.ndarray) -> int:
    return lib().oracle_bit_reverse(field, _ptr(data), data.shape[0])


def ntt(field: int, data: np.ndarray, log_n: int, root: np.ndarray, length: int | None = None) -> int:
    n = data.shape[0] if length is None else length
    return lib().oracle_ntt(field, _ptr(data), n, log_n, _ptr(_as_elem(field, root)))


def intt(field: int, data: np.ndarray, log_n: int, root: np.ndarray, length: int | None = None) -> int:
    n = data.shape[0] if length is None else length
    return lib().oracle_intt(field, _ptr(data), n, log_n, _ptr(_as_elem(field, root)))


def coset_ntt(field: int, data: np.ndarray, log_n: int, root: np.ndarray, shift: np.ndarray) -> int:
    return lib().oracle_coset_ntt(field, _ptr(data), data.shape[0], log_n, _ptr(_as_elem(field, root)),
                                  _ptr(_as_elem(field, shift)))


def precompute_twiddles(field: int, log_n: int, root: np.ndarray) -> tuple[int, list[np.ndarray]]:
    """Returns (status, stages); stage s holds 2^s elements."""
    if log_n >= 64:
        flat = empty(field, 1)
        st = lib().oracle_precompute_twiddles(field, log_n, _ptr(_as_elem(field, root)), _ptr(flat))
        return st, []
    total = (1 << log_n) - 1
    flat = empty(field, max(total, 1))
    st = lib().oracle_precompute_twiddles(field, log_n, _ptr(_as_elem(field, root)), _ptr(flat))
    stages = [flat[(1 << s) - 1:(1 << (s + 1)) - 1] for s in range(log_n)]
    return st, stages


def ntt_with_twiddles(field: int, data: np.ndarray, log_n: int, stages: list[np.ndarray],
                      inverse: bool = False) -> int:
    stages = [np.ascontiguousarray(s) for s in stages]
    ptrs = (ctypes.c_void_p * max(len(stages), 1))(*[s.ctypes.data for s in stages])
    lens = (ctypes.c_size_t * max(len(stages), 1))(*[s.shape[0] for s in stages])
    return lib().oracle_ntt_with_twiddles(field, int(inverse), _ptr(data), data.shape[0], log_n,
                                          ctypes.cast(ptrs, ctypes.c_void_p),
                                          ctypes.cast(lens, ctypes.c_void_p), len(stages))


def intt_with_twiddles(field: int, data: np.ndarray, log_n: int, stages: list[np.ndarray]) -> int:
    return ntt_with_twiddles(field, data, log_n, stages, inverse=True)


def lde(field: int, coeffs: np.ndarray, log_k: int, log_blowup: int, shift: np.ndarray) -> tuple[int, np.ndarray]:
    out = empty(field, 1 << (log_k + log_blowup))
    st = lib().oracle_lde(field, _ptr(coeffs), log_k, log_blowup, _ptr(_as_elem(field, shift)), _ptr(out))
    return st, out


def gen_uniform(field: int, seed: int, n: int) -> np.ndarray:
    out = empty(field, n)
    st = lib().oracle_gen_uniform(field, seed, _ptr(out), n)
    assert st == OK
    return out


def poly_mul(field: int, a: np.ndarray, b: np.ndarray, cap: int | None = None) -> tuple[int, np.ndarray]:
    """Schoolbook Polynomial.mul (libs/poly/src/poly.zig:197-212): (status, product)."""
    cap = a.shape[0] + b.shape[0] if cap is None else cap
    out = empty(field, max(cap, 1))
    n = ctypes.c_size_t()
    st = lib().oracle_poly_mul(field, _ptr(np.ascontiguousarray(a)), a.shape[0], _ptr(np.ascontiguousarray(b)),
                               b.shape[0], _ptr(out), cap, ctypes.byref(n))
    return st, out[:n.value].copy()


def poly_eval(field: int, coeffs: np.ndarray, point) -> np.ndarray:
    out = empty(field, 1).reshape(-1)
    st = lib().oracle_poly_eval(field, _ptr(np.ascontiguousarray(coeffs)), coeffs.shape[0],
                                _ptr(_as_elem(field, point)), _ptr(out))
    assert st == OK
    return out


def fill_counter(field: int, seed: int, stream_id: int, n: int) -> np.ndarray:
    """Counter-based stream of synthetic request coefficients."""
    out = empty(field, n)
    st = lib().oracle_fill_counter(field, seed, stream_id, _ptr(out), n)
    assert st == OK
    return out


def xoshiro_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().oracle_xoshiro_stream(seed, _ptr(out), n)
    return out


def ntt_batch_mt(field: int, data: np.ndarray, offsets: np.ndarray, log_ns: np.ndarray,
                 roots_by_log: np.ndarray, threads: int) -> int:
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    log_ns = np.ascontiguousarray(log_ns, dtype=np.uint64)
    return lib().oracle_ntt_batch_mt(field, _ptr(data), _ptr(offsets), _ptr(log_ns),
                                     _ptr(roots_by_log), len(log_ns), threads)


def _bptr(a: np.ndarray):
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"] and a.dtype == np.uint8
    return a.ctypes.data_as(ctypes.c_void_p)


def blake3(data: bytes) -> bytes:
    """BLAKE3 of `data`, 32-byte output."""
    out = np.zeros(32, dtype=np.uint8)
    lib().oracle_blake3(bytes(data), len(data), _bptr(out))
    return out.tobytes()


def fri_leaves(field: int, evals: np.ndarray) -> np.ndarray:
    """Leaf hashes of one FRI layer: (n/2, 32) uint8."""
    n = evals.shape[0]
    out = np.zeros((max(n // 2, 1), 32), dtype=np.uint8)
    st = lib().oracle_fri_leaves(field, _ptr(np.ascontiguousarray(evals)), n, _bptr(out))
    assert st == OK, st
    return out


def merkle_from_hashes(hashes: np.ndarray) -> tuple[np.ndarray, bytes]:
    """MerkleTree.initFromHashes: (nodes (2*cap, 32) uint8, root)."""
    hashes = np.ascontiguousarray(hashes, dtype=np.uint8).reshape(-1, 32)
    cap = 1
    while cap < hashes.shape[0]:
        cap *= 2
    nodes = np.zeros((2 * cap, 32), dtype=np.uint8)
    root = np.zeros(32, dtype=np.uint8)
    st = lib().oracle_merkle_from_hashes(_bptr(hashes), hashes.shape[0], _bptr(nodes), _bptr(root))
    assert st == OK, st
    return nodes, root.tobytes()


def merkle_commit(field: int, evals: np.ndarray) -> tuple[np.ndarray, bytes]:
    """Commitment of one FRI layer of n = 2^k values: (nodes (n, 32) uint8, root)."""
    n = evals.shape[0]
    nodes = np.zeros((n, 32), dtype=np.uint8)
    root = np.zeros(32, dtype=np.uint8)
    st = lib().oracle_merkle_commit(field, _ptr(np.ascontiguousarray(evals)), n, _bptr(nodes), _bptr(root))
    assert st == OK, st
    return nodes, root.tobytes()


class Transcript:
    """The oracle's Fiat-Shamir transcript (libs/transcript/src/root.zig)."""

    def __init__(self, label: bytes):
        self._h = lib().oracle_tr_new(bytes(label), len(label))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_tr_free(self._h)
            self._h = None

    def absorb_bytes(self, data: bytes) -> None:
        lib().oracle_tr_absorb(self._h, bytes(data), len(data))

    def absorb_u64(self, v: int) -> None:
        self.absorb_bytes(int(v).to_bytes(8, "little"))

    def challenge_bytes(self, n: int) -> bytes:
        out = np.zeros(max(n, 1), dtype=np.uint8)
        lib().oracle_tr_challenge_bytes(self._h, _bptr(out), n)
        return out[:n].tobytes()

    def challenge_u64(self) -> int:
        return int(lib().oracle_tr_challenge_u64(self._h))

    def challenge_field(self, field: int) -> np.ndarray:
        out = empty(field, 1)
        st = lib().oracle_tr_challenge_field(self._h, field, _ptr(out))
        assert st == OK
        return out[0]


def fri_proof_sizes(field: int, log_domain: int, log_final: int, log_residual_degree: int, num_queries: int):
    rounds = log_domain - log_final
    depth = sum(log_domain - r - 1 for r in range(rounds))
    nb = NUM_BYTES[field]
    return {"rounds": rounds, "roots": rounds * 32, "residual": (1 << log_residual_degree) * nb,
            "values": num_queries * rounds * 2 * nb, "paths": num_queries * depth * 32}


def fri_prove(field: int, transcript: Transcript, evals: np.ndarray, log_domain: int, log_initial_degree: int,
              log_final: int, log_residual_degree: int, num_queries: int):
    """prove(F, transcript, evals, config) -> (status, proof dict of flat arrays)."""
    rounds = max(log_domain - log_final, 0)
    sz = fri_proof_sizes(field, log_domain, log_final, log_residual_degree, num_queries) if rounds > 0 else None
    if sz is None:
        sz = {"roots": 0, "residual": 0, "values": 0, "paths": 0}
    roots = np.zeros(max(sz["roots"], 1), dtype=np.uint8)
    residual = np.zeros(max(sz["residual"], 1), dtype=np.uint8)
    pair = np.zeros(max(num_queries, 1), dtype=np.uint64)
    values = np.zeros(max(sz["values"], 1), dtype=np.uint8)
    paths = np.zeros(max(sz["paths"], 1), dtype=np.uint8)
    ev = np.ascontiguousarray(evals)
    st = lib().oracle_fri_prove(field, transcript._h, _ptr(ev), ev.shape[0], log_domain, log_initial_degree,
                                log_final, log_residual_degree, num_queries, _bptr(roots), _bptr(residual),
                                _ptr(pair), _bptr(values), _bptr(paths))
    if st != OK:
        return st, None
    return st, {"roots": roots[:sz["roots"]].tobytes(), "residual": residual[:sz["residual"]].tobytes(),
                "pair_index": [int(x) for x in pair[:num_queries]], "values": values[:sz["values"]].tobytes(),
                "paths": paths[:sz["paths"]].tobytes()}


def vec8_m31(data: np.ndarray, log_n: int, root: int, inverse: bool = False, length: int | None = None) -> int:
    """Vec8NttM31 (in place) on interleaved u64 values data[i * 8 + lane]."""
    assert data.dtype == np.uint64 and data.flags["C_CONTIGUOUS"]
    n = data.shape[0] if length is None else length
    return lib().oracle_vec8_m31(1 if inverse else 0, _ptr(data), n, log_n, root)
