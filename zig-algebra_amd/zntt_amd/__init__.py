"""Python host binding of libzntt_gpu.so, the MI355X NTT library.

Mirrors zig-ntt's interface (``libs/ntt/src/root.zig``): the same seven
functions, the same argument meaning and the same typed errors, raised as
exceptions named after the Zig error values.  Every call goes through the C
ABI declared in ``include/zntt_gpu.h``; there is no CPU fallback -- importing
this module without the built library raises immediately.

Arrays use the reference's element layout:
  * BN254 Fr: numpy ``uint64`` of shape ``(n, 4)`` (Montgomery limbs, R = 2^256);
  * Goldilocks: numpy ``uint64`` of shape ``(n,)`` (canonical values).
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

FIELD_FR = 1
FIELD_GL = 2
FIELD_BB = 3  # BabyBear, p = 2^31 - 2^27 + 1
FIELD_KB = 4  # KoalaBear, p = 2^31 - 2^24 + 1

FORWARD = 0
INVERSE = 1
COSET_FORWARD = 2

MAX_LOG_N = 28

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzntt_gpu.so")


class ZnttError(Exception):
    """Base class; ``status`` is the C ABI status code."""
    status = -1


class LogTooLarge(ZnttError):
    status = 1


class LengthMismatch(ZnttError):
    status = 2


class InvalidLength(ZnttError):
    status = 3


class InvalidTwiddles(ZnttError):
    status = 4


class OutOfMemory(ZnttError):
    status = 5


class DeviceError(ZnttError):
    status = 6


class NoDevice(ZnttError):
    status = 7


class InvalidArgument(ZnttError):
    status = 8


class Unsupported(ZnttError):
    status = 9


class OrderTooLarge(ZnttError):
    status = 10


class DegreeTooLarge(ZnttError):
    status = 11


class InvalidParameters(ZnttError):
    status = 12


class NotPrepared(ZnttError):
    status = 13


_ERRORS = {c.status: c for c in (LogTooLarge, LengthMismatch, InvalidLength, InvalidTwiddles, OutOfMemory,
                                 DeviceError, NoDevice, InvalidArgument, Unsupported, OrderTooLarge,
                                 DegreeTooLarge, InvalidParameters, NotPrepared)}

# Every symbol include/zntt_gpu.h declares.
EXPORTED = [
    "zntt_ctx_create", "zntt_ctx_destroy", "zntt_ctx_bytes_in_use", "zntt_ctx_trim", "zntt_ctx_reserve",
    "zntt_status_string",
    "zntt_element_size", "zntt_bit_reverse", "zntt_ntt", "zntt_intt", "zntt_precompute_twiddles",
    "zntt_free_twiddles", "zntt_ntt_with_twiddles", "zntt_intt_with_twiddles", "zntt_coset_ntt", "zntt_lde",
    "zntt_primitive_root", "zntt_execute_device", "zntt_ctx_synchronize", "zntt_debug_field_op",
    "zntt_probe_peaks", "zntt_fill_uniform", "zntt_ctx_stream",
    "zntt_batch_create", "zntt_batch_execute", "zntt_batch_status", "zntt_batch_autotune", "zntt_batch_shape",
    "zntt_batch_set_shape", "zntt_batch_destroy",
    "zntt_server_create", "zntt_server_submit", "zntt_server_query", "zntt_server_wait", "zntt_server_cancel",
    "zntt_server_read", "zntt_server_release", "zntt_server_drain", "zntt_server_stats_get", "zntt_server_sim_pump",
    "zntt_server_sim_complete", "zntt_server_destroy", "zntt_fill_counter",
    "zntt_poly_mul", "zntt_poly_mul_device", "zntt_merkle_commit", "zntt_merkle_from_hashes",
    "zntt_fri_proof_sizes", "zntt_fri_prove", "zntt_transcript_create", "zntt_transcript_destroy",
    "zntt_transcript_absorb_bytes", "zntt_transcript_challenge_bytes", "zntt_blake3",
    "zntt_ntt_vec8_m31", "zntt_intt_vec8_m31", "zntt_vec8_m31_device",
    "zntt_ntt_with_twiddles_device", "zntt_intt_with_twiddles_device", "zntt_execute_device_noalloc",
]

M31_MODULUS = 0x7FFFFFFF

BYTES_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)


class TranscriptOps(ctypes.Structure):
    _fields_ = [("absorb_bytes", BYTES_FN), ("challenge_bytes", BYTES_FN)]


class FriConfig(ctypes.Structure):
    _fields_ = [("log_domain", ctypes.c_size_t), ("log_initial_degree", ctypes.c_size_t),
                ("log_final", ctypes.c_size_t), ("log_residual_degree", ctypes.c_size_t),
                ("num_queries", ctypes.c_size_t)]


class FriProof(ctypes.Structure):
    _fields_ = [("roots", ctypes.c_void_p), ("residual", ctypes.c_void_p), ("pair_index", ctypes.c_void_p),
                ("values", ctypes.c_void_p), ("paths", ctypes.c_void_p)]

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = 0, 1, 2, 3, 4

ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p))
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)


class Allocator(ctypes.Structure):
    _fields_ = [("user", ctypes.c_void_p), ("alloc", ALLOC_FN), ("free", FREE_FN)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("hbm_budget", ctypes.c_uint64), ("max_active", ctypes.c_int),
                ("max_log_n", ctypes.c_int), ("simulate", ctypes.c_int),
                ("allocator", ctypes.POINTER(Allocator))]


class LdeRequest(ctypes.Structure):
    _fields_ = [("log_n", ctypes.c_size_t), ("len", ctypes.c_size_t), ("log_blowup", ctypes.c_size_t),
                ("shift", ctypes.c_void_p), ("coeffs", ctypes.c_void_p), ("seed", ctypes.c_uint64),
                ("stream_id", ctypes.c_uint64), ("keep_result", ctypes.c_int), ("out", ctypes.c_void_p)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int), ("queued_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("data", ctypes.c_void_p), ("log_n", ctypes.c_size_t),
                ("log_blowup", ctypes.c_size_t)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "admitted", "completed", "failed", "cancelled",
                                               "steps", "elements_done", "bytes_in_use", "peak_bytes", "queued",
                                               "active")] + [("busy_ms", ctypes.c_double)]


class Transform(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("direction", ctypes.c_int), ("log_n", ctypes.c_size_t),
                ("data", ctypes.c_void_p), ("root", ctypes.c_void_p), ("shift", ctypes.c_void_p)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C zig-algebra_amd` (or __graft_entry__.build())")
    # PyTorch-ROCm ships its own HIP/HSA runtime.  Two HIP runtimes in one
    # process cannot share the GPU, so when torch is installed it is loaded
    # first and the library's libamdhip64 dependency binds to that copy; device
    # pointers and streams from torch are then valid in the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
    pp = ctypes.POINTER(ctypes.c_void_p)
    L.zntt_ctx_create.argtypes = [ctypes.POINTER(vp), i32, u64]
    L.zntt_ctx_destroy.argtypes = [vp]
    L.zntt_ctx_destroy.restype = None
    L.zntt_ctx_bytes_in_use.argtypes = [vp]
    L.zntt_ctx_bytes_in_use.restype = u64
    L.zntt_ctx_trim.argtypes = [vp]
    L.zntt_ctx_reserve.argtypes = [vp, ctypes.c_size_t]
    L.zntt_status_string.argtypes = [i32]
    L.zntt_status_string.restype = ctypes.c_char_p
    L.zntt_element_size.argtypes = [i32]
    L.zntt_element_size.restype = sz
    L.zntt_bit_reverse.argtypes = [vp, i32, vp, sz]
    L.zntt_ntt.argtypes = [vp, i32, vp, sz, sz, vp]
    L.zntt_intt.argtypes = [vp, i32, vp, sz, sz, vp]
    L.zntt_precompute_twiddles.argtypes = [vp, i32, sz, vp, pp, sz]
    L.zntt_free_twiddles.argtypes = [vp, pp, sz]
    L.zntt_ntt_with_twiddles.argtypes = [vp, i32, vp, sz, sz, pp, ctypes.POINTER(sz), sz]
    L.zntt_intt_with_twiddles.argtypes = [vp, i32, vp, sz, sz, pp, ctypes.POINTER(sz), sz]
    L.zntt_coset_ntt.argtypes = [vp, i32, vp, sz, sz, vp, vp]
    L.zntt_ntt_with_twiddles_device.argtypes = [vp, i32, vp, sz, sz, pp, ctypes.POINTER(sz), sz, vp]
    L.zntt_intt_with_twiddles_device.argtypes = [vp, i32, vp, sz, sz, pp, ctypes.POINTER(sz), sz, vp]
    L.zntt_lde.argtypes = [vp, i32, vp, sz, sz, vp, vp]
    L.zntt_primitive_root.argtypes = [i32, sz, vp]
    L.zntt_execute_device.argtypes = [vp, ctypes.POINTER(Transform), vp]
    L.zntt_execute_device_noalloc.argtypes = [vp, ctypes.POINTER(Transform), vp]
    L.zntt_ctx_synchronize.argtypes = [vp]
    L.zntt_ctx_stream.argtypes = [vp]
    L.zntt_ctx_stream.restype = vp
    L.zntt_debug_field_op.argtypes = [vp, i32, i32, vp, vp, vp, sz]
    L.zntt_probe_peaks.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.zntt_fill_uniform.argtypes = [i32, u64, vp, sz]
    L.zntt_batch_create.argtypes = [vp, ctypes.POINTER(Transform), sz, ctypes.POINTER(vp)]
    L.zntt_batch_execute.argtypes = [vp, vp]
    L.zntt_batch_status.argtypes = [vp, vp]
    L.zntt_batch_autotune.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double)]
    ip = ctypes.POINTER(ctypes.c_int)
    L.zntt_batch_shape.argtypes = [vp, ip, ip, ip, ip, ip]
    L.zntt_batch_set_shape.argtypes = [vp, i32, i32, i32, i32, i32]
    L.zntt_batch_destroy.argtypes = [vp]
    L.zntt_batch_destroy.restype = None
    L.zntt_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    L.zntt_server_submit.argtypes = [vp, ctypes.POINTER(LdeRequest), ctypes.POINTER(u64)]
    L.zntt_server_query.argtypes = [vp, u64, ctypes.POINTER(RequestInfo)]
    L.zntt_server_wait.argtypes = [vp, u64, ctypes.c_double]
    L.zntt_server_cancel.argtypes = [vp, u64]
    L.zntt_server_read.argtypes = [vp, u64, vp, sz]
    L.zntt_server_release.argtypes = [vp, u64]
    L.zntt_server_drain.argtypes = [vp]
    L.zntt_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.zntt_server_sim_pump.argtypes = [vp, ip]
    L.zntt_server_sim_complete.argtypes = [vp, i32]
    L.zntt_server_destroy.argtypes = [vp]
    L.zntt_server_destroy.restype = None
    L.zntt_fill_counter.argtypes = [vp, i32, u64, u64, vp, sz, vp]
    L.zntt_poly_mul.argtypes = [vp, i32, vp, sz, vp, sz, vp, sz, ctypes.POINTER(sz)]
    L.zntt_poly_mul_device.argtypes = [vp, i32, vp, sz, vp, sz, vp, vp]
    L.zntt_merkle_commit.argtypes = [vp, i32, vp, sz, vp, vp, vp]
    L.zntt_merkle_from_hashes.argtypes = [vp, vp, sz, vp, vp, vp]
    L.zntt_fri_proof_sizes.argtypes = [i32, ctypes.POINTER(FriConfig)] + [ctypes.POINTER(sz)] * 4
    L.zntt_fri_prove.argtypes = [vp, i32, ctypes.POINTER(FriConfig), ctypes.POINTER(TranscriptOps), vp, vp, sz,
                                 ctypes.POINTER(FriProof), vp]
    L.zntt_transcript_create.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(vp)]
    L.zntt_transcript_destroy.argtypes = [vp]
    L.zntt_transcript_destroy.restype = None
    L.zntt_transcript_absorb_bytes.argtypes = [vp, ctypes.c_char_p, sz]
    L.zntt_transcript_absorb_bytes.restype = None
    L.zntt_transcript_challenge_bytes.argtypes = [vp, vp, sz]
    L.zntt_transcript_challenge_bytes.restype = None
    L.zntt_blake3.argtypes = [ctypes.c_char_p, sz, vp]
    L.zntt_blake3.restype = None
    L.zntt_ntt_vec8_m31.argtypes = [vp, vp, sz, sz, vp]
    L.zntt_intt_vec8_m31.argtypes = [vp, vp, sz, sz, vp]
    L.zntt_vec8_m31_device.argtypes = [vp, i32, vp, sz, vp, vp]
    return L


lib = _load()


def check(status: int) -> None:
    if status:
        cls = _ERRORS.get(status, ZnttError)
        raise cls(lib.zntt_status_string(status).decode())


def element_size(field: int) -> int:
    return int(lib.zntt_element_size(field))


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"] and a.dtype == np.uint64
    return a.ctypes.data_as(ctypes.c_void_p)


def _elem(field: int, x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64).reshape(-1))


def empty(field: int, n: int) -> np.ndarray:
    return np.zeros((n, 4) if field == FIELD_FR else (n,), dtype=np.uint64)


def fill_uniform(field: int, seed: int, n: int) -> np.ndarray:
    """Seeded uniform elements in storage layout (host-side generator)."""
    out = empty(field, n)
    check(lib.zntt_fill_uniform(field, seed, _ptr(out), n))
    return out


FR_MODULUS = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
GL_MODULUS = 0xFFFFFFFF00000001
BB_MODULUS = 0x78000001
KB_MODULUS = 0x7F000001
MODULUS = {FIELD_FR: FR_MODULUS, FIELD_GL: GL_MODULUS, FIELD_BB: BB_MODULUS, FIELD_KB: KB_MODULUS}


def from_int(field: int, value: int) -> np.ndarray:
    """One element in storage layout: Fr in Montgomery form (value * 2^256 mod r)."""
    if field == FIELD_FR:
        m = ((value % FR_MODULUS) << 256) % FR_MODULUS
        return np.array([(m >> (64 * i)) & ((1 << 64) - 1) for i in range(4)], dtype=np.uint64)
    return np.array([value % MODULUS[field]], dtype=np.uint64)


def primitive_root(field: int, log_size: int) -> np.ndarray:
    """Field's primitive 2^log_size-th root (libs/field/src/roots.zig:149-157)."""
    out = empty(field, 1).reshape(-1)
    check(lib.zntt_primitive_root(field, log_size, _ptr(out)))
    return out


def blake3(data: bytes) -> bytes:
    """BLAKE3 of `data` (32 bytes), the library's host implementation."""
    out = (ctypes.c_uint8 * 32)()
    lib.zntt_blake3(bytes(data), len(data), out)
    return bytes(out)


class Transcript:
    """The library's Fiat-Shamir transcript (zig-transcript's byte stream)."""

    def __init__(self, label: bytes):
        h = ctypes.c_void_p()
        check(lib.zntt_transcript_create(bytes(label), len(label), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and lib is not None:
            lib.zntt_transcript_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def absorb_bytes(self, data: bytes) -> None:
        lib.zntt_transcript_absorb_bytes(self._h, bytes(data), len(data))

    def challenge_bytes(self, n: int) -> bytes:
        out = (ctypes.c_uint8 * max(n, 1))()
        lib.zntt_transcript_challenge_bytes(self._h, out, n)
        return bytes(out)[:n]

    def ops(self):
        """(ops, handle) for zntt_fri_prove: the library's own functions."""
        return (TranscriptOps(ctypes.cast(lib.zntt_transcript_absorb_bytes, BYTES_FN),
                              ctypes.cast(lib.zntt_transcript_challenge_bytes, BYTES_FN)), self._h)


def callback_ops(obj):
    """(ops, handle, keepalive) that route zntt_fri_prove's transcript calls to
    any Python object with absorb_bytes(bytes) and challenge_bytes(n) -> bytes."""
    def absorb(_, data, n):
        obj.absorb_bytes(ctypes.string_at(data, n))

    def challenge(_, out, n):
        b = obj.challenge_bytes(n)
        ctypes.memmove(out, b, n)

    fa, fc = BYTES_FN(absorb), BYTES_FN(challenge)
    return TranscriptOps(fa, fc), None, (fa, fc)


def fri_proof_sizes(field: int, log_domain: int, log_initial_degree: int, log_final: int,
                    log_residual_degree: int, num_queries: int) -> dict:
    cfg = FriConfig(log_domain, log_initial_degree, log_final, log_residual_degree, num_queries)
    v = [ctypes.c_size_t() for _ in range(4)]
    check(lib.zntt_fri_proof_sizes(field, ctypes.byref(cfg), *[ctypes.byref(x) for x in v]))
    return dict(zip(["roots", "residual", "values", "paths"], [x.value for x in v]))


class Context:
    """One GPU: a HIP stream, scratch memory and the device twiddle cache.

    Batches and servers made on a context hold a reference to it in the
    library, so they stay valid whatever order handles are released in;
    close() also closes those still open, so that their device memory is
    returned at a known point rather than whenever they are collected."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = ctypes.c_void_p()
        check(lib.zntt_ctx_create(ctypes.byref(h), device, hbm_budget_bytes))
        self._h = h
        self._children = weakref.WeakSet()

    @property
    def handle(self):
        return self._h

    def _adopt(self, child) -> None:
        self._children.add(child)

    def close(self):
        if self._h:
            for child in list(self._children):
                child.close()
            lib.zntt_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes_in_use(self) -> int:
        return int(lib.zntt_ctx_bytes_in_use(self._h))

    def trim(self) -> None:
        check(lib.zntt_ctx_trim(self._h))

    def reserve(self, max_log_n: int) -> None:
        """Allocates up front what the seven zig-ntt functions use up to
        2^max_log_n, so that those calls allocate no device memory after it
        (zntt_ctx_reserve)."""
        check(lib.zntt_ctx_reserve(self._h, max_log_n))

    def synchronize(self) -> None:
        check(lib.zntt_ctx_synchronize(self._h))

    # ---- the seven zig-ntt functions (root.zig:38-212) ---------------------
    def bit_reverse(self, field: int, data: np.ndarray) -> None:
        check(lib.zntt_bit_reverse(self._h, field, _ptr(data), data.shape[0]))

    def ntt(self, field: int, data: np.ndarray, log_n: int, root) -> None:
        check(lib.zntt_ntt(self._h, field, _ptr(data), data.shape[0], log_n, _ptr(_elem(field, root))))

    def intt(self, field: int, data: np.ndarray, log_n: int, root) -> None:
        check(lib.zntt_intt(self._h, field, _ptr(data), data.shape[0], log_n, _ptr(_elem(field, root))))

    def precompute_twiddles(self, field: int, log_n: int, root) -> "Twiddles":
        if log_n >= 64:
            check(lib.zntt_precompute_twiddles(self._h, field, log_n, _ptr(_elem(field, root)), None, log_n))
        stages = [empty(field, 1 << s) for s in range(log_n)]
        ptrs = (ctypes.c_void_p * max(log_n, 1))(*[s.ctypes.data for s in stages])
        check(lib.zntt_precompute_twiddles(self._h, field, log_n, _ptr(_elem(field, root)),
                                           ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p)), log_n))
        return Twiddles(self, field, stages, registered=True)

    def free_twiddles(self, tw: "Twiddles") -> None:
        tw.free()

    def ntt_with_twiddles(self, field: int, data: np.ndarray, log_n: int, twiddles) -> None:
        self._twiddled(lib.zntt_ntt_with_twiddles, field, data, log_n, twiddles)

    def intt_with_twiddles(self, field: int, data: np.ndarray, log_n: int, twiddles) -> None:
        self._twiddled(lib.zntt_intt_with_twiddles, field, data, log_n, twiddles)

    def _twiddled(self, fn, field, data, log_n, twiddles):
        ptrs, lens, k = self._table_args(twiddles)
        check(fn(self._h, field, _ptr(data), data.shape[0], log_n, ptrs, lens, k))

    @staticmethod
    def _table_args(twiddles):
        stages = twiddles.stages if isinstance(twiddles, Twiddles) else [np.ascontiguousarray(s) for s in twiddles]
        k = len(stages)
        ptrs = (ctypes.c_void_p * max(k, 1))(*[s.ctypes.data for s in stages])
        lens = (ctypes.c_size_t * max(k, 1))(*[s.shape[0] for s in stages])
        return ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p)), lens, k

    def with_twiddles_device(self, field: int, data_ptr: int, log_n: int, twiddles, inverse: bool = False,
                             stream: int | None = None, length: int | None = None) -> None:
        """nttWithTwiddles / inttWithTwiddles on 2^log_n device elements at
        data_ptr, enqueued on `stream` (None: the context's own stream)."""
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        ptrs, lens, k = self._table_args(twiddles)
        fn = lib.zntt_intt_with_twiddles_device if inverse else lib.zntt_ntt_with_twiddles_device
        n = (1 << log_n) if length is None else length
        check(fn(self._h, field, data_ptr, n, log_n, ptrs, lens, k, stream))

    # ---- beyond zig-ntt ----------------------------------------------------
    def coset_ntt(self, field: int, data: np.ndarray, log_n: int, root, shift) -> None:
        check(lib.zntt_coset_ntt(self._h, field, _ptr(data), data.shape[0], log_n,
                                 _ptr(_elem(field, root)), _ptr(_elem(field, shift))))

    def poly_mul(self, field: int, a: np.ndarray, b: np.ndarray, cap: int | None = None) -> np.ndarray:
        """Polynomial.mul (libs/poly/src/poly.zig:197-212) by NTT; `cap` is
        the reference's max_degree + 1 (default: big enough)."""
        cap = a.shape[0] + b.shape[0] if cap is None else cap
        out = empty(field, max(cap, 1))
        n = ctypes.c_size_t()
        check(lib.zntt_poly_mul(self._h, field, _ptr(np.ascontiguousarray(a)), a.shape[0],
                                _ptr(np.ascontiguousarray(b)), b.shape[0], _ptr(out), cap, ctypes.byref(n)))
        return out[:n.value].copy()

    def poly_mul_device(self, field: int, a_ptr: int, a_len: int, b_ptr: int, b_len: int, out_ptr: int,
                        stream: int | None = None) -> None:
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        check(lib.zntt_poly_mul_device(self._h, field, a_ptr, a_len, b_ptr, b_len, out_ptr, stream))

    def merkle_commit_device(self, field: int, evals_ptr: int, log_n: int, nodes_ptr: int | None = None,
                             stream: int | None = None) -> bytes:
        """Root of the FRI layer commitment of 2^log_n device evaluations;
        with nodes_ptr the whole tree (2^log_n nodes of 32 bytes) lands there."""
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        root = (ctypes.c_uint8 * 32)()
        check(lib.zntt_merkle_commit(self._h, field, evals_ptr, log_n, nodes_ptr, root, stream))
        return bytes(root)

    def merkle_from_hashes_device(self, hashes_ptr: int, count: int, nodes_ptr: int | None = None,
                                  stream: int | None = None) -> bytes:
        """Root of MerkleTree.initFromHashes over `count` device leaf hashes."""
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        root = (ctypes.c_uint8 * 32)()
        check(lib.zntt_merkle_from_hashes(self._h, hashes_ptr, count, nodes_ptr, root, stream))
        return bytes(root)

    def fri_prove(self, field: int, transcript, evals_ptr: int, evals_len: int, log_domain: int,
                  log_initial_degree: int, log_final: int, log_residual_degree: int, num_queries: int,
                  stream: int | None = None) -> dict:
        """FRI prove over device evaluations; `transcript` is a Transcript or any
        object with absorb_bytes / challenge_bytes (called back)."""
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        cfg = FriConfig(log_domain, log_initial_degree, log_final, log_residual_degree, num_queries)
        sz = fri_proof_sizes(field, log_domain, log_initial_degree, log_final, log_residual_degree, num_queries)
        bufs = {k: (ctypes.c_uint8 * max(v, 1))() for k, v in sz.items()}
        pair = (ctypes.c_uint64 * num_queries)()
        proof = FriProof(ctypes.cast(bufs["roots"], ctypes.c_void_p), ctypes.cast(bufs["residual"], ctypes.c_void_p),
                         ctypes.cast(pair, ctypes.c_void_p), ctypes.cast(bufs["values"], ctypes.c_void_p),
                         ctypes.cast(bufs["paths"], ctypes.c_void_p))
        if isinstance(transcript, Transcript):
            ops, handle = transcript.ops()
            keep = None
        else:
            ops, handle, keep = callback_ops(transcript)
        check(lib.zntt_fri_prove(self._h, field, ctypes.byref(cfg), ctypes.byref(ops), handle, evals_ptr,
                                 evals_len, ctypes.byref(proof), stream))
        del keep
        out = {k: bytes(bufs[k])[:sz[k]] for k in sz}
        out["pair_index"] = [int(x) for x in pair]
        return out

    def ntt_vec8_m31(self, data: np.ndarray, log_n: int, root: int, inverse: bool = False,
                     length: int | None = None) -> None:
        """Vec8NttM31 in place on interleaved u64 values data[i * 8 + lane]."""
        assert data.dtype == np.uint64 and data.flags["C_CONTIGUOUS"]
        r = np.array([root], dtype=np.uint64)
        fn = lib.zntt_intt_vec8_m31 if inverse else lib.zntt_ntt_vec8_m31
        check(fn(self._h, _ptr(data), data.shape[0] if length is None else length, log_n, _ptr(r)))

    def vec8_m31_device(self, data_ptr: int, log_n: int, root: int, inverse: bool = False,
                        stream: int | None = None) -> None:
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        r = np.array([root], dtype=np.uint64)
        check(lib.zntt_vec8_m31_device(self._h, 1 if inverse else 0, data_ptr, log_n, _ptr(r), stream))

    def lde(self, field: int, coeffs: np.ndarray, log_k: int, log_blowup: int, shift) -> np.ndarray:
        out = empty(field, 1 << (log_k + log_blowup))
        check(lib.zntt_lde(self._h, field, _ptr(coeffs), log_k, log_blowup, _ptr(_elem(field, shift)), _ptr(out)))
        return out

    def execute_device(self, field: int, direction: int, log_n: int, data_ptr: int, root, shift=None,
                       stream: int | None = None) -> None:
        """Transform 2^log_n elements already in device memory at data_ptr,
        enqueued on `stream` (a hipStream_t as int, 0 = HIP's null stream;
        None = the context's own stream)."""
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        r = _elem(field, root)
        s = None if shift is None else _elem(field, shift)
        t = Transform(field, direction, log_n, data_ptr, r.ctypes.data, None if s is None else s.ctypes.data)
        check(lib.zntt_execute_device(self._h, ctypes.byref(t), stream))

    def execute_device_noalloc(self, field: int, direction: int, log_n: int, data_ptr: int, root, shift=None,
                               stream: int | None = None) -> None:
        """execute_device without any device allocation: the tables must be
        cached already (NotPrepared otherwise) and intermediate passes use the
        library's static per-GPU scratch (OutOfMemory beyond it)."""
        if stream is None:
            stream = lib.zntt_ctx_stream(self._h)
        r = _elem(field, root)
        s = None if shift is None else _elem(field, shift)
        t = Transform(field, direction, log_n, data_ptr, r.ctypes.data, None if s is None else s.ctypes.data)
        check(lib.zntt_execute_device_noalloc(self._h, ctypes.byref(t), stream))

    def debug_field_op(self, field: int, op: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        out = np.zeros_like(a)
        check(lib.zntt_debug_field_op(self._h, field, op, _ptr(a), _ptr(b), _ptr(out), a.shape[0]))
        return out

    def probe_peaks(self) -> tuple[float, float]:
        mad, copy = ctypes.c_double(), ctypes.c_double()
        check(lib.zntt_probe_peaks(self._h, ctypes.byref(mad), ctypes.byref(copy)))
        return mad.value, copy.value


class Twiddles:
    """The zig `[]const []const F` table: host stages plus, when registered,
    the device-resident copy held by the context until free()."""

    def __init__(self, ctx: Context, field: int, stages: list[np.ndarray], registered: bool):
        self.ctx, self.field, self.stages, self.registered = ctx, field, stages, registered

    def free(self) -> None:
        if self.registered and self.ctx._h:
            k = len(self.stages)
            ptrs = (ctypes.c_void_p * max(k, 1))(*[s.ctypes.data for s in self.stages])
            check(lib.zntt_free_twiddles(self.ctx._h, ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p)), k))
            self.registered = False

    def __len__(self):
        return len(self.stages)

    def __getitem__(self, i):
        return self.stages[i]


class Batch:
    """Many device transforms (uneven sizes, mixed directions) run as one
    persistent launch, replayed from a HIP graph (zntt_batch_*).

    `transforms` is a list of (direction, log_n, data_ptr, root[, shift]) for
    one field; data_ptr are device pointers that stay the batch's buffers."""

    def __init__(self, ctx: Context, field: int, transforms):
        self.ctx = ctx
        self._keep = []
        arr = (Transform * len(transforms))()
        for i, t in enumerate(transforms):
            direction, log_n, ptr, root = t[:4]
            shift = t[4] if len(t) > 4 else None
            r = _elem(field, root)
            s = None if shift is None else _elem(field, shift)
            self._keep += [r, s]
            arr[i] = Transform(field, direction, log_n, ptr, r.ctypes.data, None if s is None else s.ctypes.data)
        h = ctypes.c_void_p()
        check(lib.zntt_batch_create(ctx.handle, arr, len(transforms), ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)

    def execute(self, stream: int | None = None) -> None:
        if stream is None:
            stream = lib.zntt_ctx_stream(self.ctx.handle)
        check(lib.zntt_batch_execute(self._h, stream))

    def status(self, stream: int | None = None) -> None:
        if stream is None:
            stream = lib.zntt_ctx_stream(self.ctx.handle)
        check(lib.zntt_batch_status(self._h, stream))

    def autotune(self, reps: int = 3) -> float:
        ms = ctypes.c_double()
        check(lib.zntt_batch_autotune(self._h, reps, ctypes.byref(ms)))
        return ms.value

    def shape(self) -> dict:
        v = [ctypes.c_int() for _ in range(5)]
        check(lib.zntt_batch_shape(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("tile_log", "kmax", "wgs_per_cu", "order", "level_kernel"), [x.value for x in v]))

    def set_shape(self, tile_log: int, kmax: int, wgs_per_cu: int, order: int, level_kernel: int) -> None:
        check(lib.zntt_batch_set_shape(self._h, tile_log, kmax, wgs_per_cu, order, level_kernel))

    def close(self) -> None:
        if self._h:
            lib.zntt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_counter_device(ctx: "Context | None", field: int, seed: int, stream_id: int, data_ptr: int, n: int,
                        stream: int | None = None) -> None:
    """Device fill with the counter-based stream of generated requests."""
    if stream is None:
        stream = lib.zntt_ctx_stream(ctx.handle)
    check(lib.zntt_fill_counter(ctx.handle, field, seed, stream_id, data_ptr, n, stream))


class Server:
    """Serving loop for coset LDEs (zntt_server_*): requests are admitted
    against an HBM budget and run by continuous batching, one scheduler step
    (one persistent launch holding the next pass of every admitted request)
    at a time.

    With simulate=True nothing runs on a device (ctx may be None): the test
    drives the scheduler with sim_pump() and completes steps with
    sim_complete(), to check admission and failure handling on a CPU host.
    `allocator` is an optional object with alloc(nbytes) -> address (0 or an
    exception for failure) and free(address, nbytes)."""

    def __init__(self, ctx: "Context | None", field: int = FIELD_FR, hbm_budget: int = 0, max_active: int = 0,
                 max_log_n: int = 0, simulate: bool = False, allocator=None):
        self.ctx = ctx
        self.field = field
        self._alloc_ref = None
        alloc_p = None
        if allocator is not None:
            def c_alloc(_user, nbytes, out):
                try:
                    addr = int(allocator.alloc(int(nbytes)) or 0)
                except MemoryError:
                    return OutOfMemory.status
                except Exception:
                    return DeviceError.status
                out[0] = addr
                return 0 if addr else OutOfMemory.status

            def c_free(_user, addr, nbytes):
                allocator.free(int(addr or 0), int(nbytes))

            a = Allocator(None, ALLOC_FN(c_alloc), FREE_FN(c_free))
            self._alloc_ref = a  # the callbacks must outlive the server
            alloc_p = ctypes.pointer(a)
        cfg = ServerConfig(field, hbm_budget, max_active, max_log_n, 1 if simulate else 0, alloc_p)
        h = ctypes.c_void_p()
        check(lib.zntt_server_create(None if ctx is None else ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        if ctx is not None:
            ctx._adopt(self)
        self._shift_keep = {}

    def submit(self, log_n: int, log_blowup: int, shift, coeffs_ptr: int | None = None, seed: int = 0,
               stream_id: int = 0, keep: bool = False, length: int | None = None,
               out_ptr: int | None = None) -> int:
        sh = _elem(self.field, shift)
        key = sh.tobytes()
        sh = self._shift_keep.setdefault(key, sh)
        n = (1 << log_n) if length is None else length
        req = LdeRequest(log_n, n, log_blowup, sh.ctypes.data, coeffs_ptr, seed, stream_id, 1 if keep else 0,
                         out_ptr)
        t = ctypes.c_uint64()
        check(lib.zntt_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        return t.value

    def query(self, ticket: int) -> dict:
        info = RequestInfo()
        check(lib.zntt_server_query(self._h, ticket, ctypes.byref(info)))
        return {"state": info.state, "status": info.status, "queued_ms": info.queued_ms,
                "latency_ms": info.latency_ms, "data": info.data or 0, "log_n": info.log_n,
                "log_blowup": info.log_blowup}

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> dict:
        check(lib.zntt_server_wait(self._h, ticket, timeout_ms))
        return self.query(ticket)

    def cancel(self, ticket: int) -> None:
        check(lib.zntt_server_cancel(self._h, ticket))

    def read(self, ticket: int) -> np.ndarray:
        """A finished request's evaluations (kept or caller-owned), copied to the host."""
        q = self.query(ticket)
        n = 1 << (q["log_n"] + q["log_blowup"])
        out = empty(self.field, n)
        check(lib.zntt_server_read(self._h, ticket, _ptr(out), n))
        return out

    def release(self, ticket: int) -> None:
        check(lib.zntt_server_release(self._h, ticket))

    def drain(self) -> None:
        check(lib.zntt_server_drain(self._h))

    def stats(self) -> dict:
        st = ServerStats()
        check(lib.zntt_server_stats_get(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def sim_pump(self) -> int:
        n = ctypes.c_int()
        check(lib.zntt_server_sim_pump(self._h, ctypes.byref(n)))
        return n.value

    def sim_complete(self, n: int = 1) -> None:
        check(lib.zntt_server_sim_complete(self._h, n))

    def close(self) -> None:
        if self._h:
            lib.zntt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
