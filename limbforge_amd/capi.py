"""ctypes binding of liblimbforge_hip.so (include/limbforge_hip.h) for Python callers.

This is a thin host-side mirror of the reference's Engine interface (include/limbforge/engine.hpp):
`Engine.run / run_unary / run_ternary` on host arrays and resident `Buffer` operands with
`Batch.run` / `Batch.mul_chain` / `Batch.submit`, with the same argument meaning and error classes
(ValueError for std::invalid_argument, RuntimeError for logic/runtime errors).  Values are numpy
uint32 arrays of shape (count, n+3) (real) or (count, 2*(n+3)) (complex): each row is the public
element record (n limbs, exponent, sign, status) of engine.hpp:195.

The library must be built (python -m limbforge_amd.build); loading fails loudly otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "liblimbforge_hip.so"

OPS = {
    "add": 0, "sub": 1, "mul": 2, "div": 3, "complex_add": 4, "complex_mul": 5, "complex_div": 6,
    "square": 7, "sqrt": 8, "fma": 9, "fms": 10, "complex_fma": 11, "complex_fms": 12,
}
COMPLEX_OPS = {4, 5, 6, 11, 12}
UNARY_OPS = {7, 8}
TERNARY_OPS = {9, 10, 11, 12}


class Operand(C.Structure):
    _fields_ = [("planes", C.c_void_p), ("pitch", C.c_uint64)]


class Broadcast(C.Structure):
    _fields_ = [("stride", C.c_uint64), ("period", C.c_uint64)]


class LimbForgeLogicError(RuntimeError):
    """std::logic_error of the reference API."""


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m limbforge_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    P = C.POINTER
    sig = {
        "lf_last_error": (C.c_char_p, []),
        "lf_library_version": (None, [P(C.c_uint), P(C.c_uint), P(C.c_uint)]),
        "lf_device_count": (i32, [P(i32)]),
        "lf_device_set": (i32, [i32]),
        "lf_device_name": (i32, [C.c_char_p, C.c_size_t]),
        "lf_malloc": (i32, [P(vp), C.c_size_t]),
        "lf_free": (i32, [vp]),
        "lf_host_alloc": (i32, [P(vp), C.c_size_t]),
        "lf_host_free": (i32, [vp]),
        "lf_memcpy_async": (i32, [vp, vp, C.c_size_t, i32, vp]),
        "lf_stream_create": (i32, [P(vp)]),
        "lf_stream_destroy": (i32, [vp]),
        "lf_stream_synchronize": (i32, [vp]),
        "lf_event_create": (i32, [P(vp)]),
        "lf_event_destroy": (i32, [vp]),
        "lf_event_record": (i32, [vp, vp]),
        "lf_event_synchronize": (i32, [vp]),
        "lf_event_query": (i32, [vp, P(i32)]),
        "lf_event_elapsed_ms": (i32, [P(C.c_float), vp, vp]),
        "lf_graph_begin_capture": (i32, [vp]),
        "lf_graph_end_capture": (i32, [vp, P(vp)]),
        "lf_graph_launch": (i32, [vp, vp]),
        "lf_graph_destroy": (i32, [vp]),
        "lf_plane_pitch": (u64, [u64]),
        "lf_planes_bytes": (u64, [i32, i32, u64]),
        "lf_aos_to_planes": (i32, [i32, i32, vp, P(Operand), u64, vp]),
        "lf_planes_to_aos": (i32, [i32, i32, P(Operand), vp, u64, vp]),
        "lf_elementwise": (i32, [i32, i32, P(Operand), P(Operand), P(Operand), P(Operand), u64, P(Broadcast), P(Broadcast), vp]),
        "lf_mul_chain": (i32, [i32, P(Operand), P(Operand), P(Operand), u64, P(Broadcast), i32, vp]),
        "lf_mul_chain_lanes": (i32, [i32, P(Operand), P(Operand), P(Operand), u64, P(Broadcast), i32, C.c_uint, vp]),
        "lf_autotune_chain": (i32, [i32, u64, i32, P(C.c_uint), P(C.c_float)]),
        "lf_calibrate_mad64": (i32, [u64, i32, P(C.c_double)]),
        "lf_calibrate_copy": (i32, [u64, P(C.c_double)]),
        "lf_factor_step_cycles": (i32, [i32, i32, P(C.c_uint64)]),
        "lf_cast": (i32, [i32, i32, i32, P(Operand), P(Operand), u64, vp]),
        "lf_cast_host": (i32, [i32, i32, i32, vp, vp, u64, P(C.c_double), P(C.c_double)]),
        "lf_run_host": (i32, [i32, i32, vp, vp, vp, vp, u64, P(Broadcast), P(Broadcast), P(C.c_double), P(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc == 0:
        return
    msg = lib().lf_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    if rc == 2:
        raise LimbForgeLogicError(msg)
    raise RuntimeError(msg)


def words(bits: int) -> int:
    return bits // 32


def record_words(bits: int, complex_values: bool = False) -> int:
    return (bits // 32 + 3) * (2 if complex_values else 1)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().lf_device_count(C.byref(n)))
    return n.value


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().lf_device_name(buf, 256))
    return buf.value.decode()


class DeviceMemory:
    """Owned device allocation (lf_malloc / lf_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().lf_malloc(C.byref(p), max(self.nbytes, 4)))
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            check(lib().lf_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, host: np.ndarray, stream=None) -> None:
        host = np.ascontiguousarray(host)
        assert host.nbytes <= self.nbytes
        check(lib().lf_memcpy_async(self.ptr, host.ctypes.data, host.nbytes, 0, stream))
        check(lib().lf_stream_synchronize(stream))

    def download(self, host: np.ndarray, stream=None) -> np.ndarray:
        assert host.flags.c_contiguous and host.nbytes <= self.nbytes
        check(lib().lf_memcpy_async(host.ctypes.data, self.ptr, host.nbytes, 1, stream))
        check(lib().lf_stream_synchronize(stream))
        return host


class Buffer:
    """A resident operand in limb planes (the reference's Buffer<Float<bits>> / Buffer<Complex<n>>)."""

    def __init__(self, bits: int, count: int, complex_values: bool = False):
        self.bits, self.count, self.complex = bits, int(count), bool(complex_values)
        self.pitch = int(lib().lf_plane_pitch(self.count))
        self.mem = DeviceMemory(int(lib().lf_planes_bytes(bits, int(self.complex), self.count)))

    @property
    def operand(self) -> Operand:
        return Operand(self.mem.ptr, self.pitch)

    def upload(self, records: np.ndarray, stream=None) -> None:
        """Buffer::upload (engine.hpp:84): AoS records -> device -> limb planes."""
        rec = np.ascontiguousarray(records, dtype=np.uint32)
        assert rec.shape == (self.count, record_words(self.bits, self.complex)), rec.shape
        if not self.count:
            return
        stage = DeviceMemory(rec.nbytes)
        stage.upload(rec, stream)
        op = self.operand
        check(lib().lf_aos_to_planes(self.bits, int(self.complex), stage.ptr, C.byref(op), self.count, stream))
        check(lib().lf_stream_synchronize(stream))
        stage.free()

    def download(self, stream=None) -> np.ndarray:
        """Buffer::download (engine.hpp:85): limb planes -> AoS records on the host."""
        out = np.zeros((self.count, record_words(self.bits, self.complex)), dtype=np.uint32)
        if not self.count:
            return out
        stage = DeviceMemory(out.nbytes)
        op = self.operand
        check(lib().lf_planes_to_aos(self.bits, int(self.complex), C.byref(op), stage.ptr, self.count, stream))
        stage.download(out, stream)
        stage.free()
        return out


def _bcast(b):
    if b is None:
        return None
    stride, period = b
    return C.byref(Broadcast(stride, period))


def elementwise(bits: int, op, a: Buffer, b: Buffer | None, c: Buffer | None, out: Buffer,
                b_index=None, c_index=None, stream=None) -> None:
    """CommandBatch::run (engine.hpp:120-147) on resident buffers, enqueued on `stream`."""
    code = OPS[op] if isinstance(op, str) else int(op)
    oa, oo = a.operand, out.operand
    ob = b.operand if b is not None else oa
    oc = c.operand if c is not None else oa
    check(lib().lf_elementwise(bits, code, C.byref(oa), C.byref(ob), C.byref(oc), C.byref(oo), a.count,
                               _bcast(b_index), _bcast(c_index), stream))


def mul_chain(bits: int, x: Buffer, f: Buffer, out: Buffer, steps: int, f_index=None, stream=None, lanes: int = 0) -> None:
    """x <- RN(x * f) `steps` times in one launch; lanes per value 1, 2 or 4, 0 = the tuned choice."""
    ox, of, oo = x.operand, f.operand, out.operand
    check(lib().lf_mul_chain_lanes(bits, C.byref(ox), C.byref(of), C.byref(oo), x.count, _bcast(f_index), int(steps),
                                   int(lanes), stream))


def autotune_chain(bits: int, count: int, steps: int = 64):
    """Times the chain at 1, 2 and 4 lanes per value, records the fastest for this width and batch-size
    class, and returns (lanes, [ms at 1, 2, 4 lanes] with -1 where the width has no such form)."""
    lanes = C.c_uint(0)
    ms = (C.c_float * 3)()
    check(lib().lf_autotune_chain(bits, int(count), int(steps), C.byref(lanes), ms))
    return lanes.value, list(ms)


def factor_step_cycles(bits: int, reps: int = 32) -> dict:
    """lf_factor_step_cycles: s_memtime cycles of the factorization's per-row steps on one wavefront."""
    out = (C.c_uint64 * 6)()
    check(lib().lf_factor_step_cycles(bits, int(reps), out))
    keys = ("exact_dot_16_terms", "exact_dot_31_terms", "div_by_reciprocal", "knuth_div", "sqrt", "reciprocal")
    return dict(zip(keys, [int(v) for v in out]))


def broadcast_elements(count: int, index) -> int:
    """Elements an operand read through Broadcast{stride, period} must hold (engine.hpp:22-25)."""
    if index is None:
        return count
    stride, period = index
    if stride <= 0:
        raise ValueError("broadcast stride must be positive")
    n = count // stride + (count % stride != 0)
    return period if period and period < n else n


def run_host(bits: int, op, a: np.ndarray, b: np.ndarray | None = None, c: np.ndarray | None = None,
             b_index=None, c_index=None):
    """Engine::run / run_unary / run_ternary (engine.hpp:202-204): host arrays in, host array out.
    Returns (out, gpu_seconds, wall_seconds).  b and c must hold broadcast_elements(count, index) records of
    the operation's width (ValueError otherwise), as the reference checks its buffer sizes."""
    code = OPS[op] if isinstance(op, str) else int(op)
    rw = record_words(bits, code in COMPLEX_OPS)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    if a.ndim != 2 or a.shape[1] != rw:
        raise ValueError(f"a must be an array of {rw}-word records")
    count = a.shape[0]
    out = np.zeros_like(a)

    def operand(x, index, name, used):
        if not used:
            return None
        if x is None:
            raise ValueError(f"operation {op} needs operand {name}")
        x = np.ascontiguousarray(x, dtype=np.uint32)
        need = broadcast_elements(count, index)
        if x.ndim != 2 or x.shape[1] != rw or x.shape[0] < need:
            raise ValueError(f"operand {name} must hold {need} records of {rw} words")
        return x

    bb = operand(b, b_index, "b", code not in UNARY_OPS)
    cc = operand(c, c_index, "c", code in TERNARY_OPS)
    g, w = C.c_double(0), C.c_double(0)
    check(lib().lf_run_host(bits, code, a.ctypes.data, bb.ctypes.data if bb is not None else None,
                            cc.ctypes.data if cc is not None else None, out.ctypes.data, count,
                            _bcast(b_index), _bcast(c_index), C.byref(g), C.byref(w)))
    return out, g.value, w.value


def _reduce_sigs(L):
    if getattr(L, "_reduce", False):
        return
    P = C.POINTER
    L.lf_segmented_dot.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_int, P(Operand), P(Operand), P(Operand), C.c_void_p]
    L.lf_segmented_dot.restype = C.c_int
    L.lf_segmented_dot_host.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        P(C.c_double), P(C.c_double)]
    L.lf_segmented_dot_host.restype = C.c_int
    L.lf_tree_sum.argtypes = [C.c_int, C.c_int, P(Operand), C.c_uint64, P(Operand), C.c_void_p]
    L.lf_tree_sum.restype = C.c_int
    L._reduce = True


def segmented_dot(bits: int, segments: int, length: int, a: Buffer, b: Buffer, out: Buffer, shared_right=False, stream=None) -> None:
    """CommandBatch::segmented_dot (engine.hpp:177-180) on resident buffers."""
    L = lib()
    _reduce_sigs(L)
    oa, ob, oo = a.operand, b.operand, out.operand
    check(L.lf_segmented_dot(bits, int(out.complex), segments, length, int(shared_right), C.byref(oa), C.byref(ob), C.byref(oo),
                             stream))


def segmented_dot_host(bits: int, complex_values: bool, segments: int, length: int, a: np.ndarray, b: np.ndarray,
                       shared_right=False):
    """Engine::segmented_dot (engine.hpp:205) on host arrays: returns (out, gpu_seconds, wall_seconds)."""
    L = lib()
    _reduce_sigs(L)
    rw = record_words(bits, complex_values)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    if a.shape != (segments * length, rw) or b.shape != ((length if shared_right else segments * length), rw):
        raise ValueError("operand shapes do not match the segmented dot shape")
    out = np.zeros((segments, rw), np.uint32)
    g, w = C.c_double(), C.c_double()
    check(L.lf_segmented_dot_host(bits, int(complex_values), segments, length, int(shared_right), a.ctypes.data, b.ctypes.data,
                                  out.ctypes.data, C.byref(g), C.byref(w)))
    return out, g.value, w.value


def tree_sum(bits: int, x: Buffer, count: int, out: Buffer, stream=None) -> None:
    """CommandBatch::tree_sum (engine.hpp:150-153) on resident buffers."""
    L = lib()
    _reduce_sigs(L)
    ox, oo = x.operand, out.operand
    check(L.lf_tree_sum(bits, int(out.complex), C.byref(ox), count, C.byref(oo), stream))


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().lf_stream_create(C.byref(s)))
        self.handle = s.value

    def synchronize(self) -> None:
        check(lib().lf_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().lf_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib().lf_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None) -> None:
        check(lib().lf_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def synchronize(self) -> None:
        check(lib().lf_event_synchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        check(lib().lf_event_elapsed_ms(C.byref(ms), self.handle, end.handle))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().lf_event_destroy(self.handle)
        except Exception:
            pass


# ------------------------------------------------------------------ dense products
class LinalgOptions(C.Structure):
    """LinalgOptions (linalg.hpp:18-25) plus method: 0 residue GEMM (int8 MFMA + CRT), 1 plain limb GEMM."""
    _fields_ = [("band_bits", C.c_int), ("max_bands", C.c_int), ("max_spread", C.c_int), ("host_threads", C.c_uint),
                ("force_fallback", C.c_int), ("resident_bands", C.c_int), ("method", C.c_int)]

    @classmethod
    def default(cls, **kw):
        o = cls()
        lib().lf_linalg_default_options(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class LinalgPair(C.Structure):
    _fields_ = [("left_band", C.c_uint32), ("right_band", C.c_uint32), ("rows", C.c_uint32), ("cols", C.c_uint32),
                ("moduli", C.c_uint32)]


class LinalgReport(C.Structure):
    _fields_ = [("m", C.c_uint64), ("n", C.c_uint64), ("k", C.c_uint64), ("band_slots", C.c_uint32),
                ("left_bands", C.c_uint64 * 65), ("right_bands", C.c_uint64 * 65),
                ("left_fallback", C.c_uint64), ("right_fallback", C.c_uint64),
                ("gpu_outputs", C.c_uint64), ("multi_band_outputs", C.c_uint64), ("fallback_outputs", C.c_uint64),
                ("trivial_outputs", C.c_uint64), ("npairs", C.c_uint32), ("pairs", LinalgPair * 64),
                ("gemm_rows", C.c_uint64), ("gemm_cols", C.c_uint64), ("int8_gemms", C.c_uint64), ("moduli", C.c_uint64),
                ("crt_words", C.c_uint32), ("analysis_seconds", C.c_double), ("upload_seconds", C.c_double),
                ("gpu_seconds", C.c_double), ("assembly_seconds", C.c_double), ("fallback_seconds", C.c_double),
                ("digits_seconds", C.c_double), ("product_seconds", C.c_double), ("rebuild_seconds", C.c_double),
                ("dot_seconds", C.c_double)]

    def bands(self, side="left"):
        arr = self.left_bands if side == "left" else self.right_bands
        return [int(arr[i]) for i in range(self.band_slots)]


def _linalg_sigs(L):
    P = C.POINTER
    L.lf_linalg_default_options.argtypes = [P(LinalgOptions)]
    L.lf_linalg_default_options.restype = None
    L.lf_gemm_host.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p,
                               C.c_int, P(LinalgOptions), P(LinalgReport), P(C.c_double), P(C.c_double)]
    L.lf_gemm_host.restype = C.c_int
    L.lf_syrk_host.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_int, C.c_int, P(LinalgOptions),
                               P(LinalgReport), P(C.c_double), P(C.c_double)]
    L.lf_syrk_host.restype = C.c_int
    L.lf_gemm.argtypes = [C.c_int, C.c_int, P(Operand), P(Operand), C.c_uint64, C.c_uint64, C.c_uint64, P(Operand), C.c_int,
                          P(LinalgOptions), P(LinalgReport), P(C.c_double), C.c_void_p]
    L.lf_gemm.restype = C.c_int
    L.lf_syrk.argtypes = [C.c_int, P(Operand), C.c_uint64, C.c_uint64, P(Operand), C.c_int, C.c_int, P(LinalgOptions), P(LinalgReport),
                          P(C.c_double), C.c_void_p]
    L.lf_syrk.restype = C.c_int


def gemm_resident(bits, transpose_a, A: "Buffer", B: "Buffer", m, n, k, Cb: "Buffer", subtract=False, options=None, stream=None):
    """The resident product (Linalg::gemm on Buffers, linalg.hpp:187-190): C written in place on the device.
    Returns (report, gpu_seconds)."""
    L = lib()
    _linalg_sigs(L)
    o = options if options is not None else LinalgOptions.default()
    rep, g = LinalgReport(), C.c_double()
    oa, ob, oc = A.operand, B.operand, Cb.operand
    check(L.lf_gemm(bits, int(transpose_a), C.byref(oa), C.byref(ob), m, n, k, C.byref(oc), int(subtract), C.byref(o), C.byref(rep),
                    C.byref(g), stream))
    return rep, g.value


def syrk_resident(bits, A: "Buffer", rows, cols, Cb: "Buffer", lower_only=True, subtract=False, options=None, stream=None):
    """The resident Linalg::syrk on Buffers: C written in place on the device.  Returns (report, gpu_seconds)."""
    L = lib()
    _linalg_sigs(L)
    o = options if options is not None else LinalgOptions.default()
    rep, g = LinalgReport(), C.c_double()
    oa, oc = A.operand, Cb.operand
    check(L.lf_syrk(bits, C.byref(oa), rows, cols, C.byref(oc), int(lower_only), int(subtract), C.byref(o), C.byref(rep), C.byref(g),
                    stream))
    return rep, g.value


def gemm(bits, transpose_a, A, B, m, n, k, C_in=None, subtract=False, options=None):
    """Linalg::gemm (linalg.hpp:153-155) on host arrays: returns (C, report, gpu_seconds, wall_seconds)."""
    L = lib()
    if not getattr(L, "_linalg", False):
        _linalg_sigs(L)
        L._linalg = True
    rw = bits // 32 + 3
    Cm = np.zeros((m * n, rw), np.uint32) if C_in is None else np.array(C_in, dtype=np.uint32).reshape(m * n, rw).copy()
    A = np.ascontiguousarray(A, dtype=np.uint32)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    rep = LinalgReport()
    g, w = C.c_double(), C.c_double()
    check(L.lf_gemm_host(bits, int(transpose_a), A.ctypes.data, B.ctypes.data, m, n, k, Cm.ctypes.data, int(subtract),
                         C.byref(options) if options is not None else None, C.byref(rep), C.byref(g), C.byref(w)))
    return Cm, rep, g.value, w.value


def syrk(bits, A, rows, cols, C_in=None, lower_only=True, subtract=False, options=None):
    """Linalg::syrk (linalg.hpp:149-152) on host arrays: returns (C, report, gpu_seconds, wall_seconds)."""
    L = lib()
    if not getattr(L, "_linalg", False):
        _linalg_sigs(L)
        L._linalg = True
    rw = bits // 32 + 3
    Cm = np.zeros((cols * cols, rw), np.uint32) if C_in is None else np.array(C_in, dtype=np.uint32).reshape(cols * cols, rw).copy()
    A = np.ascontiguousarray(A, dtype=np.uint32)
    rep = LinalgReport()
    g, w = C.c_double(), C.c_double()
    check(L.lf_syrk_host(bits, A.ctypes.data, rows, cols, Cm.ctypes.data, int(lower_only), int(subtract),
                         C.byref(options) if options is not None else None, C.byref(rep), C.byref(g), C.byref(w)))
    return Cm, rep, g.value, w.value


# ---------------------------------------------------------------- Cholesky and triangular solves
class FactorOptions(C.Structure):
    """FactorOptions (linalg.hpp:73-78)."""
    _fields_ = [("block", C.c_uint64), ("host_macs", C.c_double), ("solve_host_macs", C.c_double), ("gpu", C.c_int)]

    @classmethod
    def default(cls, **kw):
        o = cls()
        L = lib()
        _factor_sigs(L)
        L.lf_factor_default_options(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class CholeskyInfo(C.Structure):
    """CholeskyInfo (linalg.hpp:82-87)."""
    _fields_ = [("pivot", C.c_uint64), ("pivot_sign", C.c_int), ("pivot_status", C.c_uint32),
                ("timing_gpu_seconds", C.c_double), ("timing_wall_seconds", C.c_double),
                ("blocks", C.c_uint64), ("gpu_updates", C.c_uint64), ("host_updates", C.c_uint64),
                ("panel_seconds", C.c_double), ("update_seconds", C.c_double)]


def _factor_sigs(L):
    P = C.POINTER
    if getattr(L, "_factor", False):
        return
    if not getattr(L, "_linalg", False):
        _linalg_sigs(L)
        L._linalg = True
    L.lf_factor_default_options.argtypes = [P(FactorOptions)]
    L.lf_factor_default_options.restype = None
    L.lf_cholesky_host.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, P(FactorOptions), P(LinalgOptions), P(CholeskyInfo)]
    L.lf_solve_host.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, P(FactorOptions),
                                P(LinalgOptions), P(C.c_double), P(C.c_double)]
    L._factor = True


def cholesky(bits, A, n, block=32, options=None, linalg=None):
    """Linalg::cholesky (linalg.hpp:157-159) on host arrays: returns (L, CholeskyInfo)."""
    L = lib()
    _factor_sigs(L)
    rw = bits // 32 + 3
    A = np.ascontiguousarray(A, dtype=np.uint32).reshape(n * n, rw)
    Lm = np.zeros((n * n, rw), np.uint32)
    fo = options if options is not None else FactorOptions.default(block=block)
    info = CholeskyInfo()
    check(L.lf_cholesky_host(bits, A.ctypes.data, n, Lm.ctypes.data, C.byref(fo),
                             C.byref(linalg) if linalg is not None else None, C.byref(info)))
    return Lm, info


def solve(bits, passes, Lf, n, B, nrhs, block=32, options=None, linalg=None):
    """Linalg::trsm (passes 1: L X = B, 2: L^T X = B) and cholesky_solve (passes 3), linalg.hpp:160-163, on
    host arrays: returns (X, gpu_seconds, wall_seconds)."""
    L = lib()
    _factor_sigs(L)
    rw = bits // 32 + 3
    Lf = np.ascontiguousarray(Lf, dtype=np.uint32).reshape(n * n, rw)
    B = np.ascontiguousarray(B, dtype=np.uint32).reshape(n * nrhs, rw)
    X = np.zeros((n * nrhs, rw), np.uint32)
    fo = options if options is not None else FactorOptions.default(block=block)
    g, w = C.c_double(), C.c_double()
    check(L.lf_solve_host(bits, int(passes), Lf.ctypes.data, n, B.ctypes.data, nrhs, X.ctypes.data, C.byref(fo),
                          C.byref(linalg) if linalg is not None else None, C.byref(g), C.byref(w)))
    return X, g.value, w.value


# ---------------------------------------------------------------- mixed-precision chain batches
class ChainBucket(C.Structure):
    _fields_ = [("bits", C.c_int), ("x", Operand), ("f", Operand), ("g", Operand), ("out", Operand),
                ("len", C.c_void_p), ("mask", C.c_void_p), ("count", C.c_uint64)]


def _mixed_sigs(L):
    if getattr(L, "_mixed", False):
        return
    P = C.POINTER
    L.lf_mixed_chains.argtypes = [P(ChainBucket), C.c_int, P(C.c_uint), C.c_void_p]
    L.lf_mixed_graph_create.argtypes = [P(ChainBucket), C.c_int, P(C.c_uint), P(C.c_void_p)]
    L.lf_autotune_mixed.argtypes = [C.c_int, C.c_uint64, P(C.c_uint), P(C.c_double)]
    L.lf_mixed_graph_create_lanes.argtypes = [P(ChainBucket), C.c_int, P(C.c_uint), P(C.c_uint), P(C.c_void_p)]
    L.lf_autotune_mixed_lanes.argtypes = [P(ChainBucket), C.c_int, P(C.c_uint), P(C.c_uint), P(C.c_float)]
    L._mixed = True


class MixedBucket:
    """The values of one precision of a mixed batch, resident: x, f, g and out planes, chain lengths and
    step masks (bit t set: step t is fma, else mul)."""

    def __init__(self, bits: int, x: np.ndarray, f: np.ndarray, g: np.ndarray, length: np.ndarray, mask: np.ndarray):
        self.bits, self.count = bits, x.shape[0]
        self.X, self.F, self.G, self.O = (Buffer(bits, self.count) for _ in range(4))
        self.X.upload(x)
        self.F.upload(f)
        self.G.upload(g)
        self.len = DeviceMemory(4 * self.count)
        self.len.upload(np.ascontiguousarray(length, dtype=np.uint32))
        self.mask = DeviceMemory(8 * self.count)
        self.mask.upload(np.ascontiguousarray(mask, dtype=np.uint64))

    def struct(self) -> ChainBucket:
        return ChainBucket(self.bits, self.X.operand, self.F.operand, self.G.operand, self.O.operand, self.len.ptr,
                           self.mask.ptr, self.count)


def _bucket_array(buckets, block):
    arr = (ChainBucket * len(buckets))(*[b.struct() for b in buckets])
    blk = None if block is None else (C.c_uint * len(buckets))(*[int(x) for x in block])
    return arr, blk


def mixed_chains(buckets, block=None, stream=None) -> None:
    """One launch per bucket, in order, on `stream` (lf_mixed_chains)."""
    L = lib()
    _mixed_sigs(L)
    arr, blk = _bucket_array(buckets, block)
    check(L.lf_mixed_chains(arr, len(buckets), blk, stream))


def mixed_graph(buckets, block=None, lanes=None) -> int:
    """A graph with one kernel node per group of widths (lf_mixed_graph_create_lanes); launch with
    graph_launch.  lanes: lanes per value of the four groups (None = the tuned choice, else one lane)."""
    L = lib()
    _mixed_sigs(L)
    arr, blk = _bucket_array(buckets, block)
    g = C.c_void_p()
    gl = None if lanes is None else (C.c_uint * 4)(*[int(x) for x in lanes])
    check(L.lf_mixed_graph_create_lanes(arr, len(buckets), blk, gl, C.byref(g)))
    return g.value


def autotune_mixed_lanes(buckets, block=None):
    """lf_autotune_mixed_lanes: the fastest lanes per value of each group for this batch's graph (cached per
    device); returns ([lanes of the four groups], ms)."""
    L = lib()
    _mixed_sigs(L)
    arr, blk = _bucket_array(buckets, block)
    gl = (C.c_uint * 4)()
    ms = C.c_float()
    check(L.lf_autotune_mixed_lanes(arr, len(buckets), blk, gl, C.byref(ms)))
    return list(gl), ms.value


def graph_launch(graph, stream=None) -> None:
    check(lib().lf_graph_launch(graph, stream))


def graph_destroy(graph) -> None:
    check(lib().lf_graph_destroy(graph))


def autotune_mixed(bits: int, sample: int = 32768):
    """lf_autotune_mixed: the fastest workgroup size of the bucket kernel at `bits` (cached per device)."""
    L = lib()
    _mixed_sigs(L)
    b, ms = C.c_uint(), C.c_double()
    check(L.lf_autotune_mixed(bits, sample, C.byref(b), C.byref(ms)))
    return b.value, ms.value
