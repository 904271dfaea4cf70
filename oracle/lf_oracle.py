"""ctypes loader of the CPU oracle (liblf_oracle.so) and reader of the golden fixtures.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the limbforge_amd package.  Records are numpy uint32 arrays of shape (count, n+3) (real) or
(count, 2*(n+3)) (complex), the public element layout.
"""
from __future__ import annotations

import ctypes as C
import struct
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblf_oracle.so"
REF_PATH = HERE / "_ref" / "libref.so"
GOLDEN = HERE.parent / "tests" / "golden"

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        vp, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
        L.lfo_elementwise.argtypes = [i32, i32, vp, vp, vp, vp, sz, sz, sz, sz, sz]
        L.lfo_elementwise.restype = i32
        L.lfo_mul_chain.argtypes = [i32, vp, vp, vp, sz, i32]
        L.lfo_mul_chain.restype = i32
        L.lfo_cast.argtypes = [i32, i32, i32, vp, vp, sz]
        L.lfo_cast.restype = i32
        L.lfo_exact_dot_add.argtypes = [i32, vp, i32, vp, C.c_ssize_t, vp, C.c_ssize_t, sz, vp]
        L.lfo_exact_dot_add.restype = None
        L.lfo_gemm.argtypes = [i32, i32, vp, vp, sz, sz, sz, vp, i32]
        L.lfo_gemm.restype = i32
        L.lfo_syrk.argtypes = [i32, vp, sz, sz, vp, i32, i32]
        L.lfo_syrk.restype = i32
        L.lfo_cholesky.argtypes = [i32, vp, sz, vp, sz, C.POINTER(i32), C.POINTER(C.c_uint32)]
        L.lfo_cholesky.restype = sz
        L.lfo_solve.argtypes = [i32, vp, sz, vp, sz, vp, sz, i32]
        L.lfo_mixed_chain.argtypes = [i32, vp, vp, vp, vp, vp, vp, sz]
        L.lfo_solve.restype = i32
        L.lfo_divide_word.argtypes = [C.c_uint32] * 4
        L.lfo_set_threads.argtypes = [i32]
        L.lfo_segmented_dot.argtypes = [i32, i32, sz, sz, i32, vp, vp, vp]
        L.lfo_segmented_dot.restype = i32
        L.lfo_tree_sum.argtypes = [i32, i32, vp, sz, vp]
        L.lfo_tree_sum.restype = i32
        L.lfo_set_threads.restype = None
        L.lfo_divide_word.restype = C.c_uint64
        _lib = L
    return _lib


def _p(x):
    return None if x is None else np.ascontiguousarray(x).ctypes.data


def elementwise(bits, op, a, b=None, c=None, b_index=(1, 0), c_index=(1, 0)):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros_like(a)
    b = np.ascontiguousarray(b if b is not None else a, dtype=np.uint32)
    c = np.ascontiguousarray(c if c is not None else a, dtype=np.uint32)
    rc = lib().lfo_elementwise(bits, int(op), a.ctypes.data, b.ctypes.data, c.ctypes.data, out.ctypes.data,
                               a.shape[0], b_index[0], b_index[1], c_index[0], c_index[1])
    assert rc == 0
    return out


def mul_chain(bits, x, f, steps):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    f = np.ascontiguousarray(f, dtype=np.uint32)
    out = np.zeros_like(x)
    assert lib().lfo_mul_chain(bits, x.ctypes.data, f.ctypes.data, out.ctypes.data, x.shape[0], steps) == 0
    return out


def mixed_chain(bits, x, f, g, length, mask):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    f = np.ascontiguousarray(f, dtype=np.uint32)
    g = np.ascontiguousarray(g, dtype=np.uint32)
    length = np.ascontiguousarray(length, dtype=np.uint32)
    mask = np.ascontiguousarray(mask, dtype=np.uint64)
    out = np.zeros_like(x)
    assert lib().lfo_mixed_chain(bits, x.ctypes.data, f.ctypes.data, g.ctypes.data, length.ctypes.data, mask.ctypes.data,
                                 out.ctypes.data, x.shape[0]) == 0
    return out


def exact_dot_add(bits, c, subtract, a, sa, b, sb, K):
    n = bits // 32
    out = np.zeros(n + 3, dtype=np.uint32)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    cc = None if c is None else np.ascontiguousarray(c, dtype=np.uint32)
    lib().lfo_exact_dot_add(bits, _p(cc), int(subtract), a.ctypes.data, sa, b.ctypes.data, sb, K, out.ctypes.data)
    return out


def gemm(bits, transpose_a, A, B, m, n, k, C_in=None, subtract=False):
    rw = bits // 32 + 3
    Cm = np.zeros((m * n, rw), dtype=np.uint32) if C_in is None else np.array(C_in, dtype=np.uint32).reshape(m * n, rw)
    A = np.ascontiguousarray(A, dtype=np.uint32)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    assert lib().lfo_gemm(bits, int(transpose_a), A.ctypes.data, B.ctypes.data, m, n, k, Cm.ctypes.data, int(subtract)) == 0
    return Cm


def gemm_rows(bits, A, B, m, n, k, rows=None, threads=None):
    """Rows `rows` (default: all m) of the untransposed product RN(A B), the row blocks spread over host
    threads.  The oracle keeps its working storage in thread-local variables and ctypes releases the
    interpreter lock during each call, so the blocks run in parallel."""
    import concurrent.futures as cf
    import os

    rw = bits // 32 + 3
    A = np.ascontiguousarray(A, dtype=np.uint32).reshape(m, k, rw)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    rows = np.arange(m) if rows is None else np.asarray(rows)
    out = np.zeros((rows.size, n, rw), dtype=np.uint32)
    if threads is None:
        try:
            threads = len(os.sched_getaffinity(0))
        except AttributeError:
            threads = os.cpu_count() or 1
        threads = max(1, min(16, threads))

    def one(t):
        i = int(rows[t])
        a = np.ascontiguousarray(A[i])
        o = np.zeros((n, rw), dtype=np.uint32)
        assert lib().lfo_gemm(bits, 0, a.ctypes.data, B.ctypes.data, 1, n, k, o.ctypes.data, 0) == 0
        out[t] = o

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, range(rows.size)))
    return out.reshape(rows.size * n, rw)


def syrk(bits, A, rows, cols, C_in=None, lower_only=True, subtract=False):
    rw = bits // 32 + 3
    Cm = np.zeros((cols * cols, rw), dtype=np.uint32) if C_in is None else np.array(C_in, dtype=np.uint32).reshape(cols * cols, rw)
    A = np.ascontiguousarray(A, dtype=np.uint32)
    assert lib().lfo_syrk(bits, A.ctypes.data, rows, cols, Cm.ctypes.data, int(lower_only), int(subtract)) == 0
    return Cm


def segmented_dot(bits, complex_values, segments, length, shared_right, a, b):
    rw = (bits // 32 + 3) * (2 if complex_values else 1)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros((segments, rw), dtype=np.uint32)
    assert lib().lfo_segmented_dot(bits, int(complex_values), segments, length, int(shared_right), a.ctypes.data, b.ctypes.data,
                                   out.ctypes.data) == 0
    return out


def tree_sum(bits, complex_values, x):
    rw = (bits // 32 + 3) * (2 if complex_values else 1)
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.zeros((1, rw), dtype=np.uint32)
    assert lib().lfo_tree_sum(bits, int(complex_values), x.ctypes.data, x.shape[0], out.ctypes.data) == 0
    return out


def set_threads(threads):
    """Host threads of cholesky() and solve() (1 = serial); results do not depend on it."""
    lib().lfo_set_threads(int(threads))


def cholesky(bits, A, n, block):
    A = np.ascontiguousarray(A, dtype=np.uint32)
    L = np.zeros_like(A)
    ps, pst = C.c_int(0), C.c_uint32(0)
    pivot = lib().lfo_cholesky(bits, A.ctypes.data, n, L.ctypes.data, block, C.byref(ps), C.byref(pst))
    return L, pivot, ps.value, pst.value


def solve(bits, L, n, B, nrhs, block, passes=3):
    L = np.ascontiguousarray(L, dtype=np.uint32)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    X = np.zeros_like(B)
    assert lib().lfo_solve(bits, L.ctypes.data, n, B.ctypes.data, nrhs, X.ctypes.data, block, passes) == 0
    return X


# ------------------------------------------------------------------ fixtures
def _records(buf, off, count, rw):
    arr = np.frombuffer(buf, dtype=np.uint32, count=count * rw, offset=off).reshape(count, rw).copy()
    return arr, off + count * rw * 4


def read_golden(name):
    """Parse one fixture file written by tests/golden/gen_golden.c into a list of section dicts."""
    buf = (GOLDEN / name).read_bytes()
    assert buf[:4] == b"LFG1", name
    kind, bits, nsec = struct.unpack_from("<3I", buf, 4)
    off = 16
    out = []
    for _ in range(nsec):
        if kind == 1:
            op, count, has_c, _r = struct.unpack_from("<4I", buf, off)
            off += 16
            rw = (bits // 32 + 3) * (2 if op in (4, 5, 6) else 1)
            a, off = _records(buf, off, count, rw)
            b, off = _records(buf, off, count, rw)
            c = None
            if has_c:
                c, off = _records(buf, off, count, rw)
            o, off = _records(buf, off, count, rw)
            out.append(dict(bits=bits, op=op, a=a, b=b, c=c, out=o))
        elif kind == 2:
            sbits, count, steps, _r = struct.unpack_from("<4I", buf, off)
            off += 16
            rw = sbits // 32 + 3
            x, off = _records(buf, off, count, rw)
            f, off = _records(buf, off, count, rw)
            o, off = _records(buf, off, count, rw)
            out.append(dict(bits=sbits, steps=steps, x=x, f=f, out=o))
        elif kind == 3:
            K, mode, _a, _b = struct.unpack_from("<4I", buf, off)
            off += 16
            rw = bits // 32 + 3
            a, off = _records(buf, off, K, rw)
            b, off = _records(buf, off, K, rw)
            c, off = _records(buf, off, 1, rw)
            o, off = _records(buf, off, 1, rw)
            out.append(dict(bits=bits, K=K, mode=mode, a=a, b=b, c=c[0], out=o[0]))
        elif kind == 4:
            sbits, m, n, k, flags = struct.unpack_from("<5I", buf, off)
            off += 20
            rw = sbits // 32 + 3
            syrk_ = bool(flags & 4)
            A, off = _records(buf, off, k * m if syrk_ else m * k, rw)
            B = None
            if not syrk_:
                B, off = _records(buf, off, k * n, rw)
            Cin, off = _records(buf, off, m * n, rw)
            Cout, off = _records(buf, off, m * n, rw)
            out.append(dict(bits=sbits, m=m, n=n, k=k, transpose_a=bool(flags & 1), subtract=bool(flags & 2),
                            syrk=syrk_, lower_only=bool(flags & 8), A=A, B=B, C_in=Cin, C_out=Cout))
        else:
            raise ValueError(kind)
    assert off == len(buf), (name, off, len(buf))
    return out
