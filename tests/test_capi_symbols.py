"""The C ABI library loads without a GPU and exports every entry point include/*.h declares; argument
checks that need no device answer with the reference's error classes."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(capi):
    lib = capi.lib()
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_argument_errors_without_device(capi):
    a = capi.Operand(None, 0)
    with pytest.raises(ValueError, match="bits"):
        capi.check(capi.lib().lf_elementwise(48, 0, ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), 1, None, None, None))
    with pytest.raises(ValueError, match="32-bit"):
        capi.check(capi.lib().lf_elementwise(64, 0, ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), 1 << 33, None, None, None))
    with pytest.raises(ValueError, match="operation"):
        capi.check(capi.lib().lf_elementwise(64, 13, ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), 1, None, None, None))
    bad = capi.Broadcast(0, 0)
    with pytest.raises(ValueError, match="stride"):
        capi.check(capi.lib().lf_elementwise(64, 0, ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), 1, ctypes.byref(bad), None, None))
    # empty batches are a no-op
    capi.check(capi.lib().lf_elementwise(64, 0, ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), 0, None, None, None))
    assert capi.lib().lf_plane_pitch(1) == 64 and capi.lib().lf_plane_pitch(64) == 64 and capi.lib().lf_plane_pitch(65) == 128
    assert capi.lib().lf_planes_bytes(1024, 1, 64) == 2 * 35 * 64 * 4


def test_factor_operand_checks_without_device(capi):
    """lf_cholesky / lf_solve reject short pitches and partly aliased operands before any launch (the
    pointers below are never dereferenced: every call fails in its argument checks)."""
    L = capi.lib()
    P = ctypes.POINTER
    L.lf_cholesky.argtypes = [ctypes.c_int, P(capi.Operand), ctypes.c_uint64, P(capi.Operand), ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p]
    L.lf_solve.argtypes = [ctypes.c_int, ctypes.c_int, P(capi.Operand), ctypes.c_uint64, P(capi.Operand), ctypes.c_uint64,
                           P(capi.Operand), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    bits, n, nrhs = 64, 8, 3
    base = 1 << 40
    span = (bits // 32 + 3) * 64 * 4  # bytes of one operand at pitch 64

    def op(k, pitch=64):
        return capi.Operand(base + k * span, pitch)

    def chol(A, Lo):
        capi.check(L.lf_cholesky(bits, ctypes.byref(A), n, ctypes.byref(Lo), None, None, None, None))

    def solve(Lo, B, X):
        capi.check(L.lf_solve(bits, 3, ctypes.byref(Lo), n, ctypes.byref(B), nrhs, ctypes.byref(X), None, None, None, None, None))

    with pytest.raises(ValueError, match="pitch"):
        chol(op(0, 63), op(1))
    with pytest.raises(ValueError, match="pitch"):
        chol(op(0), op(1, 32))
    with pytest.raises(ValueError, match="overlap"):
        chol(op(0), op(0, 128))          # same planes, different pitch
    with pytest.raises(ValueError, match="overlap"):
        chol(op(0), capi.Operand(base + 256, 64))  # partly overlapping
    with pytest.raises(ValueError, match="null"):
        chol(capi.Operand(None, 64), op(1))
    with pytest.raises(ValueError, match="pitch"):
        solve(op(0, 16), op(1), op(2))
    with pytest.raises(ValueError, match="pitch"):
        solve(op(0), op(1, 16), op(2))
    with pytest.raises(ValueError, match="pitch"):
        solve(op(0), op(1), op(2, 16))
    with pytest.raises(ValueError, match="overlap"):
        solve(op(0), op(1), op(1, 128))  # X and B: same planes, different pitch
    with pytest.raises(ValueError, match="overlap"):
        solve(op(0), op(1), op(0))       # X over L
