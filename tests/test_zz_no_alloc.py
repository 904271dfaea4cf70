"""lf_run_host_noalloc (the host-array element-wise call whose device staging is a static per-GPU array, for
callers that cannot allocate) against lf_run_host on the same inputs, bit for bit: calls in a row on one
GPU, each smaller than the last, so every call runs over the previous call's leftovers; and the size
limit, which is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu


def noalloc(capi, bits, op, a, b=None, c=None, b_index=None, c_index=None):
    L = capi.lib()
    f = L.lf_run_host_noalloc
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(capi.Broadcast),
                  C.POINTER(capi.Broadcast), C.c_void_p]
    f.restype = C.c_int
    code = capi.OPS[op]
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.full_like(a, 0xDEADBEEF)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.uint32)
    cc = None if c is None else np.ascontiguousarray(c, dtype=np.uint32)
    bi = None if b_index is None else C.byref(capi.Broadcast(*b_index))
    ci = None if c_index is None else C.byref(capi.Broadcast(*c_index))
    capi.check(f(bits, code, a.ctypes.data, None if bb is None else bb.ctypes.data, None if cc is None else cc.ctypes.data,
                 out.ctypes.data, a.shape[0], bi, ci, None))
    return out


def test_noalloc_matches_run_host(capi, oracle):
    rng = np.random.default_rng(77)
    cases = []
    a = lfcases.random_records(rng, 1024, 30000, 300)
    b = lfcases.random_records(rng, 1024, 30000, 300)
    c = lfcases.random_records(rng, 1024, 64, 300)
    cases.append((1024, "fma", a, b, c, None, (1, 64)))
    a = lfcases.random_records(rng, 256, 5000, 300)
    b = lfcases.random_records(rng, 256, 5000, 300)
    b[::17] = lfcases.zero(256)
    cases.append((256, "div", a, b, None, None, None))
    ca = np.concatenate([lfcases.random_records(rng, 512, 700, 60), lfcases.random_records(rng, 512, 700, 60)], axis=1)
    cb = np.concatenate([lfcases.random_records(rng, 512, 700, 60), lfcases.random_records(rng, 512, 700, 60)], axis=1)
    cases.append((512, "complex_div", ca, cb, None, None, None))
    a = lfcases.random_records(rng, 96, 100, 300)
    cases.append((96, "sqrt", a, None, None, None, None))  # negative operands give status invalid
    for bits, op, a, b, c, bi, ci in cases:
        print(f"[gpu] run_host_noalloc bits={bits} op={op} count={a.shape[0]}", flush=True)
        got = noalloc(capi, bits, op, a, b, c, bi, ci)
        want, _, _ = capi.run_host(bits, op, a, b, c, bi, ci)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"{bits}-bit {op}: {bad.size} results differ from lf_run_host, first {bad[:8].tolist()}"
        k = min(a.shape[0], 256)
        ref = oracle.elementwise(bits, capi.OPS[op], a[:k], None if b is None else b[:k], c, c_index=ci or (1, 0))
        assert (got[:k] == ref).all(), f"{bits}-bit {op}: differs from the oracle"


def test_noalloc_rejects_batches_beyond_the_scratch_array(capi):
    size = C.c_uint64()
    capi.check(capi.lib().lf_noalloc_scratch_bytes(C.byref(size)))
    assert size.value == 64 << 20
    a = np.zeros((100000, 2 * 35), np.uint32)
    with pytest.raises(ValueError, match="scratch"):
        noalloc(capi, 1024, "complex_div", a, a)
