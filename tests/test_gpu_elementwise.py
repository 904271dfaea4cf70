"""GPU parity of the element-wise kernels through the C ABI: bit-exact against the oracle at every
width (64..1024 bits) on the reference tests' case classes, against the MPFR golden fixtures, with
broadcast operands, aliasing, the host-array entry and empty batches (tests/test_arithmetic.cpp,
test_fused.cpp, test_broadcast.cpp of the reference)."""
import sys

import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu

WIDTHS = list(range(64, 1025, 32))
REAL_OPS = ["add", "sub", "mul", "div", "square", "sqrt", "fma", "fms"]
COMPLEX_OPS = ["complex_add", "complex_mul", "complex_div"]


def run_gpu(capi, bits, op, a, b=None, c=None, b_index=None, c_index=None):
    cplx = capi.OPS[op] in capi.COMPLEX_OPS
    count = a.shape[0]
    A = capi.Buffer(bits, count, cplx)
    A.upload(a)
    B = C = None
    if b is not None:
        B = capi.Buffer(bits, b.shape[0], cplx)
        B.upload(b)
    if c is not None:
        C = capi.Buffer(bits, c.shape[0], cplx)
        C.upload(c)
    O = capi.Buffer(bits, count, cplx)
    capi.elementwise(bits, op, A, B, C, O, b_index, c_index)
    return O.download()


def assert_same(got, want, label):
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} mismatches, first rows {bad[:8].tolist()}"


def oracle_mul(oracle, bits):
    return lambda a, b: oracle.elementwise(bits, 2, a, b)


@pytest.mark.parametrize("bits", WIDTHS)
def test_real_ops_vs_oracle(capi, oracle, bits):
    print(f"[gpu] real ops bits={bits}", flush=True)
    rng = np.random.default_rng(20261007 + bits)
    count = 600
    a, b = lfcases.binary_cases(rng, bits, count)
    roots = lfcases.root_cases(rng, bits, count)
    fa, fb, fc = lfcases.fma_cases(rng, bits, count, oracle_mul(oracle, bits))
    for op in REAL_OPS:
        code = capi.OPS[op]
        if op == "sqrt":
            x, y, z = roots, None, None
        elif op in ("fma", "fms"):
            x, y, z = fa, fb, fc
        elif op == "square":
            x, y, z = a, None, None
        else:
            x, y, z = a, b, None
        got = run_gpu(capi, bits, op, x, y, z)
        want = oracle.elementwise(bits, code, x, y, z)
        assert_same(got, want, f"{op} bits={bits}")


@pytest.mark.parametrize("bits", WIDTHS)
def test_complex_ops_vs_oracle(capi, oracle, bits):
    print(f"[gpu] complex ops bits={bits}", flush=True)
    rng = np.random.default_rng(77 + bits)
    a, b = lfcases.complex_cases(rng, bits, 160)
    for op in COMPLEX_OPS:
        got = run_gpu(capi, bits, op, a, b)
        want = oracle.elementwise(bits, capi.OPS[op], a, b)
        assert_same(got, want, f"{op} bits={bits}")


@pytest.mark.skip(reason="complex_fma / complex_fms are switched off on the GPU: their element-wise kernel took a "
                         "memory-access fault at 800 bits (complex_fms, full addend) whose cause is not established; "
                         "the entries reject the two operations (test_complex_fused_ops_are_rejected)")
@pytest.mark.parametrize("bits", WIDTHS)
def test_complex_fused_vs_oracle(capi, oracle, bits):
    """complex_fma / complex_fms (core.hpp:330-358): each part one rounding of a.re b.re - a.im b.im + c.re and
    a.re b.im + a.im b.re + c.im, with the addend broadcast, cancelling products, and addends far above and
    below the products."""
    print(f"[gpu] complex fma/fms bits={bits}", flush=True)
    rng = np.random.default_rng(1177 + bits)
    a, b = lfcases.complex_cases(rng, bits, 96)
    rw = bits // 32 + 3
    c = np.concatenate([lfcases.random_records(rng, bits, 96, 3000), lfcases.random_records(rng, bits, 96, 3000)], axis=1)
    e = lfcases.get_exp(a[:, :rw], bits) + lfcases.get_exp(b[:, :rw], bits)
    far = np.arange(96) % 4
    c[far == 1, bits // 32] = lfcases.i32(e[far == 1] + 3 * bits)
    c[far == 2, bits // 32] = lfcases.i32(e[far == 2] - 4 * bits)
    c[np.arange(96) % 11 == 3] = 0
    for op in ("complex_fma", "complex_fms"):
        got = run_gpu(capi, bits, op, a, b, c)
        assert_same(got, oracle.elementwise(bits, capi.OPS[op], a, b, c), f"{op} bits={bits}")
        got = run_gpu(capi, bits, op, a, b, c[:7], c_index=(1, 7))
        assert_same(got, oracle.elementwise(bits, capi.OPS[op], a, b, c[:7], c_index=(1, 7)), f"{op} bits={bits} broadcast c")


def test_complex_fused_ops_are_rejected(capi):
    """The device entries refuse complex_fma / complex_fms before any launch, with the error class of an
    invalid argument; the host functions cfma / cfms (core.hpp) remain their implementation."""
    rng = np.random.default_rng(5)
    a, b = lfcases.complex_cases(rng, 256, 8)
    for op in ("complex_fma", "complex_fms"):
        with pytest.raises(ValueError, match="not run on the GPU"):
            run_gpu(capi, 256, op, a, b, a)
        with pytest.raises(ValueError, match="not run on the GPU"):
            capi.run_host(256, op, a, b, a)


@pytest.mark.parametrize("bits", [64, 96, 256, 384, 512, 1024])
def test_golden_fixtures(capi, oracle, bits):
    print(f"[gpu] golden bits={bits}", flush=True)
    names = {v: k for k, v in capi.OPS.items()}
    for s in oracle.read_golden(f"elementwise_{bits}.bin"):
        op = names[s["op"]]
        unary = s["op"] in capi.UNARY_OPS
        got = run_gpu(capi, bits, op, s["a"], None if unary else s["b"], s["c"])
        assert_same(got, s["out"], f"golden {op} bits={bits}")


@pytest.mark.parametrize("bits", [64, 256, 544, 1024])
def test_broadcast_operands(capi, oracle, bits):
    """b and c read through (i / stride) % period (engine.hpp:19-25; kernels.metal:7)."""
    print(f"[gpu] broadcast bits={bits}", flush=True)
    rng = np.random.default_rng(3 + bits)
    count = 1000
    a = lfcases.random_records(rng, bits, count, 512)
    for op, stride, period in [("mul", 1, 64), ("div", 3, 0), ("add", 7, 5), ("sub", 1, 1)]:
        nb = min(period, -(-count // stride)) if period else -(-count // stride)
        b = lfcases.random_records(rng, bits, nb, 512)
        got = run_gpu(capi, bits, op, a, b, b_index=(stride, period))
        want = oracle.elementwise(bits, capi.OPS[op], a, b, b_index=(stride, period))
        assert_same(got, want, f"broadcast {op} stride={stride} period={period}")
    # config-1 fma: b broadcast with period 64, c element-wise
    b = lfcases.random_records(rng, bits, 64, 512)
    c = lfcases.random_records(rng, bits, count, 512)
    got = run_gpu(capi, bits, "fma", a, b, c, b_index=(1, 64))
    want = oracle.elementwise(bits, 9, a, b, c, b_index=(1, 64))
    assert_same(got, want, "fma period 64")
    # both broadcast
    c2 = lfcases.random_records(rng, bits, 10, 512)
    got = run_gpu(capi, bits, "fms", a, b, c2, b_index=(1, 64), c_index=(100, 0))
    want = oracle.elementwise(bits, 10, a, b, c2, b_index=(1, 64), c_index=(100, 0))
    assert_same(got, want, "fms two broadcasts")


def test_aliasing_and_in_place(capi, oracle):
    bits = 384
    rng = np.random.default_rng(11)
    a, b = lfcases.binary_cases(rng, bits, 300)
    A = capi.Buffer(bits, 300)
    B = capi.Buffer(bits, 300)
    A.upload(a)
    B.upload(b)
    capi.elementwise(bits, "mul", A, B, None, A)      # out aliases a
    capi.elementwise(bits, "add", A, A, None, A)      # all three alias
    capi.elementwise(bits, "fma", A, B, A, A)         # c aliases out
    want = oracle.elementwise(bits, 2, a, b)
    want = oracle.elementwise(bits, 0, want, want)
    want = oracle.elementwise(bits, 9, want, b, want)
    assert_same(A.download(), want, "in-place chain")


@pytest.mark.parametrize("bits", [128, 1024])
def test_host_array_entry(capi, oracle, bits):
    """Engine::run / run_unary / run_ternary on host arrays (engine.hpp:202-204)."""
    rng = np.random.default_rng(99 + bits)
    a, b = lfcases.binary_cases(rng, bits, 777)
    for op in ["add", "div", "sqrt"]:
        out, gpu_s, wall_s = capi.run_host(bits, op, np.abs(a) if False else a, None if op == "sqrt" else b)
        want = oracle.elementwise(bits, capi.OPS[op], a, b)
        assert_same(out, want, f"host {op}")
        assert wall_s >= gpu_s >= 0
    ca, cb = lfcases.complex_cases(rng, bits, 50)
    out, _, _ = capi.run_host(bits, "complex_div", ca, cb)
    assert_same(out, oracle.elementwise(bits, 6, ca, cb), "host complex_div")
    c = lfcases.random_records(rng, bits, 64, 100)
    out, _, _ = capi.run_host(bits, "fma", a, c, b, b_index=(1, 64))
    assert_same(out, oracle.elementwise(bits, 9, a, c, b, b_index=(1, 64)), "host fma broadcast")


def test_empty_and_errors(capi):
    out, g, w = capi.run_host(256, "mul", np.zeros((0, 11), np.uint32), np.zeros((0, 11), np.uint32))
    assert out.shape == (0, 11)
    with pytest.raises(ValueError):
        capi.run_host(250, "mul", np.zeros((1, 11), np.uint32), np.zeros((1, 11), np.uint32))


@pytest.mark.parametrize("from_bits,to_bits", [(64, 1024), (1024, 64), (96, 544), (544, 96), (512, 512), (1024, 992), (160, 128)])
def test_casts(capi, oracle, from_bits, to_bits):
    """Exact widening, RN narrowing with carry and overflow (kernels.metal:251-266), real and complex."""
    import ctypes as C

    rng = np.random.default_rng(from_bits * 7 + to_bits)
    n = 500
    x = lfcases.random_records(rng, from_bits, n, 1000)
    w = from_bits // 32
    x[::7, :w] = 0xFFFFFFFF                 # all-ones mantissas carry into the next binade
    x[::11, w] = 1000000000                 # ... and overflow at the top exponent
    x[::13] = lfcases.zero(from_bits, 2)
    x[::17] = lfcases.zero(from_bits)
    for cplx in (0, 1):
        src = np.concatenate([x, x[::-1]], axis=1) if cplx else x
        out = np.zeros((n, (to_bits // 32 + 3) * (2 if cplx else 1)), dtype=np.uint32)
        g, wl = C.c_double(), C.c_double()
        capi.check(capi.lib().lf_cast_host(from_bits, to_bits, cplx, src.ctypes.data, out.ctypes.data, n, C.byref(g), C.byref(wl)))
        want = np.zeros_like(out)
        assert oracle.lib().lfo_cast(from_bits, to_bits, cplx, src.ctypes.data, want.ctypes.data, n) == 0
        assert_same(out, want, f"cast {from_bits}->{to_bits} complex={cplx}")
