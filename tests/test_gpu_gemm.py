"""GPU parity of the exact dense products (Linalg::gemm / syrk, linalg.hpp:149-155) through the C ABI:
bit-exact against the MPFR golden fixtures and the oracle on the matrix kinds of the reference's
tests/test_linalg.cpp:71-106 (moderate, wide/multi-band with spread fallback, cancellation, sparse with
statuses, ties, saturated), under the option configurations of :156-157 (default, narrow bands,
zero-width bands, forced fallback) plus the plain limb GEMM method; SYRK lower/full, transposed GEMM,
updates, K = 0; and the residue and limb GEMMs against each other on a config-3-shaped product."""
import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu

OPTION_SETS = {
    "default": {},
    "narrow": {"band_bits": 8},
    "zero_width": {"band_bits": 0},
    "forced": {"force_fallback": 1},
    "limb_gemm": {"method": 1},
    "wide_bands": {"band_bits": 512},
}


def same(got, want, label):
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} mismatching outputs, first {bad[:8].tolist()}"


def matrix(rng, bits, rows, cols, kind):
    """Entries of the reference test kinds (test_linalg.cpp:71-106)."""
    n = bits // 32
    count = rows * cols
    x = lfcases.random_records(rng, bits, count, 0)
    if kind == "moderate":
        e = rng.integers(-40, 41, size=count)
    elif kind == "bands3":   # config-3 style: centres -200, 0, +200, sigma 8
        e = rng.choice([-200, 0, 200], size=count) + np.round(rng.normal(0, 8, size=count)).astype(np.int64)
    elif kind == "wide":     # 2-5 bands per line and some lines beyond max_spread
        e = rng.choice([-900, -300, 0, 250, 600], size=count) + rng.integers(-20, 21, size=count)
    elif kind == "huge":     # astronomically spread exponents (cluster path)
        e = rng.integers(-600000000, 600000001, size=count)
    elif kind == "cancel":
        e = rng.integers(-10, 11, size=count)
    elif kind == "saturated":
        e = np.zeros(count, dtype=np.int64)
        x[:, :n] = 0xFFFFFFFF
        x[:, n + 1] = 1
    else:
        e = rng.integers(-100, 101, size=count)
    x[:, n] = lfcases.i32(e)
    if kind == "sparse":
        x[rng.random(count) < 0.6] = lfcases.zero(bits)
        st = rng.random(count) < 0.01
        x[st, n + 2] = rng.choice([1, 2, 4], size=st.sum())
    if kind == "ties":
        x[:, :n] = 0
        x[:, n - 1] = 0x80000000
    return x


def cancel_pair(rng, bits, m, n, k):
    """A and B whose products cancel in pairs except for a tiny survivor (last-bit cancellation)."""
    A = matrix(rng, bits, m, k, "cancel")
    B = matrix(rng, bits, k, n, "cancel")
    w = bits // 32
    half = k // 2
    # column 2t+1 of A mirrors column 2t with the opposite sign; rows of B repeat
    for t in range(half):
        A[np.arange(m) * k + 2 * t + 1] = lfcases.negate(A[np.arange(m) * k + 2 * t], bits)
        B[(2 * t + 1) * n + np.arange(n)] = B[(2 * t) * n + np.arange(n)]
    if k % 2:
        last = (k - 1) * n + np.arange(n)
        B[last, w] = lfcases.i32(lfcases.get_exp(B[last], bits) - bits - 30)
    return A, B


def test_golden_products(capi, oracle):
    for idx, s in enumerate(oracle.read_golden("products.bin")):
        print(f"[gpu] golden product {idx} bits={s['bits']}", flush=True)
        for name, kw in OPTION_SETS.items():
            o = capi.LinalgOptions.default(**kw)
            if s["syrk"]:
                got, rep, _, _ = capi.syrk(s["bits"], s["A"], s["k"], s["m"], s["C_in"], s["lower_only"], s["subtract"], o)
            else:
                got, rep, _, _ = capi.gemm(s["bits"], s["transpose_a"], s["A"], s["B"], s["m"], s["n"], s["k"], s["C_in"],
                                           s["subtract"], o)
            same(got, s["C_out"], f"golden product {idx} options={name}")


@pytest.mark.parametrize("bits", [64, 160, 512, 1024])
@pytest.mark.parametrize("kind", ["moderate", "bands3", "wide", "sparse", "ties", "saturated"])
def test_gemm_kinds_vs_oracle(capi, oracle, bits, kind):
    print(f"[gpu] gemm kind={kind} bits={bits}", flush=True)
    rng = np.random.default_rng(bits * 1000 + sum(map(ord, kind)))
    m, n, k = 9, 11, 23
    A = matrix(rng, bits, m, k, kind)
    B = matrix(rng, bits, k, n, kind)
    want = oracle.gemm(bits, False, A, B, m, n, k)
    for name, kw in OPTION_SETS.items():
        got, rep, _, _ = capi.gemm(bits, False, A, B, m, n, k, options=capi.LinalgOptions.default(**kw))
        same(got, want, f"gemm {kind} bits={bits} options={name}")


@pytest.mark.parametrize("bits", [96, 512])
def test_huge_spread_cluster_path(capi, oracle, bits):
    rng = np.random.default_rng(bits)
    m, n, k = 5, 6, 17
    A = matrix(rng, bits, m, k, "huge")
    B = matrix(rng, bits, k, n, "huge")
    want = oracle.gemm(bits, False, A, B, m, n, k)
    got, rep, _, _ = capi.gemm(bits, False, A, B, m, n, k)
    same(got, want, "huge spread")
    assert rep.fallback_outputs == m * n


@pytest.mark.parametrize("bits", [128, 544])
def test_cancellation(capi, oracle, bits):
    rng = np.random.default_rng(5 + bits)
    m, n, k = 7, 5, 21
    A, B = cancel_pair(rng, bits, m, n, k)
    want = oracle.gemm(bits, False, A, B, m, n, k)
    for name, kw in OPTION_SETS.items():
        got, _, _, _ = capi.gemm(bits, False, A, B, m, n, k, options=capi.LinalgOptions.default(**kw))
        same(got, want, f"cancellation options={name}")


@pytest.mark.parametrize("bits", [64, 256, 1024])
def test_syrk_transpose_and_updates(capi, oracle, bits):
    print(f"[gpu] syrk/transpose/updates bits={bits}", flush=True)
    rng = np.random.default_rng(17 + bits)
    rows, cols = 19, 8
    A = matrix(rng, bits, rows, cols, "wide")
    A2 = matrix(rng, bits, rows, cols, "bands3")
    Cold = lfcases.random_records(rng, bits, cols * cols, 300)
    for name, kw in OPTION_SETS.items():
        o = capi.LinalgOptions.default(**kw)
        for AA in (A, A2):
            for lower in (True, False):
                got, _, _, _ = capi.syrk(bits, AA, rows, cols, Cold, lower, False, o)
                same(got, oracle.syrk(bits, AA, rows, cols, Cold, lower, False), f"syrk lower={lower} options={name}")
            got, _, _, _ = capi.syrk(bits, AA, rows, cols, Cold, True, True, o)
            same(got, oracle.syrk(bits, AA, rows, cols, Cold, True, True), f"syrk update options={name}")
        # transposed GEMM update: C (m x n) - A^T B with A k x m
        m, n, k = 6, 7, 13
        At = matrix(rng, bits, k, m, "bands3")
        B = matrix(rng, bits, k, n, "bands3")
        C0 = lfcases.random_records(rng, bits, m * n, 450)
        C0[3] = lfcases.zero(bits, 2)
        got, _, _, _ = capi.gemm(bits, True, At, B, m, n, k, C0, True, o)
        same(got, oracle.gemm(bits, True, At, B, m, n, k, C0, True), f"transposed update options={name}")


def test_empty_k_and_shapes(capi, oracle):
    bits = 192
    rng = np.random.default_rng(3)
    C0 = lfcases.random_records(rng, bits, 12, 10)
    empty = np.zeros((0, bits // 32 + 3), np.uint32)
    got, rep, _, _ = capi.gemm(bits, False, empty, empty, 3, 4, 0, C0, False)
    assert (got == 0).all()
    got, rep, _, _ = capi.gemm(bits, False, empty, empty, 3, 4, 0, C0, True)
    same(got, C0, "K=0 update leaves C unchanged")
    got, rep, _, _ = capi.gemm(bits, False, empty, empty, 0, 4, 5)
    assert got.shape[0] == 0
    with pytest.raises(ValueError):
        capi.gemm(80, False, C0, C0, 1, 1, 1)
    with pytest.raises(ValueError):
        capi.syrk(bits, C0, 3, 4, None, False, True)


def test_report_moduli_and_bands(capi, oracle):
    bits = 512
    rng = np.random.default_rng(8)
    m = n = k = 32
    A = matrix(rng, bits, m, k, "bands3")
    B = matrix(rng, bits, k, n, "bands3")
    got, rep, _, _ = capi.gemm(bits, False, A, B, m, n, k)
    same(got, oracle.gemm(bits, False, A, B, m, n, k), "32^3 bands3")
    # three bands per line, 9 band pairs, moduli for 2*512 + 2*pmax + log2(32) + 2 bits of 16-bit primes
    assert rep.bands("left")[3] == m and rep.bands("right")[3] == n
    assert rep.npairs == 9 and rep.int8_gemms == 3 * rep.moduli
    assert 70 <= rep.moduli <= 76
    assert rep.gpu_outputs == m * n and rep.multi_band_outputs == m * n


def test_residue_equals_limb_gemm_config3_shape(capi, oracle):
    """A config-3-shaped product (512 bits, three exponent bands) at 128^3: the residue GEMM and the plain
    limb GEMM agree bit for bit, and a sample of outputs matches the oracle."""
    bits, m = 512, 128
    rng = np.random.default_rng(20261007)
    A = matrix(rng, bits, m, m, "bands3")
    B = matrix(rng, bits, m, m, "bands3")
    r1, rep1, g1, _ = capi.gemm(bits, False, A, B, m, m, m)
    r2, rep2, g2, _ = capi.gemm(bits, False, A, B, m, m, m, options=capi.LinalgOptions.default(method=1))
    same(r1, r2, "residue vs limb GEMM")
    rows = rng.integers(0, m, size=24)
    for i in rows:
        want = oracle.gemm(bits, False, A[i * m:(i + 1) * m], B, 1, m, m)
        same(r1[i * m:(i + 1) * m], want, f"row {i} vs oracle")
    print(f"[gpu] 128^3 residue {g1*1e3:.3f} ms, limb {g2*1e3:.3f} ms", flush=True)
