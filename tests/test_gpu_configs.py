"""GPU parity at the benchmark workloads' own shapes, through the C ABI, against the oracle:

- the 512-bit exact GEMM at M = N = K = 512 with three exponent bands per line (entries c + round(8 z),
  c in {-200, 0, +200}): the residue GEMM (default options and band_bits = 512) and the plain limb GEMM
  agree on every output, and every output equals the oracle's RN(sum of exact products);
- products whose depth K runs the int8-MFMA kernel's four-stage LDS ring past its first refill (K of 160,
  300, 600 and 1000, one of them not a multiple of 64), every output against the oracle;
- the 256-bit SYRK of A (2048 x 512, single band): the full symmetric S against the oracle's, then the
  blocked Cholesky (n = 512, block 32) of S = S + I and cholesky_solve with 64 right-hand sides against
  the oracle's restatement of the reference's host order (src/linalg.mm:411-486), bit for bit;
- the mixed-precision graph over all 31 widths (64 to 1024 bits), so every group kernel runs.

The oracle's products run on all host threads (oracle.gemm_rows)."""
import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu


def same(got, want, label):
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} mismatching outputs, first {bad[:8].tolist()}"


def records(rng, bits, count, exps):
    w = bits // 32
    r = np.zeros((count, w + 3), dtype=np.uint32)
    r[:, :w] = rng.integers(0, 1 << 32, size=(count, w), dtype=np.uint64).astype(np.uint32)
    r[:, w - 1] |= np.uint32(0x80000000)
    r[:, w] = lfcases.i32(exps)
    r[:, w + 1] = lfcases.i32(np.where(rng.integers(0, 2, size=count) == 1, 1, -1))
    return r


def three_bands(rng, count):
    return rng.choice([-200, 0, 200], size=count) + np.round(8 * rng.standard_normal(count)).astype(np.int64)


def test_gemm_512bit_512cubed_every_output(capi, oracle):
    bits, n = 512, 512
    rng = np.random.default_rng(20261007 + bits)
    A = records(rng, bits, n * n, three_bands(rng, n * n))
    B = records(rng, bits, n * n, three_bands(rng, n * n))
    print("[gpu] 512-bit 512^3 GEMM: residue (default)", flush=True)
    r1, rep1, g1, _ = capi.gemm(bits, False, A, B, n, n, n)
    # three bands on most lines (a few outliers give a fourth), every output on the residue path
    assert rep1.bands("left")[3] > n // 2 and rep1.npairs >= 9 and rep1.fallback_outputs == 0
    print("[gpu] 512-bit 512^3 GEMM: residue (band_bits 512)", flush=True)
    r2, rep2, g2, _ = capi.gemm(bits, False, A, B, n, n, n, options=capi.LinalgOptions.default(band_bits=512))
    assert rep2.npairs == 1
    print("[gpu] 512-bit 512^3 GEMM: limb", flush=True)
    r3, _, g3, _ = capi.gemm(bits, False, A, B, n, n, n, options=capi.LinalgOptions.default(method=1))
    same(r1, r3, "residue vs limb GEMM")
    same(r2, r3, "residue band_bits=512 vs limb GEMM")
    print(f"[gpu] device ms: residue {g1*1e3:.3f}, band512 {g2*1e3:.3f}, limb {g3*1e3:.3f}; oracle on every output", flush=True)
    want = oracle.gemm_rows(bits, A, B, n, n, n)
    same(r1, want, "512^3 residue GEMM vs oracle")


@pytest.mark.parametrize("bits,m,n,k", [(128, 12, 20, 160), (512, 9, 7, 300), (256, 16, 16, 1000), (1024, 5, 6, 600)])
@pytest.mark.parametrize("opts", ["default", "band512", "narrow"])
def test_deep_k_vs_oracle(capi, oracle, bits, m, n, k, opts):
    kw = {"default": {}, "band512": {"band_bits": 512}, "narrow": {"band_bits": 16}}[opts]
    print(f"[gpu] deep-K product bits={bits} m={m} n={n} k={k} options={opts}", flush=True)
    rng = np.random.default_rng(bits * 7 + k)
    A = records(rng, bits, m * k, three_bands(rng, m * k))
    B = records(rng, bits, k * n, three_bands(rng, k * n))
    got, rep, _, _ = capi.gemm(bits, False, A, B, m, n, k, options=capi.LinalgOptions.default(**kw))
    if opts != "narrow":  # 16-bit bands split every exponent cluster past max_bands: those lines take the exact dots
        assert rep.gpu_outputs == m * n
    want = oracle.gemm_rows(bits, A, B, m, n, k)
    same(got, want, f"K={k} bits={bits} options={opts}")
    # the same product as an update of random old entries
    C0 = lfcases.random_records(rng, bits, m * n, 250)
    got, _, _, _ = capi.gemm(bits, False, A, B, m, n, k, C_in=C0, subtract=True, options=capi.LinalgOptions.default(**kw))
    same(got, oracle.gemm(bits, False, A, B, m, n, k, C0, True), f"K={k} update")


def test_syrk_cholesky_solve_256bit_n512(capi, oracle):
    bits, rows, n, nrhs, block = 256, 2048, 512, 64, 32
    w = bits // 32
    rng = np.random.default_rng(20261007 + 4)
    A = records(rng, bits, rows * n, np.round(8 * rng.standard_normal(rows * n)).astype(np.int64))
    B = records(rng, bits, n * nrhs, rng.integers(-10, 11, size=n * nrhs))
    print("[gpu] 256-bit SYRK 2048 x 512 (full and lower)", flush=True)
    S, rep, _, _ = capi.syrk(bits, A, rows, n, lower_only=False)
    assert rep.gpu_outputs > 0 and rep.fallback_outputs == 0
    Sl, _, _, _ = capi.syrk(bits, A, rows, n, lower_only=True)
    low = (np.arange(n)[:, None] >= np.arange(n)[None, :]).reshape(-1)
    same(Sl[low], S[low], "lower-only SYRK vs full SYRK")
    print("[gpu] SYRK vs oracle on every entry", flush=True)
    At = np.ascontiguousarray(A.reshape(rows, n, w + 3).transpose(1, 0, 2)).reshape(n * rows, w + 3)
    want = oracle.gemm_rows(bits, At, A, n, n, rows)
    same(S, want, "SYRK 2048 x 512 vs oracle")
    one = np.zeros((n, w + 3), np.uint32)
    one[:, w - 1] = 0x80000000
    one[:, w + 1] = 1
    diag = np.arange(n) * (n + 1)
    S[diag], _, _ = capi.run_host(bits, "add", S[diag], one)
    print("[gpu] Cholesky n=512 block 32", flush=True)
    L, info = capi.cholesky(bits, S, n, block=block)
    WL, piv, psign, pstat = oracle.cholesky(bits, S, n, block)
    assert piv == n and (info.pivot, info.pivot_sign, info.pivot_status) == (piv, psign, pstat)
    same(L, WL, "Cholesky n=512 vs oracle")
    print("[gpu] cholesky_solve with 64 right-hand sides", flush=True)
    X, _, _ = capi.solve(bits, 3, L, n, B, nrhs, block=block)
    WX = oracle.solve(bits, WL, n, B, nrhs, block, 3)
    same(X, WX, "solve 64 RHS vs oracle")
    print("[gpu] cholesky_solve with 64 right-hand sides, every off-block update through the update GEMM", flush=True)
    Xg, _, _ = capi.solve(bits, 3, L, n, B, nrhs, options=capi.FactorOptions.default(block=block, solve_host_macs=-1.0))
    same(Xg, WX, "solve 64 RHS (update GEMM) vs oracle")


def test_mixed_graph_all_widths(capi, oracle):
    from test_gpu_mixed import check, make_batch

    rng = np.random.default_rng(31)
    widths = list(range(64, 1025, 32))
    batch = make_batch(rng, widths, 150)
    buckets = [capi.MixedBucket(bits, *batch[bits]) for bits in widths]
    graph = capi.mixed_graph(buckets)
    stream = capi.Stream()
    try:
        for replay in range(2):
            print(f"[gpu] mixed graph over 31 widths, replay {replay}", flush=True)
            capi.graph_launch(graph, stream.handle)
            stream.synchronize()
            check(capi, oracle, batch, buckets)
            for b in buckets:
                x = lfcases.random_records(rng, b.bits, b.count, 64)
                b.X.upload(x)
                batch[b.bits] = (x,) + batch[b.bits][1:]
    finally:
        capi.graph_destroy(graph)
    print("[gpu] mixed chains over 31 widths, per-precision launches", flush=True)
    capi.mixed_chains(buckets)
    capi.lib().lf_stream_synchronize(None)
    check(capi, oracle, batch, buckets)
