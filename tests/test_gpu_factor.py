"""Blocked Cholesky and triangular solves on the GPU (lf_cholesky_host / lf_solve_host) against the oracle's
restatement of the reference's host order (src/linalg.mm:411-486), bit for bit: pivots, the factor, and
forward / transposed / combined solves, over the input kinds, failure pivots, block sizes and block
boundaries of the reference's tests/test_cholesky.cpp:57-138."""
import numpy as np
import pytest

from factor_cases import failure_cases, rhs, spd
from lfcases import zero, random_records

pytestmark = pytest.mark.gpu


def runs(n):
    r = [("default", 32), ("block 16", 16)]
    if n <= 130:
        r.append(("single block", 0))
    if n <= 40:
        r.append(("block 1", 1))
    if n <= 140:
        r.append(("block 7", 7))
    return r


def check_factor(capi, oracle, rng, bits, A, n, label, blocks, solve=True, expect_spd=True):
    for q, (name, block) in enumerate(blocks):
        print(f"[gpu] cholesky bits={bits} n={n} {label} [{name}]", flush=True)
        want, piv, psign, pstat = oracle.cholesky(bits, A, n, block)
        if expect_spd:
            assert piv == n, f"fixture is not SPD at {piv}: {label}"
        got, info = capi.cholesky(bits, A, n, block=block)
        assert (info.pivot, info.pivot_sign, info.pivot_status) == (piv, psign, pstat), (label, name)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"{label} [{name}]: {bad.size} entries differ, first ({bad[0] // n},{bad[0] % n})"
        if not solve or piv < n:
            continue
        nrhs = (1, 3, 5)[q % 3]
        B = rhs(rng, bits, n, nrhs)
        if q == 1 and n > 3:
            B[nrhs + 1, bits // 32 + 2] = 4
        for passes in (3, 1, 2):
            X = oracle.solve(bits, want, n, B, nrhs, block, passes)
            Y, _, _ = capi.solve(bits, passes, got, n, B, nrhs, block=block)
            bad = np.nonzero((X != Y).any(axis=1))[0]
            assert bad.size == 0, f"solve passes={passes} nrhs={nrhs} {label} [{name}]: {bad.size} entries differ"


@pytest.mark.parametrize("bits", [256, 64, 384])
def test_failure_pivots(capi, oracle, bits):
    rng = np.random.default_rng(4242 + bits)
    for label, A, n in failure_cases(oracle, rng, bits):
        check_factor(capi, oracle, rng, bits, A, n, label, runs(n), expect_spd=False)


@pytest.mark.parametrize("bits", [256, 224, 64, 384])
@pytest.mark.parametrize("n,kind", [(1, "moderate"), (2, "qsc"), (7, "graded"), (33, "qsc"), (48, "banded"),
                                    (70, "spread"), (128, "wide")])
def test_spd_kinds(capi, oracle, bits, n, kind):
    if bits in (64, 384) and n == 128:
        pytest.skip("the largest case runs at 224 and 256 bits")
    rng = np.random.default_rng(bits * 1000 + n)
    A = spd(oracle, rng, bits, n, kind)
    check_factor(capi, oracle, rng, bits, A, n, kind, runs(n))


@pytest.mark.parametrize("bits", [640, 704, 768, 1024])
@pytest.mark.parametrize("n,kind", [(33, "qsc"), (40, "spread"), (48, "banded")])
def test_wide_values(capi, oracle, bits, n, kind):
    """640-1024 bits: the exact dots' window holds products and the addend together at these widths (the
    addend sits about 32N bits above products of its magnitude), so the register window, not only the
    cluster path, runs here; spread inputs take both."""
    rng = np.random.default_rng(bits * 7 + n)
    A = spd(oracle, rng, bits, n, kind)
    check_factor(capi, oracle, rng, bits, A, n, f"{kind} {bits}-bit", [("default", 32), ("block 16", 16)])


@pytest.mark.parametrize("bits", [256, 224])
def test_rank_deficient(capi, oracle, bits):
    rng = np.random.default_rng(bits + 33)
    A = spd(oracle, rng, bits, 33, "rank_deficient")
    check_factor(capi, oracle, rng, bits, A, 33, "rank-deficient", runs(33), expect_spd=False)


@pytest.mark.parametrize("n", [15, 16, 17, 32, 47, 63, 64, 65, 129])
def test_block_boundaries(capi, oracle, n):
    bits = 256
    rng = np.random.default_rng(n)
    A = spd(oracle, rng, bits, n, "qsc" if n % 2 else "moderate")
    check_factor(capi, oracle, rng, bits, A, n, "block boundary", [("block 16", 16), ("block 64", 64)])


@pytest.mark.parametrize("bits", [256, 64])
def test_trsm_zero_diagonal(capi, oracle, bits):
    """A zero diagonal entry of a caller's L: division_by_zero propagates (test_cholesky.cpp:117-120)."""
    rng = np.random.default_rng(bits + 9)
    n, nrhs = 20, 2
    L = np.zeros((n * n, bits // 32 + 3), np.uint32)
    for i in range(n):
        L[i * n:i * n + i + 1] = random_records(rng, bits, i + 1, 3)
    L[9 * n + 9] = zero(bits)
    B = rhs(rng, bits, n, nrhs)
    for passes in (1, 2):
        X = oracle.solve(bits, L, n, B, nrhs, 4, passes)
        Y, _, _ = capi.solve(bits, passes, L, n, B, nrhs, block=4)
        assert np.array_equal(X, Y), passes


def test_empty_problems(capi, oracle):
    bits = 256
    L, info = capi.cholesky(bits, np.zeros((0, 11), np.uint32), 0)
    assert info.pivot == 0
    X, _, _ = capi.solve(bits, 3, np.zeros((0, 11), np.uint32), 0, np.zeros((0, 11), np.uint32), 3)
    assert X.size == 0
    rng = np.random.default_rng(5)
    A = spd(oracle, rng, bits, 5, "moderate")
    L, info = capi.cholesky(bits, A, 5)
    assert info.pivot == 5
    X, _, _ = capi.solve(bits, 3, L, 5, np.zeros((0, 11), np.uint32), 0)
    assert X.size == 0


def test_invalid_arguments(capi):
    with pytest.raises(ValueError):
        capi.cholesky(80, np.zeros((4, 5), np.uint32), 2)
    with pytest.raises(ValueError):
        capi.solve(256, 4, np.zeros((4, 11), np.uint32), 2, np.zeros((2, 11), np.uint32), 1)


@pytest.mark.parametrize("bits", [256])
def test_qsc_like_400(capi, oracle, bits):
    """QSC-like normal equations, n = 400 (test_cholesky.cpp:146-149)."""
    rng = np.random.default_rng(400 + bits)
    A = spd(oracle, rng, bits, 400, "qsc")
    check_factor(capi, oracle, rng, bits, A, 400, "QSC n=400", [("default", 32), ("block 16", 16)])
