"""Seeded symmetric positive definite inputs for the Cholesky and triangular-solve parity tests.

The kinds follow the reference's tests/test_cholesky.cpp:57-79: A = J^T J plus a Marquardt damping
a_ii += 2^(e_ii - 10) for J of 2n+1 rows with per-column exponent scales (moderate, qsc, wide, graded),
a rank-deficient J (two dependent column pairs, no damping), a banded diagonally dominant matrix, and
"spread": A = L0 L0^T (one rounding per entry) for a unit lower L0 with off-diagonal exponents down to
-700, whose trailing updates span more than any band.  The products are formed by the oracle (CPU).
"""
from __future__ import annotations

import numpy as np

from lfcases import MASK, i32, negate, one, random_records, zero

KINDS = ("moderate", "qsc", "wide", "graded", "banded", "spread", "rank_deficient")


def _rand(rng, bits, span):
    return random_records(rng, bits, 1, span)[0]


def spd(oracle, rng, bits, n, kind):
    nw = bits // 32
    rw = nw + 3
    A = np.zeros((n * n, rw), np.uint32)
    if kind == "banded":
        for i in range(n):
            d = _rand(rng, bits, 0)
            d[nw + 1] = 1
            d[nw] = 3
            A[i * n + i] = d
            for off in (1, 3):
                if off <= i:
                    x = _rand(rng, bits, 0)
                    x[nw] = i32(-1)
                    A[i * n + i - off] = x
                    A[(i - off) * n + i] = x
        return A
    if kind == "spread":
        L0 = np.zeros((n * n, rw), np.uint32)
        T = np.zeros((n * n, rw), np.uint32)
        for i in range(n):
            for j in range(i + 1):
                x = _rand(rng, bits, 0)
                if i == j:
                    x[nw + 1] = 1
                    x[nw] = 0
                elif rng.integers(0, 10) < 3:
                    x = zero(bits)
                else:
                    x[nw] = i32(-int(rng.integers(0, 701)))
                L0[i * n + j] = x
                T[j * n + i] = x
        return oracle.gemm(bits, False, L0, T, n, n, n)
    m = 2 * n + 1
    J = np.zeros((m * n, rw), np.uint32)
    for c in range(n):
        scale = {"qsc": int(rng.integers(0, 121)) - 60, "wide": int(rng.integers(0, 601)) - 300,
                 "graded": 20 if c % 2 else -20}.get(kind, 0)
        col = random_records(rng, bits, m, 15 if kind == "qsc" else 4)
        col[:, nw] = i32(col[:, nw].astype(np.int64) - ((col[:, nw].astype(np.int64) >> 31) << 32) + scale)
        if kind == "qsc":
            col[rng.integers(0, 4, size=m) == 0] = zero(bits)
        J[c::n] = col
        if kind == "qsc" and c % 17 == 5:
            r = int(rng.integers(0, m))
            J[r * n + c, nw] = i32(scale - 150)
    if kind == "rank_deficient" and n >= 3:
        for r in range(m):
            J[r * n + n - 1] = J[r * n + 1]
            J[r * n + n // 2] = negate(J[r * n + 0], bits)
    A = oracle.syrk(bits, J, m, n, lower_only=False)
    if kind != "rank_deficient":
        for i in range(n):
            a = A[i * n + i]
            if a[nw + 1] != 0:
                e = int(a[nw]) - ((int(a[nw]) >> 31) << 32)
                A[i * n + i] = oracle.elementwise(bits, 0, a[None, :], one(bits, e - 10)[None, :])[0]
    return A


def rhs(rng, bits, n, nrhs):
    B = random_records(rng, bits, n * nrhs, 8)
    B[rng.integers(0, 11, size=n * nrhs) == 0] = zero(bits)
    return B


def failure_cases(oracle, rng, bits):
    """(label, A, n) inputs whose factorization stops at a pivot (test_cholesky.cpp:106-115)."""
    nw = bits // 32
    o, t = one(bits), one(bits, 1)
    out = [("exact zero pivot", np.stack([o, o, o, o]), 2), ("indefinite", np.stack([o, t, t, o]), 2),
           ("negative 1x1", np.stack([negate(o, bits)]), 1), ("zero matrix", np.stack([zero(bits)] * 9), 3)]
    n = 40
    A = spd(oracle, rng, bits, n, "moderate")
    A[37 * n + 37] = one(bits, 40, -1)
    out.append(("negative pivot 37", A, n))
    A = spd(oracle, rng, bits, n, "moderate")
    A[20 * n + 5, nw + 2] = 4
    out.append(("status at (20,5)", A, n))
    A = spd(oracle, rng, bits, n, "qsc")
    A[0, nw + 2] = 2
    out.append(("status at (0,0)", A, n))
    A = spd(oracle, rng, bits, n, "moderate")
    for j in range(n):
        A[max(j, 23) * n + min(j, 23)] = zero(bits)
    out.append(("zero row 23", A, n))
    A = spd(oracle, rng, bits, n, "moderate")
    for j in range(17):
        A[17 * n + j] = A[16 * n + j]
    A[17 * n + 17] = A[16 * n + 16]
    A[17 * n + 16] = A[16 * n + 16]
    for i in range(18, n):
        A[i * n + 17] = A[i * n + 16]
    out.append(("duplicated row 17", A, n))
    return out
