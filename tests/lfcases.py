"""Seeded input generators for the parity tests (numpy records in the public element layout).

The case classes follow the reference tests' fixtures: random values with an exponent span
(tests/reference.hpp:21-26), alignment and cancellation (tests/test_arithmetic.cpp:20-25), ties at
half an ulp and binade boundaries (:28-36), exponent limits (:32-33), all-ones mantissas (:34),
zeros and statuses (:23-25), and the fma classes of tests/test_fused.cpp:10-28.
"""
from __future__ import annotations

import numpy as np

MASK = 0xFFFFFFFF


def words(bits):
    return bits // 32


def i32(x):
    return np.asarray(x, dtype=np.int64) & MASK


def random_records(rng, bits, count, span=1000, exponents=None):
    n = words(bits)
    r = np.zeros((count, n + 3), dtype=np.uint32)
    r[:, :n] = rng.integers(0, 1 << 32, size=(count, n), dtype=np.uint64).astype(np.uint32)
    r[:, n - 1] |= np.uint32(0x80000000)
    e = rng.integers(-span, span + 1, size=count) if exponents is None else np.asarray(exponents)
    r[:, n] = i32(e)
    r[:, n + 1] = i32(np.where(rng.integers(0, 2, size=count) == 1, 1, -1))
    return r


def one(bits, exponent=0, sign=1):
    n = words(bits)
    r = np.zeros(n + 3, dtype=np.uint32)
    r[n - 1] = 0x80000000
    r[n] = exponent & MASK
    r[n + 1] = sign & MASK
    return r


def zero(bits, status=0):
    r = np.zeros(words(bits) + 3, dtype=np.uint32)
    r[-1] = status
    return r


def get_exp(r, bits):
    return r[..., words(bits)].astype(np.int64) - ((r[..., words(bits)].astype(np.int64) >> 31) << 32)


def negate(r, bits):
    r = r.copy()
    n = words(bits)
    s = get_sign(r, bits)
    r[..., n + 1] = i32(-s)
    return r


def get_sign(r, bits):
    n = words(bits)
    return r[..., n + 1].astype(np.int64) - ((r[..., n + 1].astype(np.int64) >> 31) << 32)


def binary_cases(rng, bits, count, span=1000):
    """(a, b) pairs with the reference's binary-op fixture classes mixed into random values."""
    n = words(bits)
    a = random_records(rng, bits, count, span)
    b = random_records(rng, bits, count, span)
    idx = np.arange(count)
    ea = get_exp(a, bits)
    # alignment gaps around limb and ulp boundaries
    m = idx % 3 == 0
    b[m, n] = i32(ea[m] - rng.integers(0, bits + 80, size=m.sum()))
    # cancellation: b = -a with the low limb perturbed
    m = idx % 11 == 0
    b[m] = negate(a[m], bits)
    b[m, 0] ^= rng.integers(0, 256, size=m.sum()).astype(np.uint32)
    a[idx % 17 == 0] = zero(bits)
    b[idx % 23 == 0] = zero(bits)
    a[idx % 101 == 0, n + 2] = 4
    b[idx % 103 == 5, n + 2] = 2
    fixed = []
    half_ulp = one(bits, -bits)
    o = one(bits)
    o2 = o.copy(); o2[0] |= 1
    fixed.append((o, half_ulp))                    # tie to even (stays)
    fixed.append((o2, half_ulp))                   # tie to even (rounds up)
    fixed.append((o, negate(half_ulp, bits)))      # rounding at a binade boundary
    o3 = o.copy(); o3[0] = 1
    fixed.append((o3, negate(o, bits)))            # exact deep cancellation
    fixed.append((one(bits, 1000000000), one(bits, 1)))    # product / sum overflow
    fixed.append((one(bits, -1000000000), one(bits, -1)))  # underflow
    ones = one(bits); ones[:n] = MASK
    fixed.append((ones, ones.copy()))
    fixed.append((one(bits, 999999999), one(bits, 999999999)))
    fixed.append((one(bits, -999999999, -1), one(bits, -2)))
    for k, (x, y) in enumerate(fixed):
        if k + 1 < count:
            a[k + 1], b[k + 1] = x, y
    # products just below a binade boundary (exponent at the lower limit)
    for k in range(12, min(count, 28)):
        A = int(rng.integers(0, 1 << 62)) | (1 << 63)
        Ai = (A << (bits - 64)) | int(rng.integers(0, 1 << 32))
        Ai |= 1 << (bits - 1)
        target = 1 << (2 * bits - 1)
        Bi = target // Ai
        if Bi.bit_length() != bits:
            continue
        a[k] = _from_int(Ai, bits, -1000000000 if k % 3 else 1000000000, 1 if k % 2 else -1)
        b[k] = _from_int(Bi, bits, -1 if k % 3 else 0, 1 if k % 2 else -1)
    return a, b


def _from_int(v, bits, exponent, sign):
    n = words(bits)
    r = np.zeros(n + 3, dtype=np.uint32)
    for i in range(n):
        r[i] = (v >> (32 * i)) & MASK
    r[n] = exponent & MASK
    r[n + 1] = sign & MASK
    return r


def root_cases(rng, bits, count):
    """Positive radicands with near-midpoint roots for both exponent parities (test_arithmetic.cpp:50-62),
    plus zeros, negatives and statuses."""
    n = words(bits)
    x = random_records(rng, bits, count, 1000)
    x[:, n + 1] = 1
    idx = np.arange(count)
    x[idx % 13 == 0, n + 1] = i32(-1)
    x[idx % 17 == 0] = zero(bits)
    x[idx % 29 == 0, n + 2] = 1
    for k in range(1, min(count, 40)):
        q = int(rng.integers(0, 1 << 62)) | (1 << 63)
        q = (q << (bits - 64)) | int(rng.integers(0, 1 << 32))
        q |= 1 << (bits - 1)
        z = q * q + q  # floor((q + 1/2)^2)
        parity = 1 if z.bit_length() == 2 * bits else 0
        z >>= bits - 1 + parity
        if k % 2:
            z += 1
        if z.bit_length() > bits:
            z -= 1
        x[k] = _from_int(z, bits, parity - (0 if k % 3 else 2), 1)
    return x


def fma_cases(rng, bits, count, oracle_mul):
    """(a, b, c) with the classes of tests/test_fused.cpp:10-28; oracle_mul(a, b) gives RN(a*b) records."""
    n = words(bits)
    a = random_records(rng, bits, count, 40)
    b = random_records(rng, bits, count, 40)
    c = random_records(rng, bits, count, 80)
    idx = np.arange(count)
    cls = idx % 23
    p = oracle_mul(a, b)
    m = cls == 0
    c[m] = negate(p[m], bits)
    m = cls == 1
    c[m] = negate(p[m], bits)
    c[m, 0] ^= rng.integers(0, 16, size=m.sum()).astype(np.uint32)
    ea, eb = get_exp(a, bits), get_exp(b, bits)
    m = cls == 2
    c[m, n] = i32(ea[m] + eb[m] - rng.integers(0, 3 * bits + 80, size=m.sum()))
    m = cls == 3
    c[m, n] = i32(ea[m] + eb[m] + rng.integers(0, 2 * bits + 80, size=m.sum()))
    for k in np.nonzero((cls == 4) | (cls == 5))[0]:
        a[k] = one(bits)
        h = one(bits, int(eb[k]) - bits, 1 if rng.integers(0, 2) else -1)
        if cls[k] == 5:
            h[0] |= 1
        c[k] = h
    a[cls == 6] = zero(bits)
    c[cls == 7] = zero(bits)
    b[cls == 8] = zero(bits)
    c[cls == 8] = zero(bits)
    c[cls == 9, n + 2] = 4
    m = cls == 10
    a[m, n] = 600000000
    b[m, n] = 600000000
    if m.any():
        c[m] = negate(oracle_mul(a[m], b[m]), bits)
    m = cls == 11
    a[m, n] = 600000000
    b[m, n] = 600000000
    c[m, n] = 999999990
    m = cls == 12
    a[m, n] = i32(-600000000)
    b[m, n] = i32(-500000000)
    c[m, n] = i32(-999999999)
    m = cls == 13
    a[m, n] = 0
    b[m, n] = 0
    c[m] = negate(a[m], bits)
    # far-apart terms in both directions across the window edges
    m = cls == 14
    c[m, n] = i32(ea[m] + eb[m] - 2 * bits - rng.integers(0, 200, size=m.sum()))
    m = cls == 15
    c[m, n] = i32(ea[m] + eb[m] + 2 * bits + rng.integers(0, 200, size=m.sum()))
    return a, b, c


def complex_cases(rng, bits, count, span=1000):
    n = words(bits)
    rw = n + 3
    a = np.concatenate([random_records(rng, bits, count, span), random_records(rng, bits, count, span)], axis=1)
    b = np.concatenate([random_records(rng, bits, count, span), random_records(rng, bits, count, span)], axis=1)
    idx = np.arange(count)
    b[idx % 5 == 1, rw:] = zero(bits)
    b[idx % 7 == 2, :rw] = zero(bits)
    b[idx % 7 == 2, rw:] = zero(bits)
    m = idx % 6 == 3
    a[m, rw:] = negate(a[m, :rw], bits)
    b[m, rw:] = b[m, :rw]
    m = idx % 9 == 4
    eb = get_exp(b[:, :rw], bits)
    b[m, rw + n] = i32(eb[m] - rng.integers(0, 2 * bits, size=m.sum()))
    a[idx % 19 == 5, n + 2] = 4
    return a, b
