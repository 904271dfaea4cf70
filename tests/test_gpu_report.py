"""LinalgReport diagnostics of the exact products (linalg.hpp:33-45) against a Python restatement of the
reference's planning rules: the greedy exponent bands of every line (src/linalg.mm:82-111: status OR,
exponent range, bands [hi - G, hi] from the largest exponent down with lo the smallest exponent inside,
max_bands and max_spread sending a line to the exact-dot fallback), the band histograms, the extended
GEMM sizes, the band pairs, and the modulus count (linalg.mm:341-345: the fewest primes <= 65279, taken in
descending order, whose log2 sum in double precision exceeds 2 bits + 2 pmax + ceil(log2 K) + 2, pmax
the widest band), with the int8 GEMM count 3 L.  The modulus rule is also checked exactly: the product
of the chosen primes exceeds 2^need and that of one fewer does not."""
import math

import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu


def primes(count=200):
    out, q = [], 65279
    while len(out) < count:
        if all(q % d for d in range(2, int(q ** 0.5) + 1)):
            out.append(q)
        q -= 1
    return out


PRIMES = primes()


def line_bands(exps, signs, stats, G, max_bands, max_spread):
    """(class, bands): class 0 banded, 1 trivial (zero or status line), 2 fallback."""
    if any(stats):
        return 1, []
    nz = [int(e) for e, s, st in zip(exps, signs, stats) if s != 0 and st == 0]
    if not nz:
        return 1, []
    top, bottom = max(nz), min(nz)
    if max_bands >= 1 and top - bottom <= G:
        return 0, [(bottom, top)]
    if top - bottom > max_spread:
        return 2, []
    bands, hi = [], top
    while True:
        lo, nxt = hi, None
        for e in nz:
            if e > hi:
                continue
            if e >= hi - G:
                lo = min(lo, e)
            else:
                nxt = e if nxt is None else max(nxt, e)
        bands.append((lo, hi))
        if len(bands) > max_bands:
            return 2, []
        if nxt is None:
            return 0, bands
        hi = nxt


def fields(X, bits):
    w = bits // 32
    e = X[:, w].astype(np.int64) - ((X[:, w].astype(np.int64) >> 31) << 32)
    s = X[:, w + 1].astype(np.int64) - ((X[:, w + 1].astype(np.int64) >> 31) << 32)
    return e, s, X[:, w + 2]


def expected(bits, A, B, m, n, k, G=64, max_bands=4, max_spread=512):
    ea, sa, ta = (f.reshape(m, k) for f in fields(A, bits))
    eb, sb, tb = (f.reshape(k, n) for f in fields(B, bits))
    left = [line_bands(ea[i], sa[i], ta[i], G, max_bands, max_spread) for i in range(m)]
    right = [line_bands(eb[:, j], sb[:, j], tb[:, j], G, max_bands, max_spread) for j in range(n)]

    def hist(side):
        h = [0] * (max_bands + 1)
        fb = 0
        for cls, b in side:
            if cls == 1:
                h[0] += 1
            elif cls == 2:
                fb += 1
            else:
                h[len(b)] += 1
        return h, fb

    pmax = max([hi - lo for cls, b in left + right if cls == 0 for lo, hi in b] or [0])
    logk = (k - 1).bit_length()  # ceil(log2 K)
    need = 2.0 * bits + 2 * pmax + logk + 2
    have, L = 0.0, 0
    while have <= need:
        have += math.log2(PRIMES[L])
        L += 1
    prod = math.prod(PRIMES[:L])
    assert prod > 2 ** int(need) and math.prod(PRIMES[:L - 1]) <= 2 ** int(need), "double and exact modulus rules disagree"
    lb = max([len(b) for cls, b in left if cls == 0] or [0])
    rb = max([len(b) for cls, b in right if cls == 0] or [0])
    return {"left": hist(left), "right": hist(right), "moduli": L, "pairs": lb * rb,
            "rows": sum(len(b) for cls, b in left if cls == 0), "cols": sum(len(b) for cls, b in right if cls == 0)}


def entries(rng, bits, count, kind):
    x = lfcases.random_records(rng, bits, count, 0)
    w = bits // 32
    if kind == "single":
        e = rng.integers(-30, 31, size=count)
    elif kind == "bands3":
        e = rng.choice([-200, 0, 200], size=count) + np.round(8 * rng.standard_normal(count)).astype(np.int64)
    elif kind == "two":
        e = rng.choice([-90, 40], size=count) + rng.integers(-5, 6, size=count)
    elif kind == "mixed":  # some lines beyond max_spread or max_bands, zero lines and status lines
        e = rng.choice([-900, -300, -150, 0, 100, 250], size=count) + rng.integers(-20, 21, size=count)
    else:
        raise ValueError(kind)
    x[:, w] = lfcases.i32(e)
    if kind == "mixed":
        x[rng.random(count) < 0.05] = lfcases.zero(bits)
    return x


@pytest.mark.parametrize("bits", [64, 512, 1024])
@pytest.mark.parametrize("kind", ["single", "bands3", "two", "mixed"])
@pytest.mark.parametrize("k", [1, 2, 7, 33, 300])
def test_report_matches_the_planning_rules(capi, bits, kind, k):
    rng = np.random.default_rng(bits + 31 * k + sum(map(ord, kind)))
    m, n = 6, 5
    A = entries(rng, bits, m * k, kind)
    B = entries(rng, bits, k * n, kind)
    if kind == "mixed":
        A[2 * k:3 * k] = lfcases.zero(bits)           # a zero line
        A[4 * k, bits // 32 + 2] = 2                   # a status line
    for G in (64, 16, 512):
        got, rep, _, _ = capi.gemm(bits, False, A, B, m, n, k, options=capi.LinalgOptions.default(band_bits=G))
        want = expected(bits, A, B, m, n, k, G=G)
        assert list(rep.bands("left")) == want["left"][0] and rep.left_fallback == want["left"][1], (G, list(rep.bands("left")))
        assert list(rep.bands("right")) == want["right"][0] and rep.right_fallback == want["right"][1]
        if want["rows"] and want["cols"]:
            assert rep.moduli == want["moduli"], (G, rep.moduli, want["moduli"])
            assert rep.int8_gemms == 3 * want["moduli"]
            assert rep.npairs == min(64, want["pairs"])
            assert rep.gemm_rows == want["rows"] and rep.gemm_cols == want["cols"]
