"""GPU parity of the cooperative multiplication chain (2 or 4 lanes of a wavefront per value) through the
C ABI: bit for bit the one-lane chain and the oracle, over widths from 256 to 1024 bits, batch sizes that
leave partial wavefronts and partial groups, zero, status and overflowing values, exact products with ties
(which take the exact column fallback), a broadcast factor and an aliased output; then the lanes tuner."""
import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu


def chain(capi, bits, x, f, steps, lanes, f_index=None, alias=True):
    X = capi.Buffer(bits, x.shape[0])
    F = capi.Buffer(bits, f.shape[0])
    X.upload(x)
    F.upload(f)
    if alias:
        capi.mul_chain(bits, X, F, X, steps, f_index, lanes=lanes)
        return X.download()
    O = capi.Buffer(bits, x.shape[0])
    capi.mul_chain(bits, X, F, O, steps, f_index, lanes=lanes)
    return O.download()


def same(got, want, label):
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} mismatching values, first {bad[:8].tolist()}"


def hard_cases(rng, bits, count):
    """Random values with zeros, a status bit, overflowing exponents, and exact products with ties."""
    w = bits // 32
    x = lfcases.random_records(rng, bits, count, 64)
    f = lfcases.random_records(rng, bits, count, 64)
    x[::97] = lfcases.zero(bits)
    f[5::101, -1] = 2
    x[7::89, w] = 999999990
    # f = 1.5 exactly (one digit): 1.5 x needs one bit more than x holds, a tie when x's lowest bit is set
    f[11::23, :w] = 0
    f[11::23, w - 1] = 0xC0000000
    f[11::23, w] = 1
    x[11::23, 0] |= 1
    # f = 1 + 2^-(32w - 1): products that are exact only past the short cut
    f[13::29, :w] = 0
    f[13::29, w - 1] = 0x80000000
    f[13::29, 0] = 1
    f[13::29, w] = 1
    return x, f


@pytest.mark.parametrize("bits", [256, 288, 512, 800, 1024])
@pytest.mark.parametrize("lanes", [2, 4])
def test_lanes_match_one_lane_and_oracle(capi, oracle, bits, lanes):
    print(f"[gpu] coop chain bits={bits} lanes={lanes}", flush=True)
    rng = np.random.default_rng(bits * 10 + lanes)
    for count in (1, 3, 17, 63, 130, 1001):
        x, f = hard_cases(rng, bits, count)
        for steps in (0, 1, 9):
            got = chain(capi, bits, x, f, steps, lanes, alias=count % 2 == 1)
            one = chain(capi, bits, x, f, steps, 1)
            same(got, one, f"bits={bits} lanes={lanes} count={count} steps={steps} vs one lane")
            same(got, oracle.mul_chain(bits, x, f, steps), f"bits={bits} lanes={lanes} count={count} steps={steps} vs oracle")


@pytest.mark.parametrize("lanes", [2, 4])
def test_lanes_full_1024_bit_chain(capi, oracle, lanes):
    """65,536 values x 16 steps at 1024 bits, exponents in [-64, 64], against the oracle."""
    bits, count = 1024, 65536
    print(f"[gpu] coop full 1024-bit chain lanes={lanes}", flush=True)
    rng = np.random.default_rng(20261007 + bits)
    x = lfcases.random_records(rng, bits, count, 64)
    f = lfcases.random_records(rng, bits, count, 64)
    same(chain(capi, bits, x, f, 16, lanes), oracle.mul_chain(bits, x, f, 16), f"lanes={lanes} full chain")


@pytest.mark.parametrize("lanes", [2, 4])
def test_lanes_broadcast_factor(capi, oracle, lanes):
    bits = 640
    print(f"[gpu] coop broadcast bits={bits} lanes={lanes}", flush=True)
    rng = np.random.default_rng(lanes)
    x = lfcases.random_records(rng, bits, 700, 64)
    f = lfcases.random_records(rng, bits, 64, 64)
    got = chain(capi, bits, x, f, 5, lanes, (3, 64), alias=False)
    ff = f[(np.arange(700) // 3) % 64]
    same(got, oracle.mul_chain(bits, x, ff, 5), "broadcast")


def test_lanes_arguments(capi):
    bits = 128
    X = capi.Buffer(bits, 8)
    with pytest.raises(ValueError, match="at least 256 bits"):
        capi.mul_chain(bits, X, X, X, 1, lanes=2)
    X = capi.Buffer(512, 8)
    with pytest.raises(ValueError, match="lanes per value"):
        capi.mul_chain(512, X, X, X, 1, lanes=3)


def test_autotune_chain_choice_is_used(capi, oracle):
    """The tuner times lanes 1, 2, 4; lanes 0 then takes its choice for that width and batch-size class,
    and the result stays the one-lane result."""
    bits = 768
    rng = np.random.default_rng(77)
    for count in (256, 4096):
        lanes, ms = capi.autotune_chain(bits, count, 16)
        print(f"[gpu] autotune bits={bits} count={count}: lanes={lanes} ms={ms}", flush=True)
        assert lanes in (1, 2, 4) and all(t > 0 for t in ms)
        assert ms[(1, 2, 4).index(lanes)] == min(ms)
        x, f = hard_cases(rng, bits, count)
        same(chain(capi, bits, x, f, 7, 0), chain(capi, bits, x, f, 7, 1), f"tuned count={count}")
    lanes, ms = capi.autotune_chain(128, 512, 4)
    assert lanes == 1 and ms[1] == -1 and ms[2] == -1
    with pytest.raises(ValueError):
        capi.autotune_chain(bits, 0, 4)
