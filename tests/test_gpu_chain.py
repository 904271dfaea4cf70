"""GPU parity of the resident multiplication chain (benchmarks/benchmark.cpp:67-94; tests/test_buffers.cpp;
examples/resident.cpp): the MPFR fixtures (1024, 256 bits and the 1.01^16 known answer at 384 bits),
and the full 1024-bit 65,536 x 16 chain bit-exact against the oracle."""
import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu


def chain_gpu(capi, bits, x, f, steps, f_index=None):
    X = capi.Buffer(bits, x.shape[0])
    F = capi.Buffer(bits, f.shape[0])
    X.upload(x)
    F.upload(f)
    capi.mul_chain(bits, X, F, X, steps, f_index)
    return X.download()


def test_golden_chains(capi, oracle):
    for s in oracle.read_golden("chain.bin"):
        print(f"[gpu] golden chain bits={s['bits']}", flush=True)
        got = chain_gpu(capi, s["bits"], s["x"], s["f"], s["steps"])
        assert (got == s["out"]).all(), s["bits"]


@pytest.mark.parametrize("bits", [64, 96, 512, 544, 1024])
def test_chain_matches_separate_launches(capi, oracle, bits):
    """One fused launch equals `steps` dependent element-wise launches and the oracle."""
    print(f"[gpu] chain vs launches bits={bits}", flush=True)
    rng = np.random.default_rng(bits)
    x = lfcases.random_records(rng, bits, 3000, 64)
    f = lfcases.random_records(rng, bits, 3000, 64)
    x[::97] = lfcases.zero(bits)
    f[5::101, -1] = 2
    x[7::89, bits // 32] = 999999990  # overflows within a few steps
    fused = chain_gpu(capi, bits, x, f, 16)
    X = capi.Buffer(bits, 3000)
    F = capi.Buffer(bits, 3000)
    X.upload(x)
    F.upload(f)
    for _ in range(16):
        capi.elementwise(bits, "mul", X, F, None, X)
    assert (fused == X.download()).all()
    assert (fused == oracle.mul_chain(bits, x, f, 16)).all()


def test_full_1024_bit_chain(capi, oracle):
    """The headline workload: 65,536 values x 16 steps at 1024 bits, exponents in [-64, 64]."""
    bits, count = 1024, 65536
    print("[gpu] full 1024-bit chain", flush=True)
    rng = np.random.default_rng(20261007 + bits)
    x = lfcases.random_records(rng, bits, count, 64)
    f = lfcases.random_records(rng, bits, count, 64)
    got = chain_gpu(capi, bits, x, f, 16)
    want = oracle.mul_chain(bits, x, f, 16)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, bad[:8]


def test_broadcast_factor(capi, oracle):
    bits = 256
    rng = np.random.default_rng(4)
    x = lfcases.random_records(rng, bits, 1000, 64)
    f = lfcases.random_records(rng, bits, 64, 64)
    got = chain_gpu(capi, bits, x, f, 5, (1, 64))
    ff = f[np.arange(1000) % 64]
    assert (got == oracle.mul_chain(bits, x, ff, 5)).all()
