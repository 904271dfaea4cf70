"""The oracle against every committed golden fixture (MPFR-expected values, tests/golden/gen_golden.c),
including the reference's known answers: 1.01^16 at 384 bits (examples/resident.cpp:8-17) and
[2^100, 1, -2^100] . [1, 1, 1] = 1 (tests/test_contract_names.cpp:14-19)."""
import numpy as np
import pytest

ELEMENTWISE = ["elementwise_%d.bin" % b for b in (64, 96, 256, 384, 512, 1024)]


@pytest.mark.parametrize("name", ELEMENTWISE)
def test_elementwise_fixtures(oracle, name):
    for s in oracle.read_golden(name):
        got = oracle.elementwise(s["bits"], s["op"], s["a"], s["b"], s["c"])
        bad = np.nonzero((got != s["out"]).any(axis=1))[0]
        assert bad.size == 0, (name, s["op"], bad[:5])


def test_chain_fixtures(oracle):
    secs = oracle.read_golden("chain.bin")
    assert [s["bits"] for s in secs] == [1024, 256, 384]
    for s in secs:
        got = oracle.mul_chain(s["bits"], s["x"], s["f"], s["steps"])
        assert (got == s["out"]).all(), s["bits"]


@pytest.mark.parametrize("name", ["dot_128.bin", "dot_512.bin"])
def test_dot_fixtures(oracle, name):
    for s in oracle.read_golden(name):
        c = s["c"] if s["mode"] else None
        got = oracle.exact_dot_add(s["bits"], c, s["mode"] == 2, s["a"], 1, s["b"], 1, s["K"])
        assert (got == s["out"]).all(), (name, s["K"], s["mode"])


def test_product_fixtures(oracle):
    for i, s in enumerate(oracle.read_golden("products.bin")):
        if s["syrk"]:
            got = oracle.syrk(s["bits"], s["A"], s["k"], s["m"], s["C_in"], s["lower_only"], s["subtract"])
        else:
            got = oracle.gemm(s["bits"], s["transpose_a"], s["A"], s["B"], s["m"], s["n"], s["k"], s["C_in"], s["subtract"])
        assert (got == s["C_out"]).all(), i


def test_known_answer_contract_names(oracle):
    s = oracle.read_golden("products.bin")[4]
    assert s["bits"] == 64 and s["k"] == 3
    n = 2
    # result is exactly 1: mantissa 0x80000000 00000000, exponent 0, sign 1
    assert list(s["C_out"][0]) == [0, 0x80000000, 0, 1, 0]
