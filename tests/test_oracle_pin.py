"""Pins the CPU oracle: bit-for-bit against MPFR on the reference tests' fixture classes and random
inputs at all 31 widths, divide_word against 64-bit division, exact dot products against one mpfr_sum,
and against the reference's own core.hpp / exact_dot_add (oracle/_ref/libref.so) where that was built."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_oracle_matches_mpfr_and_reference(oracle):
    pin = ROOT / "oracle" / "mpfr_pin"
    assert pin.exists(), "oracle/mpfr_pin is not built (make -C oracle)"
    cmd = [str(pin), "all", "400"]
    ref = ROOT / "oracle" / "_ref" / "libref.so"
    if ref.exists():
        cmd.append(str(ref))
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-4000:]
    assert " 0 failures" in p.stdout


def test_divide_word_closed_form(oracle):
    import numpy as np

    rng = np.random.default_rng(5)
    for _ in range(2000):
        d = int(rng.integers(1 << 31, 1 << 32))
        high = int(rng.integers(0, d))
        low = int(rng.integers(0, 1 << 32))
        inv = ((1 << 64) - 1) // d - (1 << 32)
        qr = oracle.lib().lfo_divide_word(high, low, d, inv)
        num = (high << 32) | low
        assert qr & 0xFFFFFFFF == num // d and qr >> 32 == num % d


def test_factor_oracle_matches_mpfr_replay(oracle):
    """lfo_cholesky / lfo_solve (the reference's right-looking host order) against an independent
    left-looking MPFR program of the documented rounding sequence (oracle/mpfr_factor_pin.c)."""
    pin = ROOT / "oracle" / "mpfr_factor_pin"
    assert pin.exists(), "oracle/mpfr_factor_pin is not built (make -C oracle)"
    p = subprocess.run([str(pin), "3"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-4000:]
    assert " 0 failures" in p.stdout
