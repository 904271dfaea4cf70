"""include/limbforge/mpfr_bridge.hpp (the reference's mpfr_bridge.hpp contract): tests/bridge_check.cpp built
against MPFR/MPC and run on the CPU -- round trips, RN-even narrowing bit for bit against the oracle's MPFR
conversion, ties, NaN/Inf/zero, exponent limits, the status rule, the array and runtime-width forms on 1, 4
and an automatic number of threads, and the mpc_t forms, at 64, 96, 256, 512 and 1024 bits."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
MPFR = Path("/opt/conda")


def test_bridge_conversions(tmp_path):
    exe = tmp_path / "bridge_check"
    cmd = ["g++", "-O1", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{ROOT / 'oracle'}", f"-I{MPFR / 'include'}",
           str(ROOT / "tests" / "bridge_check.cpp"), "-x", "c", str(ROOT / "oracle" / "lf_oracle.c"), "-x", "none",
           str(MPFR / "lib" / "libmpc.so"), str(MPFR / "lib" / "libmpfr.so"), str(MPFR / "lib" / "libgmp.so"), f"-Wl,-rpath,{MPFR / 'lib'}",
           # the MPFR directory also holds an older libstdc++: link the compiler's own statically
           "-static-libstdc++", "-static-libgcc", "-pthread", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0 and " 0 failures" in p.stdout, p.stdout[-3000:]
