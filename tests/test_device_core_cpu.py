"""The device arithmetic core (limbforge_amd/csrc/limb_core.hpp) built for the host (LF_HOST_BUILD, the same
target the product's host arithmetic uses), checked bit for bit against the oracle at every width on the
reference tests' case classes.  Catches logic errors before any GPU run."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_device_core_matches_oracle_on_host(tmp_path, oracle):
    exe = tmp_path / "devcore"
    cmd = ["g++", "-O1", "-std=c++17", "-DLF_HOST_BUILD", f"-I{ROOT/'limbforge_amd'/'csrc'}",
           f"-I{ROOT/'oracle'}", "-I/opt/conda/include", str(ROOT / "tests" / "device_core_cpu.cpp"),
           "-x", "c", str(ROOT / "oracle" / "lf_oracle.c"), "-x", "none", "-L/opt/conda/lib",
           "-Wl,-rpath,/opt/conda/lib", "-lgmp", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    p = subprocess.run([str(exe), "2", "32", "250"], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:]
    assert " 0 failures" in p.stdout
