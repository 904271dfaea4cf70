"""The C++17 facade (include/limbforge/*.hpp, liblimbforge.so): a consumer program that uses the API as
the reference's tests do compiles against the headers (CPU), and runs bit-exact on the GPU."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "limbforge_amd" / "lib"
EXE = ROOT / "tests" / "cpp" / "test_engine"


def build_consumer(out: Path):
    cmd = ["g++", "-O1", "-std=c++17", f"-I{ROOT/'include'}", f"-I{ROOT/'oracle'}", str(ROOT / "tests" / "cpp" / "test_engine.cpp"),
           f"-L{LIB}", "-llimbforge", "-llimbforge_hip", f"-Wl,-rpath,{LIB}", f"-L{ROOT/'oracle'}", "-llf_oracle",
           f"-Wl,-rpath,{ROOT/'oracle'}", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)


def test_consumer_compiles_against_headers(tmp_path, oracle):
    assert (LIB / "liblimbforge.so").exists(), "liblimbforge.so is not built"
    build_consumer(tmp_path / "test_engine")


@pytest.mark.gpu
def test_engine_api_on_gpu(tmp_path, oracle):
    exe = tmp_path / "test_engine"
    build_consumer(exe)
    print("[gpu] C++ Engine API test", flush=True)
    p = subprocess.run([str(exe), str(ROOT / "tests" / "golden")], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    assert "passed 0 failures" in p.stdout
