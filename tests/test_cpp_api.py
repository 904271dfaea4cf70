"""The C++17 facade (include/limbforge/*.hpp, liblimbforge.so): consumer programs that use the API as the
reference's tests do compile against the headers (CPU), and run bit-exact on the GPU."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "limbforge_amd" / "lib"
PROGRAMS = ["test_engine", "test_linalg", "test_resident"]


def build_consumer(name: str, out: Path):
    cmd = ["g++", "-O1", "-std=c++17", f"-I{ROOT/'include'}", f"-I{ROOT/'oracle'}", str(ROOT / "tests" / "cpp" / f"{name}.cpp"),
           f"-L{LIB}", "-llimbforge", "-llimbforge_hip", f"-Wl,-rpath,{LIB}", f"-L{ROOT/'oracle'}", "-llf_oracle",
           f"-Wl,-rpath,{ROOT/'oracle'}", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("name", PROGRAMS)
def test_consumer_compiles_against_headers(tmp_path, oracle, name):
    assert (LIB / "liblimbforge.so").exists(), "liblimbforge.so is not built"
    build_consumer(name, tmp_path / name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROGRAMS)
def test_api_on_gpu(tmp_path, oracle, name):
    exe = tmp_path / name
    build_consumer(name, exe)
    print(f"[gpu] C++ API program {name}", flush=True)
    p = subprocess.run([str(exe), str(ROOT / "tests" / "golden")], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    assert "passed 0 failures" in p.stdout
