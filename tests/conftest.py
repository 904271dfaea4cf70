import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (built by `make -C oracle`; built here only when nothing is built)."""
    lib = ROOT / "oracle" / "liblf_oracle.so"
    if not lib.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "liblf_oracle.so", "mpfr_pin", "mpfr_factor_pin"], check=True)
    from oracle import lf_oracle

    lf_oracle.lib()
    return lf_oracle


@pytest.fixture(scope="session")
def capi():
    """The HIP library; GPU tests fail (not skip) when it is missing."""
    from limbforge_amd import capi as c

    c.lib()
    return c
