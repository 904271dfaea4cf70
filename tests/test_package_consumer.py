"""The CMake package: configure and install LimbForge (CMakeLists.txt at the repository root) into a scratch
prefix, build tests/package_consumer against it with find_package(LimbForge), and check what it computes on
the host -- the scalar arithmetic of limbforge/core.hpp (add ... cdiv, cfma, cfms) and exact_dot_add -- bit for bit against
the oracle at 64, 96, 256, 512 and 1024 bits.  No GPU: both run on the CPU."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import lfcases

ROOT = Path(__file__).resolve().parents[1]
WIDTHS = (2, 3, 8, 16, 32)
OPS = ("add", "sub", "mul", "div", "sqrt", "fma", "fms", "square")
CODES = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 8, "fma": 9, "fms": 10, "square": 7}


def version():
    text = (ROOT / "include" / "limbforge" / "version.hpp").read_text()
    parts = [text.split(f"#define LIMBFORGE_VERSION_{p} ")[1].split()[0] for p in ("MAJOR", "MINOR", "PATCH")]
    return ".".join(parts), parts[0]


def run(cmd, cwd=None):
    p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=900)
    assert p.returncode == 0, f"{' '.join(map(str, cmd))}\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}"
    return p.stdout


@pytest.fixture(scope="module")
def consumer(tmp_path_factory):
    if not (ROOT / "limbforge_amd" / "lib" / "liblimbforge_hip.so").exists():
        pytest.fail("liblimbforge_hip.so is not built: run python -m limbforge_amd.build")
    if not shutil.which("cmake"):
        pytest.skip("cmake is not installed")
    tmp = tmp_path_factory.mktemp("package")
    jobs = str(min(8, os.cpu_count() or 2))
    run(["cmake", "-S", str(ROOT), "-B", str(tmp / "build"), f"-DCMAKE_INSTALL_PREFIX={tmp / 'prefix'}", "-DCMAKE_BUILD_TYPE=Release"])
    run(["cmake", "--build", str(tmp / "build"), "-j", jobs])
    run(["cmake", "--install", str(tmp / "build")])
    ver, api = version()
    run(["cmake", "-S", str(ROOT / "tests" / "package_consumer"), "-B", str(tmp / "consumer"), f"-DCMAKE_PREFIX_PATH={tmp / 'prefix'}",
         f"-DLF_EXPECTED_VERSION={ver}", f"-DLF_EXPECTED_API={api}", f"-DLF_REQUIRED_VERSION={ver}"])
    run(["cmake", "--build", str(tmp / "consumer"), "-j", jobs])
    return tmp / "consumer" / "package_consumer", tmp


def test_host_arithmetic_and_exact_dot_vs_oracle(consumer, oracle):
    exe, tmp = consumer
    rng = np.random.default_rng(2026)
    C, D = 120, 6
    words = [np.array([C], np.uint32)]
    cases = {}
    for n in WIDTHS:
        bits = 32 * n
        a, b = lfcases.binary_cases(rng, bits, C, 300)
        c = lfcases.random_records(rng, bits, C, 300)
        c[::9] = lfcases.zero(bits)
        cases[n] = (a, b, c)
        words += [a.ravel(), b.ravel(), c.ravel()]
    words.append(np.array([D], np.uint32))
    dots = {}
    for n in WIDTHS:
        bits = 32 * n
        dots[n] = []
        for d in range(D):
            K = (0, 1, 3, 17, 40, 200)[d]
            span = (40, 40, 40, 2000, 600000000, 300)[d]
            a = lfcases.random_records(rng, bits, K, span)
            b = lfcases.random_records(rng, bits, K, span)
            if d == 2 and K:  # [2^100, 1, -2^100] . [1, 1, 1] style cancellation
                b[:] = lfcases.one(bits)
                a[2] = lfcases.negate(a[0], bits)
            c = lfcases.random_records(rng, bits, 1, span)[0]
            subtract, has_c = d % 2, int(d != 1)
            dots[n].append((K, subtract, has_c, c, a, b))
            words += [np.array([K, subtract, has_c], np.uint32), c, a.ravel(), b.ravel()]
    (tmp / "in.bin").write_bytes(np.concatenate(words).astype(np.uint32).tobytes())
    print(run([str(exe), str(tmp / "in.bin"), str(tmp / "out.bin")]))
    got = np.frombuffer((tmp / "out.bin").read_bytes(), dtype=np.uint32)
    at = 0
    for n in WIDTHS:
        bits, rw = 32 * n, n + 3
        a, b, c = cases[n]
        per = len(OPS) * rw + 5 * 2 * rw
        block = got[at:at + C * per].reshape(C, per)
        at += C * per
        for k, op in enumerate(OPS):
            want = oracle.elementwise(bits, CODES[op], a, b, c)
            mine = block[:, k * rw:(k + 1) * rw]
            bad = np.nonzero((mine != want).any(axis=1))[0]
            assert bad.size == 0, f"{bits}-bit host {op}: {bad.size} cases differ from the oracle, first {bad[:5].tolist()}"
        x = np.concatenate([a, b], axis=1)
        y = np.concatenate([c, b], axis=1)
        base = len(OPS) * rw
        z = np.concatenate([c, a], axis=1)
        for k, code in enumerate((4, 5, 6, 11, 12)):
            want = oracle.elementwise(bits, code, x, y, z)
            mine = block[:, base + k * 2 * rw:base + (k + 1) * 2 * rw]
            bad = np.nonzero((mine != want).any(axis=1))[0]
            assert bad.size == 0, f"{bits}-bit host complex op {code}: {bad.size} cases differ from the oracle"
    for n in WIDTHS:
        bits, rw = 32 * n, n + 3
        for d, (K, subtract, has_c, c, a, b) in enumerate(dots[n]):
            want = oracle.exact_dot_add(bits, c if has_c else None, subtract, a, 1, b, 1, K)
            mine = got[at:at + rw]
            at += rw
            assert (mine == want).all(), f"{bits}-bit exact_dot_add problem {d} (K={K}) differs from the oracle"
    assert at == got.size
