"""Build recipe for liblimbforge_hip.so (HIP kernels + C ABI, gfx950) and the C++17 facade.

Usage: python -m limbforge_amd.build [--jobs N] [--only hip|cxx]

Everything is compiled in-tree: objects under limbforge_amd/_build/, libraries under
limbforge_amd/lib/.  The element-wise kernels are compiled once per limb count (2..32) so that the 31
widths build in parallel.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the limb loops are fully unrolled (1024 products at 1024 bits) so that every limb stays in a VGPR
    "-mllvm", "-pragma-unroll-threshold=1000000",
    f"-I{CSRC}",
    f"-I{INCLUDE}",
]
WIDTHS = range(2, 33)


def _run(cmd: list[str]) -> tuple[list[str], int, str]:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return cmd, p.returncode, p.stdout


def _hip_jobs() -> list[tuple[Path, list[str]]]:
    jobs = []
    # sources compiled once per limb count (LF_WORDS), and once
    per_width = ("elementwise_w", "gemm_w", "factor_w", "mixed_w", "reduce_w")
    # every other source is one translation unit
    once = sorted(p.stem for p in CSRC.glob("*.hip") if p.stem not in per_width + ("mixed_group",))
    for n in WIDTHS:
        for name in per_width:
            out = BUILD / f"{name}{n}.o"
            jobs.append((out, [HIPCC, *HIP_FLAGS, f"-DLF_WORDS={n}", "-c", str(CSRC / f"{name}.hip"), "-o", str(out)]))
    for name in once:
        out = BUILD / f"{name}.o"
        jobs.append((out, [HIPCC, *HIP_FLAGS, "-c", str(CSRC / f"{name}.hip"), "-o", str(out)]))
    # the mixed-precision group kernels: one translation unit per group of eight widths
    for grp in range(4):
        out = BUILD / f"mixed_group{grp}.o"
        jobs.append((out, [HIPCC, *HIP_FLAGS, f"-DLF_GROUP={grp}", "-c", str(CSRC / "mixed_group.hip"), "-o", str(out)]))
    return jobs


def _deps(src: Path, seen=None) -> set[Path]:
    """The source and every quoted header it includes from csrc/ or include/, recursively."""
    seen = set() if seen is None else seen
    if src in seen or not src.exists():
        return seen
    seen.add(src)
    for line in src.read_text().splitlines():
        line = line.strip()
        if line.startswith("#include \""):
            name = line.split('"')[1]
            for d in (CSRC, INCLUDE):
                if (d / name).exists():
                    _deps(d / name, seen)
    return seen


def _stale(out: Path, cmd: list[str]) -> bool:
    if not out.exists():
        return True
    src = Path(cmd[cmd.index("-c") + 1])
    newest = max(p.stat().st_mtime for p in _deps(src) | {Path(__file__)})
    return out.stat().st_mtime < newest


def build_hip(jobs: int, incremental: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    every = _hip_jobs()
    work = [j for j in every if not incremental or _stale(*j)]
    # the heaviest translation units first
    work.sort(key=lambda j: -int(j[0].stem.split("_w")[-1]) if "_w" in j[0].stem else (-200 if "group" in j[0].stem else -100))
    failures = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, rc, log in ex.map(lambda j: _run(j[1]), work):
            if rc != 0:
                failures.append((cmd, log))
    if failures:
        for cmd, log in failures:
            sys.stderr.write(" ".join(cmd) + "\n" + log + "\n")
        raise RuntimeError(f"{len(failures)} HIP translation unit(s) failed to compile")
    objs = [str(o) for o, _ in every]
    lib = LIB / "liblimbforge_hip.so"
    tmp = LIB / "liblimbforge_hip.so.tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    _, rc, log = _run(cmd)
    if rc != 0:
        raise RuntimeError("link failed:\n" + log)
    os.replace(tmp, lib)
    return lib


def build_cxx(jobs: int) -> Path | None:
    """The C++17 facade (include/limbforge/*.hpp) over the C ABI: liblimbforge.so."""
    srcs = sorted(p for p in HOST.glob("*.cpp") if p.stem != "core_host_w")
    if not srcs:
        return None
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", "-pthread"]
    work = []
    # the scalar arithmetic of core.hpp: the GPU kernels' arithmetic core built for the host, per width
    for n in WIDTHS:
        out = BUILD / f"host_core{n}.o"
        work.append((out, [cxx, *flags, "-Wno-unknown-pragmas", "-DLF_HOST_BUILD", f"-DLF_WORDS={n}", f"-I{CSRC}", "-c", str(HOST / "core_host_w.cpp"),
                           "-o", str(out)]))
    work.sort(key=lambda j: -int(j[0].stem[len("host_core"):]))
    for s in srcs:
        out = BUILD / f"host_{s.stem}.o"
        work.append((out, [cxx, *flags, "-c", str(s), "-o", str(out)]))
    failures = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, rc, log in ex.map(lambda j: _run(j[1]), work):
            if rc != 0:
                failures.append((cmd, log))
    if failures:
        for cmd, log in failures:
            sys.stderr.write(" ".join(cmd) + "\n" + log + "\n")
        raise RuntimeError("C++ facade failed to compile")
    lib = LIB / "liblimbforge.so"
    cmd = [cxx, "-shared", "-o", str(lib), *[str(o) for o, _ in work], f"-L{LIB}", "-llimbforge_hip",
           "-Wl,-rpath,$ORIGIN", "-pthread"]
    _, rc, log = _run(cmd)
    if rc != 0:
        raise RuntimeError("C++ facade link failed:\n" + log)
    return lib


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(16, os.cpu_count() or 4))
    ap.add_argument("--only", choices=["hip", "cxx"], default=None)
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("--incremental", action="store_true", help="recompile only stale objects")
    a = ap.parse_args(argv)
    if a.clean and BUILD.exists():
        shutil.rmtree(BUILD)
    if a.only in (None, "hip"):
        print("built", build_hip(a.jobs, a.incremental))
    if a.only in (None, "cxx"):
        lib = build_cxx(a.jobs)
        if lib:
            print("built", lib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
