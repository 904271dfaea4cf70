"""Resident products (lf_gemm / lf_syrk on device limb planes: Linalg::gemm / syrk on Buffers,
linalg.hpp:187-190) against the host-array calls, bit for bit with the same LinalgReport (the reference's
tests/test_resident_linalg.cpp:51-64): transposed and update forms, multi-band and fallback lines, SYRK
lower and full, at several widths."""
import numpy as np
import pytest

import lfcases
from test_gpu_gemm import matrix

pytestmark = pytest.mark.gpu

FIELDS = ("left_fallback", "right_fallback", "gpu_outputs", "multi_band_outputs", "fallback_outputs", "trivial_outputs", "npairs",
          "gemm_rows", "gemm_cols", "int8_gemms", "moduli", "crt_words")


def same_report(a, b):
    assert a.bands("left") == b.bands("left") and a.bands("right") == b.bands("right")
    for f in FIELDS:
        assert getattr(a, f) == getattr(b, f), f


@pytest.mark.parametrize("bits", [64, 256, 512])
@pytest.mark.parametrize("kind", ["bands3", "wide", "sparse"])
def test_resident_gemm_equals_host(capi, oracle, bits, kind):
    rng = np.random.default_rng(bits + len(kind))
    m, n, k = 13, 17, 40
    print(f"[gpu] resident gemm bits={bits} kind={kind}", flush=True)
    for ta, sub in ((False, False), (True, True)):
        A = matrix(rng, bits, k * m if ta else m * k, 1, kind).reshape(-1, bits // 32 + 3)
        B = matrix(rng, bits, k, n, kind)
        C0 = lfcases.random_records(rng, bits, m * n, 200)
        want, hrep, _, _ = capi.gemm(bits, ta, A, B, m, n, k, C0 if sub else None, sub)
        bA, bB, bC = capi.Buffer(bits, A.shape[0]), capi.Buffer(bits, B.shape[0]), capi.Buffer(bits, m * n)
        bA.upload(A)
        bB.upload(B)
        bC.upload(C0 if sub else np.zeros_like(C0))
        rep, _ = capi.gemm_resident(bits, ta, bA, bB, m, n, k, bC, sub)
        capi.lib().lf_stream_synchronize(None)
        got = bC.download()
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"transpose={ta} subtract={sub}: {bad.size} outputs differ from the host-array call"
        same_report(rep, hrep)


@pytest.mark.parametrize("bits", [96, 384])
def test_resident_syrk_equals_host(capi, oracle, bits):
    rng = np.random.default_rng(bits)
    rows, cols = 45, 21
    A = matrix(rng, bits, rows, cols, "bands3")
    C0 = lfcases.random_records(rng, bits, cols * cols, 100)
    for lower, sub in ((True, False), (False, False), (True, True)):
        print(f"[gpu] resident syrk bits={bits} lower={lower} subtract={sub}", flush=True)
        want, hrep, _, _ = capi.syrk(bits, A, rows, cols, C0, lower, sub)
        bA, bC = capi.Buffer(bits, A.shape[0]), capi.Buffer(bits, cols * cols)
        bA.upload(A)
        bC.upload(C0)
        rep, _ = capi.syrk_resident(bits, bA, rows, cols, bC, lower, sub)
        capi.lib().lf_stream_synchronize(None)
        assert (bC.download() == want).all(), f"lower={lower} subtract={sub}"
        same_report(rep, hrep)
