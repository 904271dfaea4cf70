"""Segmented dot products and the adjacent-pair tree sum on the GPU (lf_segmented_dot, lf_segmented_dot_host,
lf_tree_sum) against the oracle's step sequences (docs/numerics.md "Segmented dot products" and "Pairwise
reductions"; kernels.metal:61-112,197-208), real and complex, separate and shared tables, empty segments,
statuses, cancellation, and tree sizes around the 512-value pass boundary."""
import numpy as np
import pytest

import lfcases

pytestmark = pytest.mark.gpu


def values(rng, bits, count, cplx, span=60):
    x = lfcases.random_records(rng, bits, count, span)
    if cplx:
        x = np.concatenate([x, lfcases.random_records(rng, bits, count, span)], axis=1)
    return x


@pytest.mark.parametrize("bits", [64, 224, 256, 384, 1024])
@pytest.mark.parametrize("cplx", [False, True])
def test_segmented_dot_vs_oracle(capi, oracle, bits, cplx):
    rng = np.random.default_rng(bits * 3 + cplx)
    rw = capi.record_words(bits, cplx)
    for segments, length, shared in ((37, 24, False), (300, 7, True), (5, 0, False), (64, 120, False), (1, 1, True)):
        print(f"[gpu] segmented dot bits={bits} complex={cplx} {segments}x{length} shared={shared}", flush=True)
        a = values(rng, bits, segments * length, cplx)
        b = values(rng, bits, length if shared else segments * length, cplx)
        if a.shape[0] > 4:
            a[3, bits // 32 + 2] = 4            # a status rides along its segment
            a[1] = 0                           # a zero term
        if not shared and length > 2:
            b[1::length] = b[0::length]        # cancellation inside segments
            a[1::length] = lfcases.negate(a[0::length], bits) if not cplx else a[1::length]
        want = oracle.segmented_dot(bits, cplx, segments, length, shared, a, b)
        got, _, _ = capi.segmented_dot_host(bits, cplx, segments, length, a, b, shared)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"host-array form: {bad.size} segments differ, first {bad[:5].tolist()}"
        if segments * length:
            A = capi.Buffer(bits, a.shape[0], cplx)
            A.upload(a)
            B = capi.Buffer(bits, b.shape[0], cplx)
            B.upload(b)
            O = capi.Buffer(bits, segments, cplx)
            capi.segmented_dot(bits, segments, length, A, B, O, shared)
            capi.lib().lf_stream_synchronize(None)
            assert (O.download() == want).all(), "resident form differs from the oracle"
            assert rw == a.shape[1]


@pytest.mark.parametrize("bits", [64, 256, 544, 1024])
@pytest.mark.parametrize("cplx", [False, True])
def test_tree_sum_vs_oracle(capi, oracle, bits, cplx):
    rng = np.random.default_rng(bits * 5 + cplx)
    for count in (0, 1, 2, 3, 511, 512, 513, 1023, 1025, 5000, 262145):
        if count > 6000 and (bits > 256 or cplx):
            continue
        print(f"[gpu] tree sum bits={bits} complex={cplx} count={count}", flush=True)
        x = values(rng, bits, max(count, 1), cplx, span=40)[:count]
        if count > 10:
            x[7] = 0
            x[9::13] = lfcases.negate(x[8::13][: x[9::13].shape[0]], bits) if not cplx else x[9::13]
        X = capi.Buffer(bits, max(count, 1), cplx)
        if count:
            X.upload(x)
        O = capi.Buffer(bits, 1, cplx)
        capi.tree_sum(bits, X, count, O)
        capi.lib().lf_stream_synchronize(None)
        want = oracle.tree_sum(bits, cplx, x)
        assert (O.download() == want).all(), f"tree sum of {count} values differs from the oracle"
