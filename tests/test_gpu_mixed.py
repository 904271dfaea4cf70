"""Mixed-precision chain batches (lf_mixed_chains, lf_mixed_graph_create, lf_autotune_mixed): every value of
every precision bit-identical to the oracle's step-by-step fma/mul sequence, for the per-precision launches,
explicit workgroup sizes, autotuned sizes and a replayed hipGraph whose inputs change between replays."""
import numpy as np
import pytest

from lfcases import random_records, zero

pytestmark = pytest.mark.gpu


def make_batch(rng, widths, per_width):
    batch = {}
    for bits in widths:
        k = per_width
        x = random_records(rng, bits, k, 64)
        f = random_records(rng, bits, k, 64)
        g = random_records(rng, bits, k, 64)
        x[::37] = zero(bits)
        f[5::41] = zero(bits)
        g[::29] = zero(bits)
        x[7::53, bits // 32 + 2] = 4        # a status rides along
        f[11::59, bits // 32] = 999_999_990 & 0xFFFFFFFF  # products near the exponent limit
        length = rng.integers(0, 65, size=k).astype(np.uint32)
        length[:3] = [0, 1, 64]
        mask = rng.integers(0, 1 << 63, size=k, dtype=np.uint64) * 2 + rng.integers(0, 2, size=k).astype(np.uint64)
        order = np.argsort(-length.astype(np.int64), kind="stable")  # longest chains first, as the batcher orders them
        batch[bits] = tuple(a[order] for a in (x, f, g, length, mask))
    return batch


def check(capi, oracle, batch, buckets):
    for b in buckets:
        x, f, g, length, mask = batch[b.bits]
        want = oracle.mixed_chain(b.bits, x, f, g, length, mask)
        got = b.O.download()
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"bits={b.bits}: {bad.size} values differ, first {bad[0]}"


WIDTHS = [64, 96, 128, 160, 256, 384, 512, 800, 1024]


def test_per_precision_launches(capi, oracle):
    rng = np.random.default_rng(2)
    batch = make_batch(rng, WIDTHS, 700)
    buckets = [capi.MixedBucket(bits, *batch[bits]) for bits in WIDTHS]
    print("[gpu] mixed chains, per-precision launches", flush=True)
    capi.mixed_chains(buckets)
    capi.lib().lf_stream_synchronize(None)
    check(capi, oracle, batch, buckets)
    for blk in (64, 128):
        print(f"[gpu] mixed chains, workgroup {blk}", flush=True)
        capi.mixed_chains(buckets, block=[blk] * len(buckets))
        capi.lib().lf_stream_synchronize(None)
        check(capi, oracle, batch, buckets)


def test_autotune_then_default_shape(capi, oracle):
    rng = np.random.default_rng(3)
    for bits in (64, 512, 1024):
        blk, ms = capi.autotune_mixed(bits, 4096)
        assert blk in (64, 128, 256) and ms > 0
    batch = make_batch(rng, [64, 512, 1024], 300)
    buckets = [capi.MixedBucket(bits, *batch[bits]) for bits in (64, 512, 1024)]
    capi.mixed_chains(buckets)
    capi.lib().lf_stream_synchronize(None)
    check(capi, oracle, batch, buckets)


def test_graph_replays_with_new_inputs(capi, oracle):
    rng = np.random.default_rng(4)
    widths = [64, 256, 1024]
    batch = make_batch(rng, widths, 500)
    buckets = [capi.MixedBucket(bits, *batch[bits]) for bits in widths]
    graph = capi.mixed_graph(buckets)
    stream = capi.Stream()
    other = capi.Buffer(256, 1024)
    other.upload(random_records(rng, 256, 1024))
    try:
        for replay in range(3):
            print(f"[gpu] mixed graph replay {replay}", flush=True)
            capi.graph_launch(graph, stream.handle)
            stream.synchronize()
            check(capi, oracle, batch, buckets)
            # eager work between replays, then new inputs in the same buffers
            capi.elementwise(256, "mul", other, other, None, other, stream=stream.handle)
            stream.synchronize()
            for b in buckets:
                x = random_records(rng, b.bits, b.count, 64)
                b.X.upload(x)
                batch[b.bits] = (x,) + batch[b.bits][1:]
    finally:
        capi.graph_destroy(graph)


def test_graph_lanes_exact_products(capi, oracle):
    """The 17-24 and 25-32-limb groups can split each value's products over 2 or 4 lanes: exact products with
    ties (factor 1.5 on odd mantissas, the short product's undecided case), factors 1 + 2^-(32n-1), addends
    that cancel the product exactly or lie far above or below it, with value counts that leave partial
    wavefronts; every lanes choice gives the oracle's values, as do the per-precision (one-lane) launches."""
    rng = np.random.default_rng(6)
    widths = [576, 736, 800, 832, 960, 1024]
    batch = {}
    for bits, k in zip(widths, (45, 200, 61, 130, 257, 1000)):
        n = bits // 32
        x = random_records(rng, bits, k, 64)
        f = random_records(rng, bits, k, 64)
        g = random_records(rng, bits, k, 64)
        x[:, 0] |= 1
        f[::3, :n] = 0
        f[::3, n - 1] = 0xC0000000  # 1.5 * 2^(e-1)
        f[1::7, :n] = 0
        f[1::7, n - 1] = 0x80000000
        f[1::7, 0] = 1
        g[2::5] = x[2::5]  # an addend equal to the start value: x f + x
        g[2::5, n + 1] = (-g[2::5, n + 1].view(np.int32)).view(np.uint32)
        g[3::11, n] = (g[3::11, n].view(np.int32) + 3000).view(np.uint32)  # far above
        g[4::13, n] = (g[4::13, n].view(np.int32) - 3000).view(np.uint32)  # far below
        x[::37] = zero(bits)
        length = rng.integers(0, 65, size=k).astype(np.uint32)
        length[:2] = [64, 1]
        mask = rng.integers(0, 1 << 63, size=k, dtype=np.uint64) * 2 + rng.integers(0, 2, size=k).astype(np.uint64)
        order = np.argsort(-length.astype(np.int64), kind="stable")
        batch[bits] = tuple(a[order] for a in (x, f, g, length, mask))
    buckets = [capi.MixedBucket(bits, *batch[bits]) for bits in widths]
    stream = capi.Stream()
    for lanes in ([1, 1, 1, 1], [1, 1, 2, 4], [0, 0, 4, 2]):
        print(f"[gpu] mixed graph, lanes per group {lanes}, exact products", flush=True)
        graph = capi.mixed_graph(buckets, lanes=lanes)
        try:
            capi.graph_launch(graph, stream.handle)
            stream.synchronize()
        finally:
            capi.graph_destroy(graph)
        check(capi, oracle, batch, buckets)
    print("[gpu] the same batch, per-precision launches", flush=True)
    capi.mixed_chains(buckets)
    capi.lib().lf_stream_synchronize(None)
    check(capi, oracle, batch, buckets)
    print("[gpu] lanes tuner on the batch", flush=True)
    lanes, ms = capi.autotune_mixed_lanes(buckets)
    assert lanes[0] == 1 and lanes[1] == 1 and lanes[2] in (1, 2, 4) and lanes[3] in (1, 2, 4) and ms > 0
    check(capi, oracle, batch, buckets)
    for bad in ([2, 1, 1, 1], [1, 4, 1, 1], [1, 1, 3, 1]):
        with pytest.raises(ValueError, match="lanes per value"):
            capi.mixed_graph(buckets, lanes=bad)


def test_invalid_shapes(capi):
    rng = np.random.default_rng(5)
    batch = make_batch(rng, [64], 64)
    b = capi.MixedBucket(64, *batch[64])
    with pytest.raises(ValueError):
        capi.mixed_chains([b], block=[100])
    with pytest.raises(ValueError):
        capi.autotune_mixed(80, 4096)
