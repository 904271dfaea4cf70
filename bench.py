#!/usr/bin/env python3
"""Headline benchmark: the 1024-bit resident multiplication chain on one MI355X.

Workload (one "step"): 65,536 independent 1024-bit (32 x u32 limb) values, each multiplied 16 times by
its own factor and rounded to nearest-even after every multiplication (the reference's resident
`mul_chain` benchmark, benchmarks/benchmark.cpp:67-94).  Operands: random sign, uniform random limbs
with the top bit set, exponent uniform in [-64, 64], seeded.  Inputs are resident in HBM (limb
planes) before the timed region; one step is one lf_mul_chain launch over the whole batch.

Prints ONE JSON line on stdout (diagnostics go to stderr).  Besides the contract fields it carries:
  roofline      the chain kernel against the measured v_mad_u64_u32 rate of the device (VALU-bound),
  cpu_baseline  the reference's own CPU arithmetic (core.hpp, compiled from the reference into
                oracle/_ref/libref.so) on all the host cores it may use, plus MPFR serial/all-cores,
  parity        the GPU output compared bit for bit with the CPU oracle on the same inputs.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

BITS, COUNT, CHAIN = 1024, 65536, 16
# Reference (Apple M5 Max) device time for this exact workload: 3.192 ms per chain batch
# (benchmarks/results/final.csv:145) -> 65,536 / 3.192 ms.
REFERENCE_VALUES_PER_S = COUNT / 3.192e-3
# Measured HBM traffic of one mul_chain_kernel<32> launch (counter passes described at the roofline below).
TRAFFIC_BYTES = round((9081.8 * 262144 / 131084.6 + 8960.0) * 1024)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(seed: int):
    n = BITS // 32
    rng = np.random.default_rng(seed)

    def rec():
        r = np.zeros((COUNT, n + 3), dtype=np.uint32)
        r[:, :n] = rng.integers(0, 1 << 32, size=(COUNT, n), dtype=np.uint64).astype(np.uint32)
        r[:, n - 1] |= np.uint32(0x80000000)
        r[:, n] = (rng.integers(-64, 65, size=COUNT) & 0xFFFFFFFF).astype(np.uint32)
        r[:, n + 1] = (np.where(rng.integers(0, 2, size=COUNT) == 1, 1, -1) & 0xFFFFFFFF).astype(np.uint32)
        return r

    return rec(), rec()


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(x, f, budget_s: float):
    """The reference's CPU core (oracle/_ref/libref.so) -- or the oracle port when that library is
    absent -- on a bounded sample of the same workload, serial and on all allowed host cores; then MPFR."""
    ref = ROOT / "oracle" / "_ref" / "libref.so"
    if ref.exists():
        L = C.CDLL(str(ref))
        fn, kind = L.ref_mul_chain, "reference"
    else:
        from oracle import lf_oracle

        L = lf_oracle.lib()
        fn, kind = L.lfo_mul_chain, "port"
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    fn.restype = C.c_int
    out = np.zeros_like(x)
    # size the serial sample to the budget from a short probe
    probe = 256
    t0 = time.perf_counter()
    fn(BITS, x.ctypes.data, f.ctypes.data, out.ctypes.data, probe, CHAIN)
    per_value = (time.perf_counter() - t0) / probe
    sample = int(min(COUNT, max(1024, budget_s / 2 / max(per_value, 1e-9))))
    t0 = time.perf_counter()
    fn(BITS, x.ctypes.data, f.ctypes.data, out.ctypes.data, sample, CHAIN)
    serial = time.perf_counter() - t0
    threads = host_threads()
    chunks = np.array_split(np.arange(sample), threads)

    def part(ix):
        lo = int(ix[0]) if len(ix) else 0
        k = len(ix)
        if k:
            stride = x.shape[1] * 4
            fn(BITS, x.ctypes.data + lo * stride, f.ctypes.data + lo * stride, out.ctypes.data + lo * stride, k, CHAIN)

    ths = [threading.Thread(target=part, args=(c,)) for c in chunks]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    parallel = time.perf_counter() - t0
    res = {
        "value": sample / parallel,
        "unit": "values/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{sample} of the {COUNT} 1024-bit values, 16-step chain each, oracle/_ref libref.so ref_mul_chain",
        "serial_values_per_s": sample / serial,
    }
    mpfr = ROOT / "oracle" / "mpfr_chain_bench"
    if mpfr.exists():
        try:
            p = subprocess.run([str(mpfr), str(BITS), str(COUNT), str(CHAIN), str(threads), "3"], capture_output=True,
                               text=True, timeout=300)
            m = json.loads(p.stdout)
            res["mpfr"] = {"version": m["mpfr"], "threads": threads, "serial_values_per_s": COUNT / m["serial_s"],
                           "parallel_values_per_s": COUNT / m["parallel_s"], "sample": f"all {COUNT} values, median of 3"}
        except Exception as e:  # the MPFR leg is informative only
            res["mpfr"] = {"error": str(e)}
    return res


def gemm_inputs(bits: int, n: int, seed: int):
    """Config-3 entries: random sign, uniform limbs with the top bit set, exponent c + round(8 z) with c uniform
    in {-200, 0, +200} per entry and z standard normal, seeded."""
    w = bits // 32
    rng = np.random.default_rng(seed)

    def mat():
        r = np.zeros((n * n, w + 3), dtype=np.uint32)
        r[:, :w] = rng.integers(0, 1 << 32, size=(n * n, w), dtype=np.uint64).astype(np.uint32)
        r[:, w - 1] |= np.uint32(0x80000000)
        e = rng.choice([-200, 0, 200], size=n * n) + np.round(8 * rng.standard_normal(n * n)).astype(np.int64)
        r[:, w] = (e & 0xFFFFFFFF).astype(np.uint32)
        r[:, w + 1] = (np.where(rng.integers(0, 2, size=n * n) == 1, 1, -1) & 0xFFFFFFFF).astype(np.uint32)
        return r

    return mat(), mat()


def isa_mad_peak() -> float:
    """The clock-derived issue bound of v_mad_u64_u32 on MI355X: 256 CUs x 64 lanes per clock x 2.4 GHz peak
    engine clock (one 32x32->64 multiply-add per lane per clock)."""
    return 256 * 64 * 2.4e9


MFMA_INT8_PEAK = 5.0e15  # MI355X dense int8 MFMA, ops/s (AMD's published figure without sparsity)


def stage_report(rep) -> dict:
    return {"analysis_ms": round(rep.analysis_seconds * 1e3, 3), "upload_ms": round(rep.upload_seconds * 1e3, 3),
            "device_ms": round(rep.gpu_seconds * 1e3, 3), "assembly_ms": round(rep.assembly_seconds * 1e3, 3),
            "digits_ms": round(rep.digits_seconds * 1e3, 3), "product_ms": round(rep.product_seconds * 1e3, 3),
            "rebuild_ms": round(rep.rebuild_seconds * 1e3, 3), "dot_ms": round(rep.dot_seconds * 1e3, 3)}


def mpfr_gemm_baseline(bits, A, B, m, n, k, budget_s, got=None):
    """oracle/mpfr_gemm_bench on the same operands: exact MPFR dots (2*bits products + one mpfr_sum) and the
    reference benchmark's rounded consumer loop, serial and on all allowed cores, on bounded row samples;
    the exact rows are compared with the GPU's output `got`."""
    import tempfile

    exe = ROOT / "oracle" / "mpfr_gemm_bench"
    if not exe.exists():
        return {"error": "oracle/mpfr_gemm_bench is not built"}
    threads = host_threads()
    with tempfile.TemporaryDirectory() as d:
        src, dst = Path(d) / "in.bin", Path(d) / "out.bin"
        src.write_bytes(np.concatenate([np.array([bits, m, n, k], np.uint32), A.ravel(), B.ravel()]).astype(np.uint32).tobytes())
        p = subprocess.run([str(exe), str(src), str(threads), str(budget_s), str(dst)], capture_output=True, text=True, timeout=600)
        r = json.loads(p.stdout)
        out = {"threads": threads, "mpfr": r["mpfr"],
               "exact_serial_ms": round(r["exact_serial_s_per_output"] * m * n * 1e3, 1),
               "exact_parallel_ms": round(r["exact_parallel_s_per_output"] * m * n * 1e3, 1),
               "consumer_serial_ms": round(r["consumer_serial_s_per_output"] * m * n * 1e3, 1),
               "consumer_parallel_ms": round(r["consumer_parallel_s_per_output"] * m * n * 1e3, 1),
               "sample": (f"exact: {r['exact_serial_rows']} rows serial, {r['exact_parallel_rows']} rows on {threads} threads; "
                          f"consumer: {r['consumer_serial_rows']} / {r['consumer_parallel_rows']} rows; of {m} rows of {n} outputs, "
                          "scaled to the whole product")}
        if got is not None:
            buf = np.frombuffer(dst.read_bytes(), dtype=np.uint32)
            rows, rw = int(buf[0]), bits // 32 + 3
            bad, at = 0, 1
            for _ in range(rows):
                i = int(buf[at])
                want = buf[at + 1:at + 1 + n * rw].reshape(n, rw)
                bad += int((got[i * n:(i + 1) * n] != want).any(axis=1).sum())
                at += 1 + n * rw
            out["mpfr_vs_gpu"] = {"outputs": rows * n, "mismatches": bad}
    return out


def gemm_section(reps: int, oracle_samples: int, cpu_budget: float, peak_tps: float):
    """512-bit exact GEMM C = A B at M = N = K = 512 (three exponent bands per line): the int8-MFMA residue
    GEMM (default options and the one-band lever band_bits = 512) against the plain on-GPU limb GEMM; all
    three bit-identical, a sample of outputs checked against the oracle, the product kernel against the
    int8 MFMA peak, and MPFR exact dots on the host as the CPU baseline."""
    from limbforge_amd import capi

    bits, n = 512, 512
    A, B = gemm_inputs(bits, n, 20261007 + bits)
    res = {}
    outs = {}
    reports = {}
    for name, kw in (("residue", {}), ("residue_band512", {"band_bits": 512}), ("limb", {"method": 1})):
        o = capi.LinalgOptions.default(**kw)
        times, walls, stages = [], [], []
        r = rep = None
        for i in range(reps if name != "limb" else max(2, reps // 3)):
            r, rep, g, w = capi.gemm(bits, False, A, B, n, n, n, options=o)
            if i:  # first call warms
                times.append(g)
                walls.append(w)
                stages.append(stage_report(rep))
        outs[name] = r
        reports[name] = rep
        med = stages[len(stages) // 2]
        res[name] = {"device_ms": round(1e3 * float(np.median(times)), 4), "wall_ms": round(1e3 * float(np.median(walls)), 3),
                     "moduli": int(rep.moduli), "int8_gemms": int(rep.int8_gemms), "gemm_rows": int(rep.gemm_rows),
                     "band_pairs": int(rep.npairs), "fallback_outputs": int(rep.fallback_outputs), "stages": med}
        log(f"[bench] gemm {name}: {res[name]}")
    same = all((outs[k] == outs["residue"]).all() for k in outs)
    check = None
    if oracle_samples:
        from oracle import lf_oracle

        rng = np.random.default_rng(1)
        rows = rng.choice(n, size=max(1, oracle_samples // n + 1), replace=False)
        want = lf_oracle.gemm_rows(bits, A, B, n, n, n, rows=rows)
        got = np.concatenate([outs["residue"][i * n:(i + 1) * n] for i in rows])
        check = {"rows": len(rows), "outputs": len(rows) * n, "mismatches": int((got != want).any(axis=1).sum())}
    # roofline of the dominant kernel, the int8-MFMA residue product: 4 int8 products (d1 d1, d1 d0, d0 d1,
    # d0 d0) of 2 ops each per modulus, extended row, extended column and K step, as launched (padded K)
    rr = reports["residue"]
    kp = max(64, (n + 63) // 64 * 64)
    ops = 8 * rr.moduli * rr.gemm_rows * rr.gemm_cols * kp
    prod_s = res["residue"]["stages"]["product_ms"] * 1e-3
    roof = {"bound": "mfma", "achieved": round(ops / prod_s / 1e12, 2), "peak": MFMA_INT8_PEAK / 1e12, "unit": "Tops/s (int8)",
            "frac": round(ops / prod_s / MFMA_INT8_PEAK, 4), "traffic": None, "kernel": "product_kernel (v_mfma_i32_32x32x32_i8)",
            "kernel_ms": res["residue"]["stages"]["product_ms"], "int8_ops": int(ops)}
    limb_wp = n * n * n * (bits // 32) ** 2
    limb_s = res["limb"]["device_ms"] * 1e-3
    limb_roof = {"bound": "valu", "achieved": round(limb_wp / limb_s / 1e12, 3), "unit": "Tword-products/s",
                 "frac_of_measured_peak": round(limb_wp / limb_s / peak_tps, 4), "frac_of_isa_peak": round(limb_wp / limb_s / isa_mad_peak(), 4),
                 "word_products": limb_wp}
    base = None
    if cpu_budget > 0:
        log("[bench] gemm: MPFR CPU baseline")
        base = mpfr_gemm_baseline(bits, A, B, n, n, n, cpu_budget, outs["residue"])
        base.update({"value": base.get("exact_parallel_ms"), "unit": "ms", "cores": base.get("threads"), "kind": "port"})
    return {
        "workload": "512-bit exact GEMM C = A B, M = N = K = 512, exponents in bands -200/0/+200 (sigma 8) (config 3)",
        "residue_ms": res["residue"]["device_ms"],
        "limb_ms": res["limb"]["device_ms"],
        "speedup_residue_vs_limb": round(res["limb"]["device_ms"] / res["residue"]["device_ms"], 2),
        "speedup_band512_vs_limb": round(res["limb"]["device_ms"] / res["residue_band512"]["device_ms"], 2),
        "wall_speedup_residue_vs_limb": round(res["limb"]["wall_ms"] / res["residue"]["wall_ms"], 2),
        "bit_identical_all_paths": bool(same),
        "oracle_check": check,
        "roofline": roof,
        "limb_gemm_roofline": limb_roof,
        "cpu_baseline": base,
        "paths": res,
        "timing": "device time of the product kernels (HIP events), median; wall includes host transfers and band analysis",
    }


def elementwise_section(reps: int, peak_tps: float, oracle_check: bool, cpu: bool):
    """Element-wise batches of 2^20 values at 256, 512 and 1024 bits: div, fma (coefficient broadcast with
    period 64), sqrt of |x| and the composed complex division, resident; operands as the chain's with
    exponents uniform in [-512, 512].  Device time per launch (HIP events, median of `reps`), the algorithmic
    word products and HBM bytes per element, a sample of outputs against the oracle, and the MPFR baseline
    on a 65,536-element sample (oracle/mpfr_elementwise_bench)."""
    from limbforge_amd import capi

    count = 1 << 20
    rng = np.random.default_rng(20261007 + 1)
    stream = capi.Stream()
    out = {}
    threads = host_threads()
    for bits in (256, 512, 1024):
        n = bits // 32
        E = 4 * n + 12

        def records(k, positive=False):
            r = np.zeros((k, n + 3), dtype=np.uint32)
            r[:, :n] = rng.integers(0, 1 << 32, size=(k, n), dtype=np.uint64).astype(np.uint32)
            r[:, n - 1] |= np.uint32(0x80000000)
            r[:, n] = (rng.integers(-512, 513, size=k) & 0xFFFFFFFF).astype(np.uint32)
            sg = np.ones(k, np.int64) if positive else np.where(rng.integers(0, 2, size=k) == 1, 1, -1)
            r[:, n + 1] = (sg & 0xFFFFFFFF).astype(np.uint32)
            return r

        for op in ("div", "fma", "sqrt", "cdiv"):
            cplx = op == "cdiv"
            if cplx:
                a = np.concatenate([records(count), records(count)], axis=1)
                b = np.concatenate([records(count), records(count)], axis=1)
            else:
                a = records(count, positive=op == "sqrt")
                b = records(64 if op == "fma" else count)
            c = records(count) if op == "fma" else None
            A, O = capi.Buffer(bits, count, cplx), capi.Buffer(bits, count, cplx)
            A.upload(a)
            B = C_ = None
            if op != "sqrt":
                B = capi.Buffer(bits, b.shape[0], cplx)
                B.upload(b)
            if c is not None:
                C_ = capi.Buffer(bits, count)
                C_.upload(c)
            code = {"div": "div", "fma": "fma", "sqrt": "sqrt", "cdiv": "complex_div"}[op]
            bidx = (1, 64) if op == "fma" else None
            times = []
            for i in range(reps + 2):
                e0, e1 = capi.Event(), capi.Event()
                e0.record(stream)
                capi.elementwise(bits, code, A, B, C_, O, b_index=bidx, stream=stream.handle)
                e1.record(stream)
                stream.synchronize()
                if i >= 2:
                    times.append(e0.elapsed_ms(e1))
            ms = float(np.median(times))
            wp = {"div": n * (n + 1), "fma": n * n, "sqrt": n * n, "cdiv": 6 * n * n + 2 * n * (n + 1)}[op]
            nbytes = {"div": 3 * E, "fma": 3 * E, "sqrt": 2 * E, "cdiv": 6 * E}[op]
            rate = count / (ms * 1e-3)
            row = {"device_ms": round(ms, 4), "values_per_s": round(rate, 1),
                   "word_products_per_value": wp, "valu_frac": round(rate * wp / peak_tps, 4),
                   "hbm_bytes_per_value": nbytes, "hbm_gbps": round(rate * nbytes / 1e9, 1),
                   "hbm_frac": round(rate * nbytes / 8e12, 4)}
            if oracle_check:
                from oracle import lf_oracle

                got = O.download()
                k = 4096
                want = lf_oracle.elementwise(bits, capi.OPS[code], a[:k], b if op == "fma" else (b[:k] if B is not None else None),
                                             c[:k] if c is not None else None, b_index=(1, 64) if op == "fma" else (1, 0))
                row["oracle_mismatches"] = int((got[:k] != want).any(axis=1).sum())
                row["oracle_checked"] = k
            if cpu:
                exe = ROOT / "oracle" / "mpfr_elementwise_bench"
                if exe.exists():
                    try:
                        p = subprocess.run([str(exe), str(bits), op, "65536", str(threads), "3"], capture_output=True, text=True, timeout=300)
                        m = json.loads(p.stdout)
                        row["mpfr"] = {"threads": threads, "serial_values_per_s": round(65536 / m["serial_s"], 1),
                                       "parallel_values_per_s": round(65536 / m["parallel_s"], 1), "sample": "65,536 values, median of 3"}
                    except Exception as e:
                        row["mpfr"] = {"error": str(e)}
            out[f"{op}_{bits}"] = row
            log(f"[bench] elementwise {op} {bits}: {row}")
            del A, O, B, C_
    return {"workload": "element-wise div / fma (b period 64) / sqrt |x| / complex div, 2^20 values at 256, 512, 1024 bits (config 1)",
            "peak_word_products_per_s": peak_tps, "ops": out}


def mixed_section(peak_tps: float, oracle_check: bool, cpu: bool, reps: int = 3):
    """10^6 values with precisions 64 + 32 k bits, k in [0, 30] drawn with P(k) ~ 0.85^k, chains of uniform
    length 1-64 whose steps are fma or mul by a seeded mask, operands as the chain's.  Bucketed by precision
    (longest chains first inside a bucket); timed as one launch per precision on a stream (workgroup 256,
    then the autotuned sizes) and as one hipGraph of concurrent bucket nodes with the autotuned sizes and
    lanes per value."""
    from limbforge_amd import capi

    count, p = 10 ** 6, 0.15
    rng = np.random.default_rng(20261007 + 2)
    probs = (1 - p) ** np.arange(31)
    probs /= probs.sum()
    ks = rng.choice(31, size=count, p=probs)
    length = rng.integers(1, 65, size=count).astype(np.uint32)
    mask = rng.integers(0, 1 << 63, size=count, dtype=np.uint64) * 2 + rng.integers(0, 2, size=count).astype(np.uint64)
    buckets, host, work = [], {}, 0
    for k in np.unique(ks):
        bits = 64 + 32 * int(k)
        n = bits // 32
        idx = np.nonzero(ks == k)[0]
        idx = idx[np.argsort(-length[idx].astype(np.int64), kind="stable")]

        def records(m):
            r = np.zeros((m, n + 3), dtype=np.uint32)
            r[:, :n] = rng.integers(0, 1 << 32, size=(m, n), dtype=np.uint64).astype(np.uint32)
            r[:, n - 1] |= np.uint32(0x80000000)
            r[:, n] = (rng.integers(-64, 65, size=m) & 0xFFFFFFFF).astype(np.uint32)
            r[:, n + 1] = (np.where(rng.integers(0, 2, size=m) == 1, 1, -1) & 0xFFFFFFFF).astype(np.uint32)
            return r

        x, f, g = records(idx.size), records(idx.size), records(idx.size)
        host[bits] = (x, f, g, length[idx], mask[idx])
        buckets.append(capi.MixedBucket(bits, *host[bits]))
        work += int(length[idx].astype(np.int64).sum()) * n * n
    tuned = {}
    for b in buckets:
        tuned[b.bits] = capi.autotune_mixed(b.bits, 32768)[0]
    stream = capi.Stream()

    def timed(run):
        ts = []
        for i in range(reps + 1):
            e0, e1 = capi.Event(), capi.Event()
            e0.record(stream)
            run()
            e1.record(stream)
            stream.synchronize()
            if i:
                ts.append(e0.elapsed_ms(e1))
        return float(np.median(ts))

    plain_ms = timed(lambda: capi.mixed_chains(buckets, block=[256] * len(buckets), stream=stream.handle))
    tuned_ms = timed(lambda: capi.mixed_chains(buckets, block=[tuned[b.bits] for b in buckets], stream=stream.handle))
    # lanes per value of the two widest groups, tuned on this batch's graph (1 lane keeps the narrower groups)
    lanes, _ = capi.autotune_mixed_lanes(buckets, block=[tuned[b.bits] for b in buckets])
    graph = capi.mixed_graph(buckets, block=[tuned[b.bits] for b in buckets], lanes=lanes)
    try:
        graph_ms = timed(lambda: capi.graph_launch(graph, stream.handle))
    finally:
        capi.graph_destroy(graph)
    res = {
        "workload": "10^6 values, 64-1024 bits (P(k) ~ 0.85^k), fma/mul chains of length 1-64 (config 2)",
        "buckets": len(buckets),
        "word_products": work,
        "per_precision_launches_ms": round(plain_ms, 4),
        "per_precision_tuned_ms": round(tuned_ms, 4),
        "graph_tuned_ms": round(graph_ms, 4),
        "values_per_s_graph": round(count / (graph_ms * 1e-3), 1),
        "speedup_graph_vs_per_precision": round(plain_ms / graph_ms, 3),
        "valu_frac_graph": round(work / (graph_ms * 1e-3) / peak_tps, 4),
        "autotuned_workgroup": {str(k): v for k, v in sorted(tuned.items())},
        "autotuned_lanes_per_group": lanes,
    }
    res["roofline"] = {"bound": "valu", "achieved": round(work / (graph_ms * 1e-3) / 1e12, 3), "peak": round(peak_tps / 1e12, 3),
                       "unit": "Tword-products/s", "frac": round(work / (graph_ms * 1e-3) / peak_tps, 4),
                       "frac_of_isa_peak": round(work / (graph_ms * 1e-3) / isa_mad_peak(), 4), "traffic": None,
                       "kernel": "mixed_group_kernel<0..3> (one graph, four concurrent group nodes)"}
    if oracle_check or cpu:
        import tempfile

        from oracle import lf_oracle

        # a sample spread over every bucket: the oracle's step sequence checks the GPU, and MPFR (mpfr_fma /
        # mpfr_mul at each value's precision) is the CPU baseline on the same sample, serial and on all cores
        bad = checked = 0
        sample_work = 0
        parts = [np.array([len(buckets)], np.uint32)]
        order = []
        for b in buckets:
            x, f, g, ln, mk = host[b.bits]
            sel = np.linspace(0, b.count - 1, num=min(b.count, 300)).astype(np.int64)
            want = lf_oracle.mixed_chain(b.bits, x[sel], f[sel], g[sel], ln[sel], mk[sel])
            got = b.O.download()[sel]
            bad += int((got != want).any(axis=1).sum())
            checked += sel.size
            sample_work += int(ln[sel].astype(np.int64).sum()) * (b.bits // 32) ** 2
            parts += [np.array([b.bits, sel.size], np.uint32), x[sel].ravel(), f[sel].ravel(), g[sel].ravel(),
                      ln[sel].astype(np.uint32), mk[sel].astype(np.uint64).view(np.uint32)]
            order.append((b.bits, got))
        res["oracle_check"] = {"values": checked, "mismatches": bad}
        exe = ROOT / "oracle" / "mpfr_mixed_bench"
        if cpu and exe.exists():
            threads = host_threads()
            with tempfile.TemporaryDirectory() as d:
                src, dst = Path(d) / "in.bin", Path(d) / "out.bin"
                src.write_bytes(np.concatenate([q.astype(np.uint32).ravel() for q in parts]).tobytes())
                p = subprocess.run([str(exe), str(src), str(threads), "3", str(dst)], capture_output=True, text=True, timeout=600)
                m = json.loads(p.stdout)
                outw = np.frombuffer(dst.read_bytes(), dtype=np.uint32)
                at, mism = 0, 0
                for bits, got in order:
                    rw = bits // 32 + 3
                    want = outw[at:at + got.shape[0] * rw].reshape(got.shape[0], rw)
                    at += got.shape[0] * rw
                    mism += int((got != want).any(axis=1).sum())
            ser = sample_work / m["serial_s"]
            par = sample_work / m["parallel_s"]
            res["cpu_baseline"] = {"value": round(count / (work / par), 1), "unit": "values/s at the batch's precision mix",
                                   "cores": threads, "kind": "port", "mpfr": m["mpfr"],
                                   "serial_values_per_s": round(count / (work / ser), 1),
                                   "word_products_per_s": {"serial": round(ser, 1), "parallel": round(par, 1)},
                                   "sample": f"{checked} values spread over every bucket (mpfr_fma / mpfr_mul per step at each value's "
                                             "precision, median of 3), scaled to the batch by its word products",
                                   "mpfr_vs_gpu_mismatches": mism}
    log(f"[bench] mixed: {res}")
    return res


def to_float(rec: np.ndarray, bits: int) -> float:
    """A record's value as a double (top two limbs; exponent clamped to double range)."""
    w = bits // 32
    s = int(rec[w + 1]) - ((int(rec[w + 1]) >> 31) << 32)
    if not s or rec[w + 2]:
        return 0.0
    e = int(rec[w]) - ((int(rec[w]) >> 31) << 32)
    top = (int(rec[w - 1]) << 32) | int(rec[w - 2])
    return float(s) * float(np.ldexp(float(top), max(min(e - 63, 1000), -1000)))


def cholesky_section(oracle_check: bool):
    """256-bit SYRK + blocked Cholesky + solve: A (2048 x 512) single band (exponent round(8 z)), S = A^T A
    then + I on the diagonal (two roundings, the reference's syrk-then-add idiom), Cholesky n = 512 with
    block 32, cholesky_solve with 64 right-hand sides (exponent span 20); the scaled residual
    max |b - S x| / (|S| |x| + |b|) with r rounded once by an exact update GEMM.  L and X are compared
    bit for bit with the oracle's restatement of the reference's fixed-order host path."""
    from limbforge_amd import capi

    bits, rows, n, nrhs, block = 256, 2048, 512, 64, 32
    w = bits // 32
    rng = np.random.default_rng(20261007 + 4)

    def records(count, exps):
        r = np.zeros((count, w + 3), dtype=np.uint32)
        r[:, :w] = rng.integers(0, 1 << 32, size=(count, w), dtype=np.uint64).astype(np.uint32)
        r[:, w - 1] |= np.uint32(0x80000000)
        r[:, w] = (np.asarray(exps, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32)
        r[:, w + 1] = (np.where(rng.integers(0, 2, size=count) == 1, 1, -1) & 0xFFFFFFFF).astype(np.uint32)
        return r

    A = records(rows * n, np.round(8 * rng.standard_normal(rows * n)))
    B = records(n * nrhs, rng.integers(-10, 11, size=n * nrhs))
    # one untimed call (the first launch of each kernel loads its code), then the median of three
    capi.syrk(bits, A, rows, n, lower_only=False)
    syrks = sorted((capi.syrk(bits, A, rows, n, lower_only=False) for _ in range(3)), key=lambda x: x[2])
    S, srep, syrk_g, syrk_w = syrks[1]
    S0 = S.copy()
    one = np.zeros((n, w + 3), np.uint32)
    one[:, w - 1] = 0x80000000
    one[:, w + 1] = 1
    diag = np.arange(n) * (n + 1)
    S[diag], _, _ = capi.run_host(bits, "add", S[diag], one)
    res = {"syrk": {"device_ms": round(syrk_g * 1e3, 3), "wall_ms": round(syrk_w * 1e3, 3), "moduli": int(srep.moduli),
                    "band_pairs": int(srep.npairs), "stages": stage_report(srep)}}
    chol = []
    for _ in range(3):
        t = time.perf_counter()
        L, info = capi.cholesky(bits, S, n, block=block)
        chol.append((time.perf_counter() - t, info.timing_gpu_seconds, info.panel_seconds, info.update_seconds))
    chol.sort()
    res["cholesky"] = {"wall_ms": round(chol[1][0] * 1e3, 3), "gpu_ms": round(chol[1][1] * 1e3, 3),
                       "panel_ms": round(chol[1][2] * 1e3, 3), "update_ms": round(chol[1][3] * 1e3, 3),
                       "pivot": int(info.pivot), "blocks": int(info.blocks)}
    sol = []
    for _ in range(3):
        t = time.perf_counter()
        X, g, _ = capi.solve(bits, 3, L, n, B, nrhs, block=block)
        sol.append((time.perf_counter() - t, g))
    sol.sort()
    res["solve"] = {"wall_ms": round(sol[1][0] * 1e3, 3), "gpu_ms": round(sol[1][1] * 1e3, 3), "nrhs": nrhs}
    # scaled residual: r = RN(b - S x) (one rounding), scale = RN(RN(|S| |x|) + |b|), ratio = RN(|r| / scale)
    r, _, _, _ = capi.gemm(bits, False, S, X, n, nrhs, n, C_in=B, subtract=True)

    def absval(v):
        v = v.copy()
        v[:, w + 1] = np.where(v[:, w + 1] != 0, 1, 0).astype(np.uint32)
        return v

    t, _, _, _ = capi.gemm(bits, False, absval(S), absval(X), n, nrhs, n)
    scale, _, _ = capi.run_host(bits, "add", t, absval(B))
    ratio, _, _ = capi.run_host(bits, "div", absval(r), scale)
    res["scaled_residual"] = max(abs(to_float(q, bits)) for q in ratio)
    log(f"[bench] cholesky: {res}")
    # roofline of the SYRK's int8-MFMA residue product (lower-triangle tiles only)
    kp = (rows + 63) // 64 * 64
    tiles = (srep.gemm_rows + 127) // 128
    ops = 8 * srep.moduli * 128 * 128 * (tiles * (tiles + 1) // 2) * kp
    res["roofline"] = {"bound": "mfma", "achieved": round(ops / srep.product_seconds / 1e12, 2), "peak": MFMA_INT8_PEAK / 1e12,
                       "unit": "Tops/s (int8)", "frac": round(ops / srep.product_seconds / MFMA_INT8_PEAK, 4), "traffic": None,
                       "kernel": "product_kernel (SYRK, lower-triangle 128 x 128 tiles)",
                       "kernel_ms": round(srep.product_seconds * 1e3, 3), "int8_ops": int(ops),
                       "note": "the factorization itself is latency-bound (single-wave exact dots, square roots and divisions per column)"}
    if oracle_check:
        from oracle import lf_oracle

        # the oracle's own S = A^T A (every entry, on all host threads) + I, then the oracle's Cholesky and
        # solves of it in the reference's fixed host order, serially and on all host threads
        threads = host_threads()
        At = np.ascontiguousarray(A.reshape(rows, n, w + 3).transpose(1, 0, 2)).reshape(n * rows, w + 3)
        t = time.perf_counter()
        WS = lf_oracle.gemm_rows(bits, At, A, n, n, rows)
        t_syrk = time.perf_counter() - t
        syrk_bad = int((WS != S0).any(axis=1).sum())
        WS[diag] = lf_oracle.elementwise(bits, 0, WS[diag], one)
        times = {}
        for th in (1, threads):
            lf_oracle.set_threads(th)
            t = time.perf_counter()
            WL, piv, _, _ = lf_oracle.cholesky(bits, WS, n, block)
            t_chol = time.perf_counter() - t
            t = time.perf_counter()
            WX = lf_oracle.solve(bits, WL, n, B, nrhs, block, 3)
            times[th] = (t_chol, time.perf_counter() - t)
        lf_oracle.set_threads(1)
        res["oracle_check"] = {"syrk_mismatches": syrk_bad, "factor_mismatches": int((WL != L).any(axis=1).sum()),
                               "solve_mismatches": int((WX != X).any(axis=1).sum()), "pivot": int(piv),
                               "from": "the oracle's own S (SYRK of every entry, then + I)"}
        res["cpu_baseline"] = {"value": round((times[threads][0] + times[threads][1]) * 1e3, 1), "unit": "ms (Cholesky + solve)",
                               "cores": threads, "kind": "port",
                               "cholesky_s": {"serial": round(times[1][0], 3), "parallel": round(times[threads][0], 3)},
                               "solve_s": {"serial": round(times[1][1], 3), "parallel": round(times[threads][1], 3)},
                               "syrk_s_parallel": round(t_syrk, 3),
                               "sample": "the whole workload: the oracle's restatement of the reference's fixed-order host path "
                                         "(panel rows, trailing-update rows and right-hand sides on the host threads)"}
        log(f"[bench] cholesky oracle check: {res['oracle_check']}, {res['cpu_baseline']}")
    # the per-row steps the factorization and the solves chain, timed on one wavefront (lf_factor_step_cycles)
    res["step_cycles_one_wave"] = capi.factor_step_cycles(bits, 32)
    log(f"[bench] cholesky per-row step cycles: {res['step_cycles_one_wave']}")
    res["workload"] = ("256-bit SYRK (A 2048 x 512, single band) + blocked Cholesky n = 512 (block 32) + cholesky_solve "
                       "with 64 right-hand sides (config 4)")
    return res


# The reference's published times for its own linalg workloads (Apple M5 Max GPU, milliseconds):
# benchmarks/results/round20_linalg.csv:16-17 (256-bit QSC-like SYRK and GEMM A^T B, n = 1000, K = 2000) and
# round23_cholesky.csv:7,9 (normal equations J^T J + damping, Cholesky block 32 and cholesky_solve, 4 RHS).
REFERENCE_LINALG = {
    "syrk_1000": {"device_ms": 18.501, "wall_ms": 22.177},
    "gemm_1000": {"device_ms": 31.945, "wall_ms": 38.345},
    "cholesky_1000": {"syrk_wall_ms": 31.064, "factor_wall_ms": 163.2, "solve_wall_ms": 73.788, "e2e_wall_ms": 268.09},
    "cholesky_400": {"syrk_wall_ms": 5.8664, "factor_wall_ms": 37.132, "solve_wall_ms": 8.8681, "e2e_wall_ms": 51.386},
}


def qsc_jacobian(rng, bits: int, rows: int, cols: int) -> np.ndarray:
    """QSC-like Jacobian columns as the reference's linalg and Cholesky benchmarks draw them
    (benchmarks/linalg.mm:39-46): column scale in [-60, 60], entries random_number(15) (tests/reference.hpp:21-26)
    shifted by scale - 10, a quarter of them zero, and in every 17th column (c % 17 == 5) one entry 150 below
    the column scale.  Row-major rows x cols records; numpy's generator stands in for mt19937_64."""
    w = bits // 32
    r = np.zeros((rows, cols, w + 3), dtype=np.uint32)
    r[:, :, :w] = rng.integers(0, 1 << 32, size=(rows, cols, w), dtype=np.uint64).astype(np.uint32)
    r[:, :, w - 1] |= np.uint32(0x80000000)
    scale = rng.integers(-60, 61, size=cols)
    e = rng.integers(-15, 16, size=(rows, cols)) + scale[None, :] - 10
    odd = np.nonzero(np.arange(cols) % 17 == 5)[0]
    e[rng.integers(0, rows, size=odd.size), odd] = scale[odd] - 150
    r[:, :, w] = (e & 0xFFFFFFFF).astype(np.uint32)
    r[:, :, w + 1] = (np.where(rng.integers(0, 2, size=(rows, cols)) == 1, 1, -1) & 0xFFFFFFFF).astype(np.uint32)
    r[rng.integers(0, 4, size=(rows, cols)) == 0] = 0
    return r.reshape(rows * cols, w + 3)


# The reference's published element-wise times for 65,536 values (Apple M5 Max; milliseconds: GPU device
# and wall of Engine::run on host arrays), benchmarks/results/final.csv:40,46,134,136,142 and
# round12_reduction_policy.csv:52,172.
REFERENCE_ELEMENTWISE = {
    ("mul", 1024): (0.482, 1.622), ("div", 1024): (0.787, 2.088), ("complex_div", 1024): (6.04, 7.96),
    ("div", 256): (0.278, 0.681), ("complex_div", 256): (0.767, 1.705), ("sqrt", 256): (1.50, 2.13), ("sqrt", 1024): (5.47, 6.74),
}


def reference_elementwise_section():
    """The reference's element-wise benchmark (benchmarks/benchmark.cpp:3-60): 65,536 values with
    random_number(5) operands (exponents in [-5, 5]), sqrt of |a|, Engine::run on host arrays; device time
    (the kernel, HIP events) and wall time (host arrays in and out) of lf_run_host, median of 5 after two
    warm-ups, beside the reference's published M5 Max times."""
    from limbforge_amd import capi

    count = 65536
    rng = np.random.default_rng(20261007 + 5)
    out = {}
    for (op, bits), (ref_dev, ref_wall) in REFERENCE_ELEMENTWISE.items():
        n = bits // 32
        cplx = op.startswith("complex")

        def records(positive=False):
            r = np.zeros((count, n + 3), dtype=np.uint32)
            r[:, :n] = rng.integers(0, 1 << 32, size=(count, n), dtype=np.uint64).astype(np.uint32)
            r[:, n - 1] |= np.uint32(0x80000000)
            r[:, n] = (rng.integers(-5, 6, size=count) & 0xFFFFFFFF).astype(np.uint32)
            sg = np.ones(count, np.int64) if positive else np.where(rng.integers(0, 2, size=count) == 1, 1, -1)
            r[:, n + 1] = (sg & 0xFFFFFFFF).astype(np.uint32)
            return r

        a = np.concatenate([records(), records()], axis=1) if cplx else records(positive=op == "sqrt")
        b = None if op == "sqrt" else (np.concatenate([records(), records()], axis=1) if cplx else records())
        runs = [capi.run_host(bits, op, a, b) for _ in range(7)][2:]
        dev = sorted(r[1] for r in runs)[2] * 1e3
        wall = sorted(r[2] for r in runs)[2] * 1e3
        row = {"device_ms": round(dev, 4), "wall_ms": round(wall, 4), "reference_device_ms": ref_dev, "reference_wall_ms": ref_wall,
               "speedup_device": round(ref_dev / dev, 1), "speedup_wall": round(ref_wall / wall, 2)}
        out[f"{op}_{bits}"] = row
        log(f"[bench] reference element-wise {op} {bits}: {row}")
    return {"workload": "65,536 values, random_number(5) operands, host arrays (benchmarks/benchmark.cpp:3-60)",
            "reference_gpu": "Apple M5 Max", "ops": out}


def small_batch_chain_section():
    """Multiplication chains on batches too small to fill the GPU with one lane per value (1,024 to 16,384
    values, 64 steps, 512 and 1024 bits) at 1, 2 and 4 lanes per value, timed by the library's tuner
    (lf_autotune_chain: HIP events, best of three after a warm-up); the 65,536-value row is the headline
    batch size, where one lane per value is fastest.  The cooperative forms are checked bit for bit against
    one lane by the GPU tests."""
    from limbforge_amd import capi

    out = {}
    for bits in (512, 1024):
        for count in (1024, 4096, 16384, 65536):
            lanes, ms = capi.autotune_chain(bits, count, 64)
            row = {"ms_lanes_1": round(ms[0], 4), "ms_lanes_2": round(ms[1], 4), "ms_lanes_4": round(ms[2], 4), "tuned_lanes": lanes,
                   "speedup_vs_one_lane": round(ms[0] / min(ms), 2)}
            out[f"{bits}_{count}"] = row
            log(f"[bench] small-batch chain {bits} bits x {count}: {row}")
    return {"workload": "mul chain, 64 steps, random operands (exponents in [-64, 64])", "rows": out}


def reference_linalg_section():
    """The reference's own linalg benchmark workloads at 256 bits, timed as it times them (median of 5 after
    a warm-up): SYRK C = J^T J (lower) and GEMM C = A^T B of QSC-like Jacobians with n = 1000, K = 2000, and the
    normal-equation solve: A = J^T J with Marquardt damping a_ii = RN(a_ii + 2^(e_ii - 10)), Cholesky
    (block 32) and cholesky_solve with 4 right-hand sides (exponent span 20), n = 1000 and 400.  Reports
    this GPU's times beside the reference's published M5 Max times."""
    from limbforge_amd import capi

    bits, w = 256, 8
    rng = np.random.default_rng(20261007 + 20)
    out = {"data": "QSC-like Jacobians (benchmarks/linalg.mm:39-46), 256 bits", "reference_gpu": "Apple M5 Max"}

    def med(xs):
        return sorted(xs)[len(xs) // 2]

    n, K = 1000, 2000
    J = qsc_jacobian(rng, bits, K, n)
    J2 = qsc_jacobian(rng, bits, K, n)
    capi.syrk(bits, J, K, n, lower_only=True)
    runs = [capi.syrk(bits, J, K, n, lower_only=True) for _ in range(5)]
    sd, sw = med([r[2] for r in runs]) * 1e3, med([r[3] for r in runs]) * 1e3
    ref = REFERENCE_LINALG["syrk_1000"]
    out["syrk_1000"] = {"device_ms": round(sd, 3), "wall_ms": round(sw, 3), "moduli": int(runs[0][1].moduli),
                        "reference_device_ms": ref["device_ms"], "reference_wall_ms": ref["wall_ms"],
                        "speedup_device": round(ref["device_ms"] / sd, 2), "speedup_wall": round(ref["wall_ms"] / sw, 2)}
    log(f"[bench] reference workload SYRK n=1000: {out['syrk_1000']}")
    capi.gemm(bits, True, J, J2, n, n, K)
    runs = [capi.gemm(bits, True, J, J2, n, n, K) for _ in range(5)]
    gd, gw = med([r[2] for r in runs]) * 1e3, med([r[3] for r in runs]) * 1e3
    ref = REFERENCE_LINALG["gemm_1000"]
    out["gemm_1000"] = {"device_ms": round(gd, 3), "wall_ms": round(gw, 3), "moduli": int(runs[0][1].moduli),
                        "reference_device_ms": ref["device_ms"], "reference_wall_ms": ref["wall_ms"],
                        "speedup_device": round(ref["device_ms"] / gd, 2), "speedup_wall": round(ref["wall_ms"] / gw, 2)}
    log(f"[bench] reference workload GEMM A^T B n=1000: {out['gemm_1000']}")
    del J2

    for n in (1000, 400):
        Jn = J if n == 1000 else qsc_jacobian(rng, bits, 2 * n, n)
        B = np.zeros((n * 4, w + 3), np.uint32)
        B[:, :w] = rng.integers(0, 1 << 32, size=(n * 4, w), dtype=np.uint64).astype(np.uint32)
        B[:, w - 1] |= np.uint32(0x80000000)
        B[:, w] = (rng.integers(-20, 21, size=n * 4) & 0xFFFFFFFF).astype(np.uint32)
        B[:, w + 1] = (np.where(rng.integers(0, 2, size=n * 4) == 1, 1, -1) & 0xFFFFFFFF).astype(np.uint32)
        diag = np.arange(n) * (n + 1)

        def normal():
            t = time.perf_counter()
            A, _, _, _ = capi.syrk(bits, Jn, 2 * n, n, lower_only=True)
            d = np.zeros((n, w + 3), np.uint32)
            nz = A[diag, w + 1] != 0
            d[nz, w - 1] = 0x80000000
            d[nz, w] = ((A[diag, w].view(np.int32)[nz].astype(np.int64) - 10) & 0xFFFFFFFF).astype(np.uint32)
            d[nz, w + 1] = 1
            A[diag], _, _ = capi.run_host(bits, "add", A[diag], d)
            return A, time.perf_counter() - t

        A, _ = normal()
        L, info = capi.cholesky(bits, A, n, block=32)
        capi.solve(bits, 3, L, n, B, 4, block=32)
        rows = []
        for _ in range(5):
            t = time.perf_counter()
            A, ts = normal()
            tc = time.perf_counter()
            L, info = capi.cholesky(bits, A, n, block=32)
            tf = time.perf_counter() - tc
            if int(info.pivot) != n:
                raise RuntimeError(f"normal matrix n={n} is not positive definite (pivot {int(info.pivot)})")
            X, _, sw_ = capi.solve(bits, 3, L, n, B, 4, block=32)
            rows.append((ts, tf, sw_, time.perf_counter() - t))
        ref = REFERENCE_LINALG[f"cholesky_{n}"]
        res = {"syrk_wall_ms": round(med([r[0] for r in rows]) * 1e3, 3), "factor_wall_ms": round(med([r[1] for r in rows]) * 1e3, 3),
               "solve_wall_ms": round(med([r[2] for r in rows]) * 1e3, 3), "e2e_wall_ms": round(med([r[3] for r in rows]) * 1e3, 3)}
        res.update({f"reference_{k}": v for k, v in ref.items()})
        res["speedup_factor"] = round(ref["factor_wall_ms"] / res["factor_wall_ms"], 2)
        res["speedup_e2e"] = round(ref["e2e_wall_ms"] / res["e2e_wall_ms"], 2)
        out[f"cholesky_{n}"] = res
        log(f"[bench] reference workload Cholesky n={n} + 4-RHS solve: {res}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU work for the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-gemm", action="store_true")
    ap.add_argument("--gemm-reps", type=int, default=6)
    ap.add_argument("--no-cholesky", action="store_true")
    ap.add_argument("--no-elementwise", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-reference-linalg", action="store_true")
    ap.add_argument("--elementwise-reps", type=int, default=5)
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})", file=sys.stderr)
        return 2

    from limbforge_amd import capi

    log(f"[bench] device: {capi.device_name()}")
    x, f = make_inputs(20261007 + BITS)
    X = capi.Buffer(BITS, COUNT)
    F = capi.Buffer(BITS, COUNT)
    O = capi.Buffer(BITS, COUNT)
    X.upload(x)
    F.upload(f)
    stream = capi.Stream()
    e0, e1 = capi.Event(), capi.Event()
    for _ in range(a.warmup):
        capi.mul_chain(BITS, X, F, O, CHAIN, stream=stream.handle)
    stream.synchronize()
    log(f"[bench] warmup done ({a.warmup} steps); timing {a.steps} steps")
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        capi.mul_chain(BITS, X, F, O, CHAIN, stream=stream.handle)
    e1.record(stream)
    stream.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    kernel_ms = e0.elapsed_ms(e1) / a.steps
    value = COUNT / (ms_per_step * 1e-3)
    log(f"[bench] {ms_per_step:.4f} ms/step wall, {kernel_ms:.4f} ms/launch (events)")

    # roofline: algorithmic 32x32->64 word products per launch / launch time, against the measured
    # sustained v_mad_u64_u32 rate (one such instruction per word product is the minimum)
    products = COUNT * CHAIN * (BITS // 32) ** 2
    achieved = products / (kernel_ms * 1e-3) / 1e12
    rate = C.c_double(0)
    capi.check(capi.lib().lf_calibrate_mad64(256 * 64, 4096, C.byref(rate)))
    # a 256 MiB device copy: the reference kernel that scales the HBM traffic counters
    copy_rate = C.c_double(0)
    capi.check(capi.lib().lf_calibrate_copy(256 << 20, C.byref(copy_rate)))
    peak = rate.value / 1e12
    roofline = {
        "bound": "valu",
        "achieved": round(achieved, 4),
        "peak": round(peak, 4),
        "unit": "Tword-products/s",
        "frac": round(achieved / peak, 4),
        # HBM bytes per launch from the FETCH_SIZE and WRITE_SIZE counter passes of this command
        # (rocprofv3 --pmc, one counter per run), each scaled by the same counter's reading for the 256 MiB
        # reference copy lf_calibrate_copy: FETCH 9081.8 KiB x (262144 / 131084.6) + WRITE 8960 KiB x 1.0
        "traffic": TRAFFIC_BYTES,
        "kernel": "mul_chain_kernel<32>",
        "kernel_ms": round(kernel_ms, 5),
        "peak_source": "measured: lf_calibrate_mad64 (8 independent v_mad_u64_u32 chains per lane, 16384 workgroups)",
        # the clock-derived issue bound beside the measured one: 256 CUs x 64 lanes x 2.4 GHz peak engine clock
        "isa_peak": round(isa_mad_peak() / 1e12, 4),
        "frac_of_isa_peak": round(achieved * 1e12 / isa_mad_peak(), 4),
        # achieved counts the schoolbook's N^2 32x32 word products per multiplication (the reference's
        # algorithm); the kernel issues fewer v_mad_u64_u32: a short product of 29-bit digits
        "executed_mad64_per_multiplication": 735,
        "algorithmic_bytes": COUNT * 3 * (4 * 32 + 12),
        "traffic_over_algorithmic": round(TRAFFIC_BYTES / (COUNT * 3 * (4 * 32 + 12)), 3),
        "copy_probe_gbps": round(copy_rate.value / 1e9, 1),
    }

    parity = None
    if not a.no_parity:
        from oracle import lf_oracle

        got = O.download()
        want = lf_oracle.mul_chain(BITS, x, f, CHAIN)
        mism = int((got != want).any(axis=1).sum())
        parity = {"checked_values": COUNT, "mismatches": mism, "against": "CPU oracle (MPFR-pinned)"}
        log(f"[bench] parity: {mism} mismatching values of {COUNT}")

    gemm = None
    if not a.no_gemm:
        log("[bench] 512-bit exact GEMM 512^3")
        gemm = gemm_section(a.gemm_reps, 0 if a.no_parity else 1024, 0 if a.no_cpu_baseline else a.cpu_budget / 3, rate.value)

    ew = None
    if not a.no_elementwise:
        log("[bench] element-wise div / fma / sqrt / complex div, 2^20 values at 256 / 512 / 1024 bits")
        ew = elementwise_section(a.elementwise_reps, rate.value, not a.no_parity, not a.no_cpu_baseline)

    mixed = None
    if not a.no_mixed:
        log("[bench] mixed-precision chains, 10^6 values, bucketed graph vs per-precision launches")
        mixed = mixed_section(rate.value, not a.no_parity, not a.no_cpu_baseline)

    chol = None
    if not a.no_cholesky:
        log("[bench] 256-bit SYRK + Cholesky n=512 + solve (64 right-hand sides)")
        chol = cholesky_section(not a.no_parity)

    small = None
    if not a.no_mixed:
        log("[bench] small-batch chains at 1, 2 and 4 lanes per value")
        small = small_batch_chain_section()

    ref_ew = None
    if not a.no_reference_linalg:
        log("[bench] the reference's element-wise workloads: 65,536 values, host arrays")
        ref_ew = reference_elementwise_section()

    ref_linalg = None
    if not a.no_reference_linalg:
        log("[bench] the reference's linalg workloads: 256-bit QSC-like SYRK / GEMM n=1000, normal-equation Cholesky n=1000 and 400")
        ref_linalg = reference_linalg_section()

    base = None
    if not a.no_cpu_baseline:
        log("[bench] timing the CPU baseline")
        base = cpu_baseline(x, f, a.cpu_budget)

    line = {
        "metric": "1024-bit mul-chain values/s",
        "value": round(value, 1),
        "unit": "values/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_VALUES_PER_S, 3),
        "dtype": "u32",
        "data": "synthetic (seeded random limbs, exponents in [-64, 64])",
        "config": {"workload": "1024-bit mul chain: 65,536 values x 16 steps, resident (config 0)", "bits": BITS,
                   "values": COUNT, "chain_steps": CHAIN, "parallelism": "single GPU"},
        "baseline_note": "vs Apple M5 Max device time 3.192 ms per batch (benchmarks/results/final.csv:145)",
        "roofline": roofline,
        "cpu_baseline": base,
        "parity": parity,
        "gemm": gemm,
        "cholesky": chol,
        "elementwise": ew,
        "mixed": mixed,
        "small_batch_chains": small,
        "reference_linalg": ref_linalg,
        "reference_elementwise": ref_ew,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
