#!/usr/bin/env python3
"""Benchmark of the pynear_amd search path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hamming|l2]

A step is one search of the whole query batch, with the database and the queries already
resident in HBM (the device entry point ``pn_exact_search_device`` on a HIP stream of the
library's own runtime, see pynear_amd/hipdev.py).
Default workload: exact Hamming k-NN, N = 1,000,000 rows x 512 bits, 10,000 queries, k = 10
(uniform random rows; each query is a random row with a Binomial(512, 0.1)-sized random set of
bits flipped; seed 1). ``--workload l2``: exact L2 k-NN, N = 100,000 x 32 float32 U[0,1),
1,000 queries, k = 5 (seed 0).

Prints one JSON line with the throughput (queries/s), the roofline of the dominant kernel
(measured with HIP events around each launch on its stream) and a CPU baseline: the reference's
own VP-tree (built from its sources into oracle/_ref) timed on a bounded query sample on this
host, or the repository's C++ restatement when the reference module is absent.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "QPS @Recall@10>=0.95 (IVF-L2 d768 N1M) + exact L2/Hamming QPS, % roofline"

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_INT32_VALU_TOPS = 39.3   # 256 CUs x 64 lanes x 2.4 GHz, one 32-bit op per lane per clock


# ---------------------------------------------------------------- workloads
def hamming_workload(n=1_000_000, nq=10_000, seed=1, nbytes=64, flip_p=0.1):
    """Uniform random rows; each query = a random row with Binomial(nbits, flip_p) bits flipped."""
    rng = np.random.default_rng(seed)
    X = rng.integers(0, 256, (n, nbytes), dtype=np.uint8)
    nbits = nbytes * 8
    src = rng.integers(0, n, nq)
    nflip = rng.binomial(nbits, flip_p, nq)
    order = np.argsort(rng.random((nq, nbits)), axis=1)
    mask_bits = np.zeros((nq, nbits), np.uint8)
    rows = np.repeat(np.arange(nq), nflip)
    cols = order[np.arange(nbits)[None, :] < nflip[:, None]]
    mask_bits[rows, cols] = 1
    Q = X[src] ^ np.packbits(mask_bits, axis=1, bitorder="little")
    return X, Q


def l2_workload(n=100_000, nq=1_000, d=32, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.random((n, d), dtype=np.float32)
    Q = rng.random((nq, d), dtype=np.float32)
    return X, Q


WORKLOADS = {
    "hamming": dict(desc="exact Hamming kNN", n=1_000_000, nq=10_000, k=10, width="512-bit", seed=1),
    "l2": dict(desc="exact L2 kNN float32 U[0,1)", n=100_000, nq=1_000, k=5, width="d=32", seed=0),
}


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline(workload, X, Q, k, gpu_idx):
    """Reference VP-tree (oracle/_ref) on a bounded query sample; falls back to the C++ port."""
    sys.path.insert(0, ROOT)
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    R = oracle.load_ref()
    sample = 128 if workload == "hamming" else 1000
    sample = min(sample, len(Q))
    Qs = Q[:sample]
    out = {"cores": cores, "unit": "queries/s"}
    if R is not None:
        idx = R.VPTreeBinaryIndex512() if workload == "hamming" else R.VPTreeL2Index()
        t0 = time.perf_counter()
        idx.set(X)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        ri, rd = idx.searchKNN(Qs, k)
        t = time.perf_counter() - t0
        out.update(value=sample / t, kind="reference",
                   sample=f"{sample} of {len(Q)} queries vs all {len(X)} rows; reference VPTree "
                          f"{type(idx).__name__} (OpenMP, {cores} threads), search only; build {t_build:.1f} s excluded")
        ref_d = np.sort(np.asarray(rd, dtype=np.float64), axis=1)
    else:
        t0 = time.perf_counter()
        if workload == "hamming":
            oi, od = oracle.knn_hamming(X, Qs, k)
        else:
            oi, od = oracle.knn_float(oracle.L2, X, Qs, k)
        t = time.perf_counter() - t0
        out.update(value=sample / t, kind="port",
                   sample=f"{sample} of {len(Q)} queries vs all {len(X)} rows; C++ brute-force restatement "
                          f"(OpenMP, {cores} threads)")
        ref_d = np.asarray(od, dtype=np.float64)
    return out, ref_d


# ---------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hamming", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on a single "
              "MI355X; multi-GPU sharding is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    from pynear_amd import _native as N
    from pynear_amd import hipdev

    W = WORKLOADS[args.workload]
    k = W["k"]
    if args.workload == "hamming":
        X, Q = hamming_workload(W["n"], W["nq"], seed=W["seed"])
        metric = N.METRIC_HAMMING
        dist_dtype = np.int32
    else:
        X, Q = l2_workload(W["n"], W["nq"], seed=W["seed"])
        metric = N.METRIC_L2
        dist_dtype = np.float32
    n, nq = len(X), len(Q)

    h = N.ExactHandle(metric)
    h.set(X)
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((nq, k), np.int32)
    od = hipdev.DeviceArray((nq, k), dist_dtype)
    stream = hipdev.Stream()

    def step():
        h.search_device(dq.ptr, nq, k, oi.ptr, od.ptr, stream.handle)

    for _ in range(args.warmup):
        step()
    stream.synchronize()

    N.stage_timing(True)
    N.stage_times()
    ev0, ev1 = hipdev.Event(), hipdev.Event()
    hipdev.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    ev1.synchronize()
    hipdev.synchronize()
    wall = time.perf_counter() - t0
    N.stage_timing(False)
    stages = N.stage_times()
    gpu_ms = ev0.elapsed_ms(ev1)
    ms_per_step = gpu_ms / args.steps
    qps = nq * args.steps / (gpu_ms / 1e3)

    scan_ms, scan_n = stages["scan"]
    scan_avg_s = scan_ms / max(scan_n, 1) / 1e3
    if args.workload == "hamming":
        words = X.shape[1] // 4
        ops = 2.0 * words * nq * n          # XOR + popcount-accumulate per 32-bit word per pair
        achieved = ops / scan_avg_s / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT32_VALU_TOPS, "unit": "TOP/s",
                "frac": round(achieved / PEAK_INT32_VALU_TOPS, 4), "traffic": None,
                "kernel": "hamming_scan_kernel<16,16,4>", "kernel_ms": round(scan_avg_s * 1e3, 4),
                "work": f"{ops:.3e} int32 ops per launch (16 words x (xor + popcount-add) x Q x N)"}
        dtype = "u32"
    else:
        d = X.shape[1]
        flops = 2.0 * d * nq * n
        achieved = flops / scan_avg_s / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "kernel": "float_scan_kernel<32,8,L2,2>", "kernel_ms": round(scan_avg_s * 1e3, 4),
                "work": f"{flops:.3e} flop per launch (2 x d x Q x N)"}
        dtype = "f32"

    result = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"{W['desc']}, N={n:,} x {W['width']}, Q={nq:,}, k={k}", "global_batch": nq,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "stages_ms_per_step": {s: round(v[0] / max(args.steps, 1), 4) for s, v in stages.items() if v[1]},
        "wall_s_timed": round(wall, 4),
        "tuning": {name: N.tune_get(name) for name in ("hamming.wg_per_cu", "float.wg_per_cu", "merge.wave_per_query")},
    }
    if not args.no_cpu_baseline:
        got_d = od.numpy().astype(np.float64)
        cb, ref_d = cpu_baseline(args.workload, X, Q, k, oi)
        cb["value"] = round(cb["value"], 2)
        result["cpu_baseline"] = cb
        result["cpu_sample_distances_equal"] = bool(np.array_equal(got_d[: len(ref_d)], ref_d))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
