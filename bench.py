#!/usr/bin/env python3
"""Benchmark of the pynear_amd search path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload ivf|hamming|l2|batch|learned|ivf_binary|range|mih] [--no-cpu-baseline]

A step is one search of the whole query batch with the database and the queries already resident
in HBM (the device entry points on a HIP stream of the library's own runtime). Workloads:

* ivf (default; the headline): IVFFlatL2Index, N = 1,000,000 x d = 768 float32 drawn from a
  1,024-component Gaussian mixture (centres ~ N(0,1), sigma 0.3, seed 2), nlist = 1,024 (the
  reference's kmeans_l2 with 100 iterations, seed 42, run on the GPU), 10,000 queries from the same
  mixture, k = 10, n_probe in {8, 16, 32}. `value` is the QPS at the smallest n_probe whose
  Recall@10 against the exact search is >= 0.95. The exact Hamming and exact L2 workloads below
  are measured in the same run and reported under "also".
* hamming: exact Hamming k-NN, N = 1,000,000 x 512 bits (uniform bytes), 10,000 queries (a random
  row with a Binomial(512, 0.1)-sized random set of bits flipped), k = 10, seed 1.
* l2: exact L2 k-NN, N = 100,000 x 32 float32 U[0,1), 1,000 queries, k = 5, seed 0.
* batch: 64 independent exact L2 indices (N ~ LogU[1k, 200k], d in {64, 128, 256}, N(0, 1) rows,
  Q ~ LogU[1, 4096], k in {1, 10, 32}, seed 3) served by one planned graph (autotuned schedule);
  results are checked equal to the per-index plain searches.
* learned: IVF-L2 with a learned probe predictor (MLP 512 -> 256 -> 4096, bf16, trained on the GPU
  on 200,000 queries' exact top-10) and adaptive n_probe, N = 2,000,000 x 512 from a Zipf(1.1)
  weighted 4,096-component mixture (sigma 0.25), nlist = 4,096, 10,000 held-out queries, k = 10;
  compared with fixed-n_probe IVF (bf16x3 ~ fp32 candidate distances, and bf16) at Recall@10 >= 0.95.
* ivf_binary: IVFFlatBinaryIndex over the hamming rows and queries, nlist 512, n_probe 16, k = 10.
* range: Hamming range search (the BK-tree classes' find_threshold) over the hamming rows, 10,000
  queries with Binomial(512, 0.06) flips, threshold 60 (host entry point; device time reported).
* mih: MIHBinaryIndex over the hamming rows, m = 8, radius 16, k = 10, queries with
  Binomial(512, 0.02) flips (host entry point; device time reported).

The roofline of the dominant kernel is measured with HIP events around each launch on its
stream. The CPU baseline runs on this host: the reference's own VP-tree (built from its sources
into oracle/_ref) for the exact workloads, and the numpy restatement of the reference's
IVFFlatL2Index search over the same clusters for the IVF workload.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "QPS @Recall@10>=0.95 (IVF-L2 d768 N1M) + exact L2/Hamming QPS, % roofline"

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP8_TFLOPS = 5000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_INT32_VALU_TOPS = 39.3   # 256 CUs x 64 lanes x 2.4 GHz, one 32-bit op per lane per clock


# ---------------------------------------------------------------- workloads
def hamming_workload(n=1_000_000, nq=10_000, seed=1, nbytes=64, flip_p=0.1, with_source=False):
    """Uniform random rows; each query = a random row with Binomial(nbits, flip_p) bits flipped
    (distinct bits, so the query's distance to its source row is its flip count). with_source
    also returns the source rows and the flip counts."""
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
    if with_source:
        return X, Q, src, nflip
    return X, Q


def l2_workload(n=100_000, nq=1_000, d=32, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.random((n, d), dtype=np.float32)
    Q = rng.random((nq, d), dtype=np.float32)
    return X, Q


def gmm_workload(n=1_000_000, nq=10_000, d=768, comps=1024, sigma=0.3, seed=2):
    """Gaussian mixture: centres ~ N(0,1)^d, equal weights, per-component sigma; rows then queries."""
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((comps, d), dtype=np.float32)

    def draw(m):
        out = np.empty((m, d), np.float32)
        for s in range(0, m, 65536):
            e = min(m, s + 65536)
            lab = rng.integers(0, comps, e - s)
            out[s:e] = centres[lab] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        return out

    return draw(n), draw(nq)


def batch_workload(n_idx=64, seed=3):
    """64 independent exact L2 indices: N ~ LogU[1k, 200k], d in {64, 128, 256}, rows ~ N(0, 1),
    Q ~ LogU[1, 4096] queries, k in {1, 10, 32}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_idx):
        n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        d = int(rng.choice([64, 128, 256]))
        nq = max(1, int(round(np.exp(rng.uniform(0.0, np.log(4096))))))
        k = int(rng.choice([1, 10, 32]))
        X = rng.standard_normal((n, d), dtype=np.float32)
        Q = rng.standard_normal((nq, d), dtype=np.float32)
        out.append((X, Q, k))
    return out


def learned_workload(n=2_000_000, d=512, comps=4096, sigma=0.25, n_train=200_000, n_test=10_000):
    """4096-component Gaussian mixture in d = 512 with Zipf(1.1) component weights; rows (seed 4),
    training queries (seed 5) and held-out queries (seed 6) from the same mixture."""
    crng = np.random.default_rng(4)
    centres = crng.standard_normal((comps, d), dtype=np.float32)
    w = 1.0 / np.arange(1, comps + 1) ** 1.1
    w /= w.sum()

    def draw(m, seed):
        rng = np.random.default_rng(seed)
        out = np.empty((m, d), np.float32)
        for s in range(0, m, 65536):
            e = min(m, s + 65536)
            lab = rng.choice(comps, e - s, p=w)
            out[s:e] = centres[lab] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        return out

    return draw(n, 4), draw(n_train, 5), draw(n_test, 6)


# ---------------------------------------------------------------- helpers
def _time_device(N, hipdev, step, steps, warmup, stream):
    """Times `steps` steps between two HIP events with nothing else on the stream, then runs the
    same number of steps again with per-stage HIP events (the stage breakdown and per-launch kernel
    times; those events add their own gaps, which stay out of the timed steps)."""
    for _ in range(warmup):
        step()
    stream.synchronize()
    N.stage_timing(False)
    ev0, ev1 = hipdev.Event(), hipdev.Event()
    hipdev.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    ev1.synchronize()
    hipdev.synchronize()
    wall = time.perf_counter() - t0
    N.stage_timing(True)
    N.stage_times()
    for _ in range(steps):
        step()
    stream.synchronize()
    N.stage_timing(False)
    stages = N.stage_times()
    return ev0.elapsed_ms(ev1), wall, stages


def _warm(step, stream, seconds):
    """Runs step until `seconds` of wall time have passed; returns the seconds per step."""
    t0 = time.perf_counter()
    n = 0
    while True:
        step()
        n += 1
        if n % 8 == 0:
            stream.synchronize()
            el = time.perf_counter() - t0
            if el >= seconds:
                return el / n


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _cpu_model():
    """The host CPU's model name (as lscpu prints it), for the CPU baselines."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_exact(workload, steps, warmup, cpu, settle=False):
    from pynear_amd import _native as N
    from pynear_amd import hipdev
    if workload == "hamming":
        X, Q = _hamming_data()[:2]
        k, metric, dt = 10, N.METRIC_HAMMING, np.int32
        desc = f"exact Hamming kNN, N={len(X):,} x 512-bit, Q={len(Q):,}, k={k}"
    else:
        X, Q = l2_workload()
        k, metric, dt = 5, N.METRIC_L2, np.float32
        desc = f"exact L2 kNN float32 U[0,1), N={len(X):,} x d=32, Q={len(Q):,}, k={k}"
    n, nq = len(X), len(Q)
    _log(f"{workload}: data {X.shape}")
    h = N.ExactHandle(metric)
    h.set(X)
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((nq, k), np.int32)
    od = hipdev.DeviceArray((nq, k), dt)
    stream = hipdev.Stream()
    step = lambda: h.search_device(dq.ptr, nq, k, oi.ptr, od.ptr, stream.handle)
    if settle:
        # (secondary line) a step here is short (0.1 ms for L2): warm the GPU up for 0.3 s so the
        # timed steps run at full clocks, and time at least 0.2 s of steps
        per = _warm(step, stream, 0.3)
        steps = max(steps, int(math.ceil(0.2 / max(per, 1e-6))))
    gpu_ms, wall, stages = _time_device(N, hipdev, step, steps, warmup, stream)
    scan_ms, scan_n = stages["scan"]
    scan_s = scan_ms / max(scan_n, 1) / 1e3
    if workload == "hamming":
        ops = 2.0 * (X.shape[1] // 4) * nq * n
        ach = ops / scan_s / 1e12
        roof = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_INT32_VALU_TOPS, "unit": "TOP/s",
                "frac": round(ach / PEAK_INT32_VALU_TOPS, 4),
                "traffic": measured_traffic("hamming_scan_kernel<16, 16, 1, 4, 0>"),
                "kernel": "hamming_scan_kernel<16, 16, 1, 4, 0>", "kernel_ms": round(scan_s * 1e3, 4),
                "work": f"{ops:.3e} int32 ops per launch: 16 words x (xor + popcount-accumulate) x Q x N"}
        _traffic_ratio(roof, X.nbytes)
    else:
        flops = 2.0 * X.shape[1] * nq * n
        ach = flops / scan_s / 1e12
        # the result is fp32-exact (bit-identical to the reference), so the work is priced as fp32
        # flops against the fp32 peak; the kernel gets there with 3 bf16 MFMA products per term
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                "traffic": measured_traffic("l2_mfma_scan_small_kernel<4, 3, 2>"),
                "kernel": "l2_mfma_scan_small_kernel<4, 3, 2>", "kernel_ms": round(scan_s * 1e3, 4),
                "work": f"{flops:.3e} flop per launch (2 d Q N)"}
    out = {"workload": desc, "value": round(nq * steps / (gpu_ms / 1e3), 1), "unit": "queries/s",
           "ms_per_step": round(gpu_ms / steps, 4), "steps": steps, "roofline": roof,
           "stages_ms_per_step": {s: round(v[0] / steps, 4) for s, v in stages.items() if v[1]}}
    o1i = o1d = None
    if workload == "l2":
        # search1NN over the same rows and queries (k = 1), timed the same way
        o1i = hipdev.DeviceArray((nq, 1), np.int32)
        o1d = hipdev.DeviceArray((nq, 1), dt)
        step1 = lambda: h.search_device(dq.ptr, nq, 1, o1i.ptr, o1d.ptr, stream.handle)
        if settle:
            _warm(step1, stream, 0.2)
        ms1, _, _ = _time_device(N, hipdev, step1, steps, warmup, stream)
        out["search1NN"] = {"value": round(nq * steps / (ms1 / 1e3), 1), "unit": "queries/s",
                            "ms_per_step": round(ms1 / steps, 4), "steps": steps}
    if cpu:
        sys.path.insert(0, ROOT)
        import oracle
        R = oracle.load_ref()
        sample = min(128 if workload == "hamming" else 1000, nq)
        if R is not None:
            idx = R.VPTreeBinaryIndex512() if workload == "hamming" else R.VPTreeL2Index()
            t0 = time.perf_counter()
            idx.set(X)
            tb = time.perf_counter() - t0
            t0 = time.perf_counter()
            _, rd = idx.searchKNN(Q[:sample], k)
            t = time.perf_counter() - t0
            kind = "reference"
            samp = (f"{sample} of {nq} queries vs all {n} rows; reference {type(idx).__name__} (OpenMP, "
                    f"{_cores()} threads), search only, build {tb:.1f} s excluded")
            ref_d = np.sort(np.asarray(rd, np.float64), axis=1)
        else:
            t0 = time.perf_counter()
            _, rd = (oracle.knn_hamming(X, Q[:sample], k) if workload == "hamming"
                     else oracle.knn_float(oracle.L2, X, Q[:sample], k))
            t = time.perf_counter() - t0
            kind, samp = "port", f"{sample} of {nq} queries vs all {n} rows; C++ brute-force restatement"
            ref_d = np.asarray(rd, np.float64)
        out["cpu_baseline"] = {"value": round(sample / t, 2), "unit": "queries/s", "cores": _cores(), "kind": kind,
                               "sample": samp}
        out["cpu_sample_distances_equal"] = bool(np.array_equal(od.numpy()[:sample].astype(np.float64), ref_d))
        if o1d is not None and R is not None:
            t0 = time.perf_counter()
            _, rd1 = idx.search1NN(Q[:sample])
            t1 = time.perf_counter() - t0
            out["search1NN"]["cpu_baseline"] = {"value": round(sample / t1, 2), "unit": "queries/s", "cores": _cores(),
                                                "kind": "reference", "sample": f"{sample} of {nq} queries, "
                                                f"reference {type(idx).__name__}.search1NN, same tree"}
            out["search1NN"]["cpu_sample_distances_equal"] = bool(
                np.array_equal(o1d.numpy()[:sample, 0], np.asarray(rd1, np.float32)))
    return out


def run_ivf(steps, warmup, cpu, probes=(8, 16, 32), precisions=("bf16", "fp8")):
    from pynear_amd import _native as N
    from pynear_amd import hipdev
    t0 = time.perf_counter()
    X, Q = gmm_workload()
    t_gen = time.perf_counter() - t0
    _log(f"ivf: generated {X.shape} in {t_gen:.1f} s")
    n, d = X.shape
    nq, k, nlist = len(Q), 10, 1024
    h = N.IvfL2Handle()
    t0 = time.perf_counter()
    labels, cent = h.build(X, nlist, 100, 42)
    t_build = time.perf_counter() - t0
    _log(f"ivf: built (kmeans + lists) in {t_build:.1f} s")
    # ground truth: exact GPU search (bit-exact reference arithmetic)
    ex = N.ExactHandle(N.METRIC_L2)
    t0 = time.perf_counter()
    ex.set(X)
    gt, _ = ex.search(Q, k)
    t_exact = time.perf_counter() - t0
    _log(f"ivf: exact ground truth in {t_exact:.1f} s")
    del ex
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((nq, k), np.int32)
    od = hipdev.DeviceArray((nq, k), np.float32)
    oc = hipdev.DeviceArray((nq,), np.int32)
    stream = hipdev.Stream()
    sizes = np.bincount(labels, minlength=nlist)
    sweeps, results = {}, {}
    for prec in precisions:  # candidate distances: bf16 or fp8 e4m3 products, fp32 rerank either way
        sweep = sweeps[prec] = {}
        for npb in probes:
            hi, hd, hc = h.search(Q, npb, k, prec)
            results[prec, npb] = (hi, hd, hc)
            recall = float(np.mean([len(set(hi[i, :hc[i]].tolist()) & set(gt[i].tolist())) / k for i in range(nq)]))
            step = lambda: h.search_device(dq.ptr, nq, npb, k, oi.ptr, od.ptr, oc.ptr, prec, stream.handle)
            # warm the GPU up for 0.3 s (untimed) so the timed steps run at full clocks: the warmup
            # steps alone last a few ms
            _warm(step, stream, 0.3)
            nst = steps
            gpu_ms, wall, stages = _time_device(N, hipdev, step, nst, warmup, stream)
            sweep[npb] = {"qps": round(nq * nst / (gpu_ms / 1e3), 1), "ms_per_step": round(gpu_ms / nst, 4),
                          "steps": nst, "recall_at_10": round(recall, 4),
                          "stages_ms_per_step": {s: round(v[0] / nst, 4) for s, v in stages.items() if v[1]},
                          "scan_launch_ms": stages["scan"][0] / max(stages["scan"][1], 1)}
            _log(f"ivf: {prec} n_probe={npb} {sweep[npb]}")
    # per precision the smallest n_probe reaching Recall@10 >= 0.95; the faster of the two is reported
    best = {}
    for prec, sweep in sweeps.items():
        ok = [p for p in probes if sweep[p]["recall_at_10"] >= 0.95]
        if ok:
            best[prec] = ok[0]
    if not best:
        best = {precisions[0]: probes[-1]}
    prec = max(best, key=lambda pr_: sweeps[pr_][best[pr_]]["qps"])
    chosen = best[prec]
    S = sweeps[prec][chosen]
    scan_s = S["scan_launch_ms"] / 1e3
    dp = (d + 63) // 64 * 64
    ebytes = 1 if prec == "fp8" else 2
    kname = "ivfl2_scan_fp8_kernel<8, 3>" if prec == "fp8" else "ivfl2_scan_kernel<1, 8>"
    peak_mm = PEAK_FP8_TFLOPS if prec == "fp8" else PEAK_BF16_TFLOPS
    # algorithmic work of one scan launch: the rows of every probed cluster once (bf16: 2 B, fp8: 1 B
    # per dim) and 2 d flops per (query, probed row); the probes are recomputed here in fp64
    # (exact top-n_probe)
    keep = np.flatnonzero(sizes)
    C64 = cent[keep].astype(np.float64)
    pr = np.argpartition((C64 * C64).sum(1)[None, :] - 2.0 * (Q.astype(np.float64) @ C64.T), chosen - 1,
                         axis=1)[:, :chosen]
    probed_rows = sizes[keep][pr]
    bytes_ = float(sizes[keep][np.unique(pr)].sum() * dp * ebytes)
    flops = 2.0 * d * float(probed_rows.sum())
    executed = None
    if prec == "fp8":
        # the fp8 scan skips the row groups that no query can need (segment bounds); what the launch
        # executed is counted on the device (MFMA instructions of 2 x 32 x 32 x 64 flops, fp8 row
        # bytes its work items read, each row block once per item) and reported beside the
        # algorithmic work, which the roofline uses (each probed list's bytes once, no re-reads)
        h.search_device(dq.ptr, nq, chosen, k, oi.ptr, od.ptr, oc.ptr, prec, stream.handle)
        stream.synchronize()
        mfma, fetched = h.scan_work()
        executed = {"mfma_flops": float(mfma) * 2 * 32 * 32 * 64, "row_bytes_read": float(fetched)}
    bw = bytes_ / scan_s / 1e9
    tf = flops / scan_s / 1e12
    if bw / PEAK_HBM_GBS >= tf / peak_mm:
        roof = {"bound": "hbm", "achieved": round(bw, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(bw / PEAK_HBM_GBS, 4)}
    else:
        roof = {"bound": "mfma", "achieved": round(tf, 2), "peak": peak_mm, "unit": "TFLOP/s",
                "frac": round(tf / peak_mm, 4)}
    what = f"{prec} rows of the probed clusters once, 2 d flops per (query, probed row)"
    roof.update({"traffic": measured_traffic(kname), "kernel": kname, "kernel_ms": round(scan_s * 1e3, 4),
                 "work": f"{bytes_:.3e} B and {flops:.3e} flop per launch ({what})",
                 "other_bound": {"hbm_GBs": round(bw, 1), "mfma_TFLOPs": round(tf, 2)}})
    if executed:
        roof["executed_work"] = {k_: float(f"{v:.4e}") for k_, v in executed.items()}
        roof["executed_work"]["note"] = ("after group pruning; rows re-read once per work item of "
                                         "up to 128 queries of a list")
    out = {"workload": (f"IVFFlatL2Index GMM N={n:,} x d={d} (1,024 comps, sigma 0.3), nlist={nlist}, "
                        f"Q={nq:,}, k={k}, n_probe={chosen} (smallest with Recall@10 >= 0.95), "
                        f"{prec} candidate distances + fp32 rerank of max(4k, 40)"),
           "value": S["qps"], "unit": "queries/s", "ms_per_step": S["ms_per_step"], "steps": S["steps"],
           "roofline": roof,
           "dtype": "fp8e4m3+f32" if prec == "fp8" else "bf16+f32",
           "precision": prec, "recall_at_10": S["recall_at_10"], "n_probe": chosen,
           "stages_ms_per_step": S["stages_ms_per_step"],
           "sweep": {pr_: {str(p): {kk: v for kk, v in s.items() if kk != "scan_launch_ms"} for p, s in sw.items()}
                     for pr_, sw in sweeps.items()},
           "setup_s": {"generate": round(t_gen, 1), "build_kmeans_lists": round(t_build, 1),
                       "exact_ground_truth": round(t_exact, 1)}}
    # error of the returned squared distances against fp64 recomputation
    hi, hd, hc = results[prec, chosen]
    ok = hi >= 0
    rows = X[np.where(ok, hi, 0)].astype(np.float64)
    ref = ((rows - Q[:, None, :].astype(np.float64)) ** 2).sum(-1)
    rel = np.abs(hd.astype(np.float64) - ref) / np.maximum(ref, 1e-30)
    out["fp64_rel_err"] = {"max": float(f"{rel[ok].max():.3e}"), "mean": float(f"{rel[ok].mean():.3e}"),
                           "what": "returned squared L2 vs fp64 recomputation, all (query, neighbour) pairs"}
    del rows, ref
    if cpu:
        sys.path.insert(0, ROOT)
        import oracle
        data = [X[labels == c] for c in range(nlist)]
        ids = [np.where(labels == c)[0] for c in range(nlist)]
        keep = [c for c in range(nlist) if len(ids[c])]
        data = [data[c] for c in keep]
        ids = [ids[c] for c in keep]
        norms = [(x * x).sum(1) for x in data]
        sample = 100
        t0 = time.perf_counter()
        ci, cd = oracle.ivf_l2_search(cent[keep], data, norms, ids, Q[:sample], chosen, k)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sample / t, 2), "unit": "queries/s", "cores": _cores(), "kind": "port",
                               "sample": (f"{sample} of {nq} queries; numpy restatement of the reference's "
                                          f"IVFFlatL2Index.searchKNN over the same {len(keep)} clusters at "
                                          f"n_probe={chosen} (BLAS threads = host cores)")}
        hi, _, hc = h.search(Q[:sample], chosen, k, prec)
        out["cpu_sample_recall_vs_port"] = round(float(np.mean(
            [len(set(hi[i, :hc[i]].tolist()) & set(ci[i])) / k for i in range(sample)])), 4)
    return out


_HAM_CACHE = {}


def _hamming_data(flip_p=0.1):
    """The 1M x 512-bit rows and 10k queries (flip probability flip_p), generated once per value."""
    if flip_p not in _HAM_CACHE:
        _HAM_CACHE[flip_p] = hamming_workload(flip_p=flip_p, with_source=True)
    return _HAM_CACHE[flip_p]


def _host_stage_timing(N, call, reps):
    """Runs a host entry point `reps` times with per-stage HIP events; returns (wall s per call,
    {stage: ms per call})."""
    call()
    N.stage_timing(True)
    N.stage_times()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    wall = (time.perf_counter() - t0) / reps
    st = N.stage_times()
    N.stage_timing(False)
    return wall, {s: v[0] / reps for s, v in st.items() if v[1]}


def run_ivf_binary(steps, warmup, cpu, nlist=512, nprobe=16, k=10):
    """IVFFlatBinaryIndex: N = 1M x 512-bit, nlist 512, n_probe 16, k = 10, 10k queries."""
    from pynear_amd import _native as N
    from pynear_amd import hipdev
    X, Q, src, _ = _hamming_data()
    n, nq = len(X), len(Q)
    t0 = time.perf_counter()
    h = N.IvfBinHandle(nlist, nprobe, 20, 42)
    h.set(X)
    t_build = time.perf_counter() - t0
    cent, labels = h.export()
    _log(f"ivf_binary: built in {t_build:.1f} s")
    # rows scanned per query: the n_probe nearest centroids by (distance, id), summed list sizes
    sizes = np.bincount(labels, minlength=len(cent))
    probed = 0
    for s in range(0, nq, 500):
        dc = np.bitwise_count(Q[s:s + 500, None, :] ^ cent[None]).sum(-1)
        order = np.lexsort((np.broadcast_to(np.arange(len(cent)), dc.shape), dc), axis=1)[:, :nprobe]
        probed += int(sizes[order].sum())
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((nq, k), np.int32)
    od = hipdev.DeviceArray((nq, k), np.int32)
    oc = hipdev.DeviceArray((nq,), np.int32)
    stream = hipdev.Stream()
    step = lambda: h.search_device(dq.ptr, nq, k, oi.ptr, od.ptr, oc.ptr, stream.handle)
    gpu_ms, _, stages = _time_device(N, hipdev, step, steps, warmup, stream)
    scan_ms, scan_n = stages["scan"]
    scan_s = scan_ms / max(scan_n, 1) / 1e3
    ops = 2.0 * (X.shape[1] // 4) * probed
    ach = ops / scan_s / 1e12
    ex = N.ExactHandle(N.METRIC_HAMMING)
    ex.set(X)
    gt = ex.search(Q, k)[0]
    got = oi.numpy()
    recall = float(np.mean([len(set(got[i].tolist()) & set(gt[i].tolist())) / k for i in range(nq)]))
    # uniform random rows have no cluster structure: beyond the query's source row its exact
    # neighbours are arbitrary rows ~230 bits away, so Recall@10 says little; the source row
    # (the true nearest, ~51 bits away) is what the probe must find
    found = float(np.mean(got[:, 0] == src))
    out = {"workload": f"IVFFlatBinaryIndex N={n:,} x 512-bit, nlist={nlist}, n_probe={nprobe}, Q={nq:,}, k={k}",
           "value": round(nq * steps / (gpu_ms / 1e3), 1), "unit": "queries/s", "ms_per_step": round(gpu_ms / steps, 4),
           "steps": steps, "recall_at_10": round(recall, 4), "source_found": round(found, 4), "setup_s": round(t_build, 2),
           "stages_ms_per_step": {s: round(v[0] / steps, 4) for s, v in stages.items() if v[1]},
           "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_INT32_VALU_TOPS, "unit": "TOP/s",
                        "frac": round(ach / PEAK_INT32_VALU_TOPS, 4),
                        "traffic": measured_traffic("ivfbin_list_scan_kernel<16, 10, 4>"),
                        "kernel": "ivfbin_list_scan_kernel<16, 10, 4>", "kernel_ms": round(scan_s * 1e3, 4),
                        "work": (f"{ops:.3e} int32 ops per launch: 16 words x (xor + popcount-accumulate) x "
                                 f"{probed:,} probed rows (all queries), one distance per candidate")}}
    if cpu:
        sys.path.insert(0, ROOT)
        import oracle
        R = oracle.load_ref()
        sample = 1000
        if R is not None:
            t0 = time.perf_counter()
            ref = R.IVFFlatBinaryIndex(nlist=nlist, nprobe=nprobe, max_iter=20, seed=42)
            ref.set(X)
            tb = time.perf_counter() - t0
            t0 = time.perf_counter()
            ri, rd = ref.searchKNN(Q[:sample], k)
            t = time.perf_counter() - t0
            kind, samp = "reference", (f"{sample} of {nq} queries; reference IVFFlatBinaryIndex (serial query loop), "
                                       f"search only, build {tb:.1f} s excluded")
        else:
            off, ids = oracle.lists_from_labels(labels, len(cent))
            t0 = time.perf_counter()
            ri, rd = oracle.ivfbin_search(cent, off, ids, X, Q[:sample], nprobe, k)
            t = time.perf_counter() - t0
            kind, samp = "port", f"{sample} of {nq} queries; C++ restatement of the reference search (serial)"
        out["cpu_baseline"] = {"value": round(sample / t, 2), "unit": "queries/s", "cores": 1, "kind": kind,
                               "sample": samp}
        gc = oc.numpy()
        out["cpu_sample_results_equal"] = bool(all(got[i, :gc[i]].tolist() == list(ri[i]) for i in range(sample)))
    return out


def run_hamming_range(cpu, threshold=60, reps=5):
    """Hamming range search (BK-tree find_threshold) over the 1M x 512-bit rows, 10k queries with
    6% of their bits flipped, threshold 60."""
    from pynear_amd import _native as N
    X, Q, src, nflip = _hamming_data(0.06)
    n, nq = len(X), len(Q)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    res = [None]

    def call():
        res[0] = h.range(Q, threshold)

    wall, st = _host_stage_timing(N, call, reps)
    off, idx, dist = res[0]
    dev_ms = st.get("scan", 0.0) + st.get("select", 0.0)
    ops = 2.0 * (X.shape[1] // 4) * n * nq
    ach = ops / (st["scan"] / 1e3) / 1e12
    found = float(np.mean([src[q] in idx[off[q]:off[q + 1]] for q in range(nq) if nflip[q] <= threshold]))
    out = {"workload": f"Hamming range search N={n:,} x 512-bit, Q={nq:,}, threshold {threshold}",
           "value": round(nq / (dev_ms / 1e3), 1), "unit": "queries/s", "ms_per_call_device": round(dev_ms, 4),
           "host_call_qps": round(nq / wall, 1), "matches": int(off[-1]), "source_found": round(found, 4),
           "stages_ms_per_call": {s: round(v, 4) for s, v in st.items()},
           "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_INT32_VALU_TOPS, "unit": "TOP/s",
                        "frac": round(ach / PEAK_INT32_VALU_TOPS, 4),
                        "traffic": measured_traffic("range_scan_kernel<16, 4, false>"),
                        "kernel": "range_scan_kernel<16, 4, false>", "kernel_ms": round(st["scan"], 4),
                        "work": f"{ops:.3e} int32 ops per count pass: 16 words x 2 ops x Q x N"}}
    if cpu:
        sys.path.insert(0, ROOT)
        import oracle
        R = oracle.load_ref()
        sample = 200
        if R is not None:
            t0 = time.perf_counter()
            bk = R.BKTreeBinaryIndex512()
            bk.set(X)
            tb = time.perf_counter() - t0
            t0 = time.perf_counter()
            ri, rd, _ = bk.find_threshold(Q[:sample], threshold)
            t = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(sample / t, 2), "unit": "queries/s", "cores": _cores(),
                                   "kind": "reference",
                                   "sample": (f"{sample} of {nq} queries; reference BKTreeBinaryIndex512.find_threshold "
                                              f"(OpenMP), build {tb:.1f} s excluded")}
            out["cpu_sample_sets_equal"] = bool(all(sorted(ri[q]) == idx[off[q]:off[q + 1]].tolist()
                                                    for q in range(sample)))
    return out


def run_mih(cpu, m=8, radius=16, k=10, reps=5):
    """MIHBinaryIndex over the 1M x 512-bit rows: m = 8, radius 16, k = 10, 10k queries with 2% of
    their bits flipped."""
    from pynear_amd import _native as N
    X, Q, src, nflip = _hamming_data(0.02)
    n, nq = len(X), len(Q)
    h = N.MihHandle(m)
    h.set(X)
    res = [None]

    def call():
        res[0] = h.search(Q, k, radius)

    wall, st = _host_stage_timing(N, call, reps)
    idx, dist, cnt = res[0]
    dev_ms = sum(st.values())
    scan = st["scan"]
    ops = (2.0 * (X.shape[1] // 4) + 2.0 * m) * n * nq  # xor + popcount per word, add + min per sub-string
    ach = ops / (scan / 1e3) / 1e12
    found = float(np.mean([idx[q, 0] == src[q] for q in range(nq) if nflip[q] <= radius]))
    out = {"workload": f"MIHBinaryIndex N={n:,} x 512-bit, m={m}, radius {radius}, Q={nq:,}, k={k}",
           "value": round(nq / (dev_ms / 1e3), 1), "unit": "queries/s", "ms_per_call_device": round(dev_ms, 4),
           "host_call_qps": round(nq / wall, 1), "source_first": round(found, 4),
           "stages_ms_per_call": {s: round(v, 4) for s, v in st.items()},
           "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_INT32_VALU_TOPS, "unit": "TOP/s",
                        "frac": round(ach / PEAK_INT32_VALU_TOPS, 4),
                        "traffic": measured_traffic("hamming_scan_kernel<16, 16, 1, 4, 2>"),
                        "kernel": "hamming_scan_kernel<16, 16, 1, 4, 2>", "kernel_ms": round(scan, 4),
                        "work": f"{ops:.3e} int32 ops per launch: (16 words x 2 + 8 groups x 2) x Q x N"}}
    _traffic_ratio(out["roofline"], X.nbytes)
    if cpu:
        sys.path.insert(0, ROOT)
        import oracle
        R = oracle.load_ref()
        sample = 4000
        if R is not None:
            t0 = time.perf_counter()
            ref = R.MIHBinaryIndex(m)
            ref.set(X)
            tb = time.perf_counter() - t0
            t0 = time.perf_counter()
            ri, rd = ref.searchKNN(Q[:sample], k, radius)
            t = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(sample / t, 2), "unit": "queries/s", "cores": 1, "kind": "reference",
                                   "sample": (f"{sample} of {nq} queries; reference MIHBinaryIndex.searchKNN (serial), "
                                              f"build {tb:.1f} s excluded")}
            out["cpu_sample_distances_equal"] = bool(all(list(rd[q]) == dist[q, :cnt[q]].tolist() for q in range(sample)))
    return out


def run_batch(steps, warmup, cpu):
    from pynear_amd import _native as N
    from pynear_amd import hipdev, VPTreeL2Index
    from pynear_amd.batch import BatchSearch
    subs = batch_workload()
    _log(f"batch: {len(subs)} indices, {sum(len(X) for X, _, _ in subs):,} rows, "
         f"{sum(len(Q) for _, Q, _ in subs):,} queries")
    idx = []
    for X, Q, k in subs:
        h = VPTreeL2Index()
        h.set(X)
        idx.append(h)
    b = BatchSearch([(h, Q, k) for h, (X, Q, k) in zip(idx, subs)])
    t0 = time.perf_counter()
    b._plan.tune(None, reps=3)  # untimed: the autotuner picks the schedule
    t_tune = time.perf_counter() - t0
    tiles, cfg = b.info()
    stream = hipdev.Stream()
    gpu_ms, wall, _ = _time_device(N, hipdev, lambda: b.run(stream), steps, warmup, stream)
    ms = gpu_ms / steps
    nq_tot = sum(len(Q) for _, Q, _ in subs)
    # the same 64 searches as separate plain device searches, back to back on one stream
    dev = [(hipdev.DeviceArray.from_numpy(Q), hipdev.DeviceArray((len(Q), min(k, len(X))), np.int32),
            hipdev.DeviceArray((len(Q), min(k, len(X))), np.float32)) for X, Q, k in subs]

    def sequential():
        for h, (X, Q, k), (dq, di, dd) in zip(idx, subs, dev):
            h._h.search_device(dq.ptr, len(Q), k, di.ptr, dd.ptr, stream.handle)
    seq_ms, _, _ = _time_device(N, hipdev, sequential, max(3, steps // 4), 1, stream)
    seq_ms /= max(3, steps // 4)
    # results must equal the per-index plain searches
    res = b.results()
    equal = all(np.array_equal(bi, h._h.search(Q, k)[0]) and np.array_equal(bd, h._h.search(Q, k)[1])
                for h, (X, Q, k), (bi, bd) in zip(idx, subs, res))
    flops = sum(2.0 * len(Q) * len(X) * X.shape[1] for X, Q, _ in subs)
    ideal_ms = sum(max(2.0 * len(Q) * len(X) * X.shape[1] / (PEAK_FP32_TFLOPS * 1e12),
                       X.nbytes / (PEAK_HBM_GBS * 1e9)) for X, Q, _ in subs) * 1e3
    out = {"workload": "64 exact L2 indices, N~LogU[1k,200k], d in {64,128,256}, Q~LogU[1,4096], k in {1,10,32}",
           "value": round(nq_tot / (ms / 1e3), 1), "unit": "queries/s", "ms_per_step": round(ms, 4),
           "roofline": {"bound": "mfma", "achieved": round(flops / (ms / 1e3) / 1e12, 3), "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(ideal_ms / ms, 4),
                        "traffic": measured_traffic("batch graph (all kernels of one replay)"),
                        "kernel": "whole batch graph (split, persistent scan, merge, rerank, fallback)",
                        "kernel_ms": round(ms, 4),
                        "work": (f"{flops:.3e} flop (sum 2 Q N d); frac = sum over indices of max(flops/fp32 peak, "
                                 f"fp32 rows/HBM) = {ideal_ms:.4f} ms over the measured time")},
           "schedule": cfg, "scan_tiles": tiles, "tune_s": round(t_tune, 2), "equal_to_plain_searches": bool(equal),
           "plain_searches_ms": round(seq_ms, 4), "queries": nq_tot}
    if cpu:
        sys.path.insert(0, ROOT)
        import oracle
        R = oracle.load_ref()
        if R is not None:
            t_total, q_done = 0.0, 0
            for X, Q, k in subs:
                m = min(len(Q), 16)
                r = R.VPTreeL2Index()
                r.set(X)
                t0 = time.perf_counter()
                r.searchKNN(Q[:m], k)
                t_total += (time.perf_counter() - t0) * len(Q) / m
                q_done += m
            out["cpu_baseline"] = {"value": round(nq_tot / t_total, 2), "unit": "queries/s", "cores": _cores(),
                                   "kind": "reference",
                                   "sample": (f"{q_done} queries (up to 16 per index) on each index's reference "
                                              f"VPTreeL2Index (OpenMP), time scaled to all {nq_tot} queries")}
    return out


def run_learned(steps, warmup, cpu, target=0.95):
    """Learned-probe IVF-L2 against fixed-nprobe IVF with ~fp32 candidate distances (bf16x3), both at
    the smallest setting reaching Recall@10 >= target on held-out queries."""
    from pynear_amd import _native as N
    from pynear_amd import hipdev
    from pynear_amd.learned import LearnedProbe
    k, nlist = 10, 4096
    t0 = time.perf_counter()
    X, QT, QE = learned_workload()
    _log(f"learned: data {X.shape}, train {QT.shape}, test {QE.shape} in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    h = N.IvfL2Handle()
    labels, cent = h.build(X, nlist, 20, 42)
    t_build = time.perf_counter() - t0
    _log(f"learned: ivf built in {t_build:.1f} s")
    t0 = time.perf_counter()
    ex = N.ExactHandle(N.METRIC_L2)
    ex.set(X)
    gt_train = np.concatenate([ex.search(QT[s:s + 20000], k)[0] for s in range(0, len(QT), 20000)])
    gt_test = ex.search(QE, k)[0]
    del ex
    t_gt = time.perf_counter() - t0
    _log(f"learned: exact ground truth for {len(QT) + len(QE):,} queries in {t_gt:.1f} s")
    t0 = time.perf_counter()
    probe = LearnedProbe(X.shape[1], nlist, hidden=256, seed=0)
    hist = probe.fit(QT, labels[gt_train].astype(np.int32), epochs=8, batch=1024, lr=1e-3)
    t_train = time.perf_counter() - t0
    _log(f"learned: trained in {t_train:.1f} s, loss {hist}")
    nq = len(QE)
    dq = hipdev.DeviceArray.from_numpy(QE)
    oi = hipdev.DeviceArray((nq, k), np.int32)
    od = hipdev.DeviceArray((nq, k), np.float32)
    oc = hipdev.DeviceArray((nq,), np.int32)
    stream = hipdev.Stream()

    def recall():
        got = oi.numpy()
        return float(np.mean([len(set(got[i].tolist()) & set(gt_test[i].tolist())) / k for i in range(nq)]))

    def fixed(npb, prec):
        return lambda: h.search_device(dq.ptr, nq, npb, k, oi.ptr, od.ptr, oc.ptr, prec, stream.handle)

    def sweep_fixed(prec):
        for npb in (1, 2, 4, 8, 16, 32, 48, 64, 80, 96, 112, 128, 144, 160, 192, 224, 256):
            fixed(npb, prec)()
            stream.synchronize()
            r = recall()
            if r >= target:
                ms = _time_device(N, hipdev, fixed(npb, prec), steps, warmup, stream)[0] / steps
                return {"n_probe": npb, "recall_at_10": round(r, 4), "ms_per_step": round(ms, 4),
                        "qps": round(nq / (ms / 1e3), 1)}
        return None

    base = sweep_fixed("bf16x3")
    fast = sweep_fixed("bf16")
    fast8 = sweep_fixed("fp8")
    _log(f"learned: fixed n_probe bf16x3 {base}, bf16 {fast}, fp8 {fast8}")
    np_max = 256

    def learned(mass, prec):
        def step():
            probes, counts = probe.select(dq.ptr, nq, mass, 1, np_max, stream)
            h.search_probes_device(dq.ptr, nq, probes.ptr, counts.ptr, np_max, k, oi.ptr, od.ptr, oc.ptr, prec,
                                   stream.handle)
        return step

    by_prec = {}
    for prec in ("fp8", "bf16"):
        for mass in (0.8, 0.85, 0.9, 0.92, 0.94, 0.95, 0.96, 0.97, 0.98, 0.99, 0.995, 0.998, 0.999, 0.9995, 0.9999):
            learned(mass, prec)()
            stream.synchronize()
            r = recall()
            if r >= target:
                probes, counts = probe.select(dq.ptr, nq, mass, 1, np_max)
                hp, hc = probes.numpy(), counts.numpy()
                gms, _, stages = _time_device(N, hipdev, learned(mass, prec), steps, warmup, stream)
                ms = gms / steps
                # the list scan's algorithmic flops: 2 d per (query, row of a selected list)
                sizes = np.bincount(labels, minlength=nlist)
                sel = np.arange(np_max)[None, :] < hc[:, None]
                rows = int(np.where(sel, sizes[np.maximum(hp, 0)] * (hp >= 0), 0).sum())
                scan_ms, scan_n = stages["scan"]
                scan_s = scan_ms / max(scan_n, 1) / 1e3
                peak = PEAK_FP8_TFLOPS if prec == "fp8" else PEAK_BF16_TFLOPS
                tf = 2.0 * X.shape[1] * rows / scan_s / 1e12
                by_prec[prec] = {"precision": prec, "mass": mass, "mean_n_probe": round(float(hc.mean()), 2),
                                 "recall_at_10": round(r, 4), "ms_per_step": round(ms, 4),
                                 "qps": round(nq / (ms / 1e3), 1),
                                 "stages_ms_per_step": {s: round(v[0] / steps, 4) for s, v in stages.items() if v[1]},
                                 "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": peak,
                                              "unit": "TFLOP/s", "frac": round(tf / peak, 4), "traffic": None,
                                              "kernel": "ivfl2_scan_fp8_kernel<8, 3>" if prec == "fp8"
                                              else "ivfl2_scan_kernel<1, 8>",
                                              "kernel_ms": round(scan_s * 1e3, 4),
                                              "work": f"{2.0 * X.shape[1] * rows:.3e} flop (2 d x {rows:,} rows of the "
                                                      f"selected lists)"}}
                break
        _log(f"learned: adaptive {prec} {by_prec.get(prec)}")
    lrn = max(by_prec.values(), key=lambda v: v["qps"]) if by_prec else None
    out = {"workload": (f"learned-probe IVF-L2, Zipf(1.1) GMM N={len(X):,} x d={X.shape[1]}, nlist={nlist}, "
                        f"MLP {X.shape[1]}->256->{nlist} bf16 trained on {len(QT):,} queries, {nq:,} held-out, k={k}; "
                        f"list scan in {lrn['precision'] if lrn else '-'}, fp32 rerank of max(4k, 40)"),
           "value": lrn["qps"] if lrn else None, "unit": "queries/s",
           "ms_per_step": lrn["ms_per_step"] if lrn else None,
           "roofline": lrn["roofline"] if lrn else None,
           "kmeans_iters": 20,  # (the fixed and learned searches share this index; forest.py's set uses 100)
           "learned": lrn, "learned_by_precision": by_prec, "fixed_nprobe_fp32": base, "fixed_nprobe_bf16": fast,
           "fixed_nprobe_fp8": fast8,
           "speedup_vs_fixed_fp32": round(lrn["qps"] / base["qps"], 3) if lrn and base else None,
           "speedup_vs_fixed_bf16": round(lrn["qps"] / fast["qps"], 3) if lrn and fast else None,
           "train": {"epochs": len(hist), "loss": [round(v, 4) for v in hist], "seconds": round(t_train, 2)},
           "setup_s": {"build": round(t_build, 1), "exact_ground_truth": round(t_gt, 1)}}
    if cpu and base:
        sys.path.insert(0, ROOT)
        import oracle
        sizes = np.bincount(labels, minlength=nlist)
        keep = np.flatnonzero(sizes)
        order = np.argsort(labels, kind="stable")
        starts = np.concatenate([[0], np.cumsum(sizes)])
        data = [X[order[starts[c]:starts[c + 1]]] for c in keep]
        ids = [order[starts[c]:starts[c + 1]] for c in keep]
        norms = [(x * x).sum(1) for x in data]
        sample = 50
        t0 = time.perf_counter()
        oracle.ivf_l2_search(cent[keep], data, norms, ids, QE[:sample], base["n_probe"], k)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sample / t, 2), "unit": "queries/s", "cores": _cores(), "kind": "port",
                               "sample": (f"{sample} of {nq} held-out queries; numpy restatement of the reference's "
                                          f"IVFFlatL2Index.searchKNN over the same lists at the fixed "
                                          f"n_probe={base['n_probe']} (the reference has no learned probe)")}
    del X
    return out


# ---------------------------------------------------------------- HBM traffic
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def copy_calibration(nbytes=1 << 30, reps=5):
    """Device-to-device copies of a known size: FETCH_SIZE / WRITE_SIZE of these dispatches give the
    counters' scale (bytes moved per KiB counted)."""
    from pynear_amd import hipdev
    a = hipdev.DeviceArray((nbytes,), np.uint8)
    b = hipdev.DeviceArray((nbytes,), np.uint8)
    for _ in range(reps):
        hipdev._ok(hipdev.hip().hipMemcpy(hipdev._vp(b.ptr), hipdev._vp(a.ptr), nbytes, 3))
    hipdev.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _traffic_ratio(roof, db_bytes):
    """Adds the HBM traffic of a Hamming scan as a multiple of its database bytes: the scan re-reads
    each row block once per query block (scalar loads that mostly hit L2), harmless while the
    kernel is VALU-bound far below the HBM rate."""
    t = roof.get("traffic")
    roof["traffic_over_db_bytes"] = round(t / db_bytes, 1) if t else None
    roof["hbm_rate_tbs"] = round(t / (roof["kernel_ms"] / 1e3) / 1e12, 3) if t else None


# ---------------------------------------------------------------- main
def _tag_cpu_model(node, model):
    """Adds the host CPU model to every cpu_baseline object in the result."""
    if isinstance(node, dict):
        for key, v in node.items():
            if key == "cpu_baseline" and isinstance(v, dict):
                v["cpu_model"] = model
            else:
                _tag_cpu_model(v, model)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ivf",
                    choices=["ivf", "hamming", "l2", "batch", "learned", "ivf_binary", "range", "mih"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="ivf only: skip the exact workloads")
    ap.add_argument("--ivf-probes", default="8,16,32", help="ivf only: n_probe values swept")
    ap.add_argument("--ivf-precisions", default="bf16,fp8", help="ivf only: candidate precisions swept")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 1 GiB (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the search path runs on a single "
              "MI355X and the multi-GPU sharding is a design, not built", file=sys.stderr)
        sys.exit(2)
    cpu = not args.no_cpu_baseline
    if args.workload == "ivf":
        main_r = run_ivf(args.steps, args.warmup, cpu, tuple(int(v) for v in args.ivf_probes.split(",")),
                         tuple(args.ivf_precisions.split(",")))
        dtype = main_r["dtype"]
    elif args.workload == "batch":
        main_r = run_batch(args.steps, args.warmup, cpu)
        dtype = "f32"
    elif args.workload == "learned":
        main_r = run_learned(args.steps, args.warmup, cpu)
        main_r.setdefault("roofline", None)
        dtype = "bf16"
    elif args.workload in ("ivf_binary", "range", "mih"):
        main_r = {"ivf_binary": lambda: run_ivf_binary(args.steps, args.warmup, cpu),
                  "range": lambda: run_hamming_range(cpu), "mih": lambda: run_mih(cpu)}[args.workload]()
        main_r.setdefault("ms_per_step", main_r.get("ms_per_call_device"))
        dtype = "u32"
    else:
        main_r = run_exact(args.workload, args.steps, args.warmup, cpu)
        dtype = "u32" if args.workload == "hamming" else "f32"
    from pynear_amd import _native as N
    result = {
        "metric": METRIC,
        "value": main_r["value"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": main_r.get("steps", args.steps),
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": main_r["workload"],
                   "global_batch": {"l2": 1_000, "batch": main_r.get("queries")}.get(args.workload, 10_000),
                   "parallelism": "single GPU"},
        "roofline": main_r["roofline"],
    }
    for key in ("cpu_baseline", "precision", "recall_at_10", "n_probe", "fp64_rel_err", "stages_ms_per_step", "sweep", "setup_s",
                "cpu_sample_recall_vs_port", "cpu_sample_distances_equal", "schedule", "scan_tiles",
                "equal_to_plain_searches", "plain_searches_ms", "sweep_fp8", "learned", "fixed_nprobe_fp32", "fixed_nprobe_bf16",
                "speedup_vs_fixed_fp32", "speedup_vs_fixed_bf16", "train", "learned_by_precision", "fixed_nprobe_fp8",
                "cpu_sample_results_equal", "host_call_qps", "matches", "source_found", "source_first",
                "stages_ms_per_call", "cpu_sample_sets_equal", "search1NN", "ms_per_call_device"):
        if key in main_r:
            result[key] = main_r[key]
    if args.workload == "ivf" and not args.no_secondary:
        also = {}
        for name, fn in (("exact_hamming", lambda: run_exact("hamming", args.steps, args.warmup, cpu, settle=True)),
                         ("exact_l2", lambda: run_exact("l2", max(args.steps, 50), args.warmup, cpu, settle=True)),
                         ("ivf_binary", lambda: run_ivf_binary(args.steps, args.warmup, cpu)),
                         ("hamming_range", lambda: run_hamming_range(cpu)),
                         ("mih", lambda: run_mih(cpu)),
                         ("batched_exact_l2", lambda: run_batch(args.steps, args.warmup, cpu)),
                         ("learned_probe_ivf", lambda: run_learned(args.steps, args.warmup, cpu))):
            try:  # a secondary workload that fails is reported, not fatal to the headline line
                also[name] = fn()
            except Exception as e:  # noqa: BLE001
                also[name] = {"error": f"{type(e).__name__}: {e}"}
        result["also"] = also
    _tag_cpu_model(result, _cpu_model())
    result["tuning"] = {name: N.tune_get(name) for name in
                        ("hamming.wg_per_cu", "hamming.qpl", "float.wg_per_cu", "l2.mfma", "l2.wg_per_cu",
                         "l2.mfma_terms")}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
