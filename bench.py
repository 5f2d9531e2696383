"""Benchmark: pair-fits/sec of the pairwise Jacquard maximum-likelihood fit on
one MI355X.

Workload (default): 1,000 samples x 100,000 biallelic sites (seeded synthetic
cohort, 20% of samples in relative pairs, 2% inbred founders, 1% missing),
every pair x <= y fitted in FP64 (500,500 fits; the metric counts the 499,500
off-diagonal ones).  One step = one full fit from the device-resident 2-bit
cohort + allele frequencies to Output records in host memory.

Prints one JSON line with the throughput, the roofline of the dominant kernel
(the sweep: fused objective + gradient + Hessian passes, VALU FP64 bound), a
CPU baseline (the oracle's restatement of the reference CPU path in the
reference's loop shape, timed on a bounded sample of the same workload on this
host's cores), every output's error against that reference-order fit
("errors"), the FP32-mixed form on the same workload ("mixed": throughput,
ratio to the FP64 form, errors) and the serving workload ("serving": a
device-resident 10,000 x 100,000 cohort in the FP64 form, saturated
pair-fits/s, p50/p99 request latency at 0.7 load over 100 Poisson requests,
1-query latency).
Other modes:
--serving (device-resident cohort answering query requests, p50/p99),
--soft (likelihood input), --cli (the kestrel binary end to end on the cohort
written as a VCF), --autotune / --graph / --batch (launch shape).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X vendor FP32 vector (packed)
FP64_VALU_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak
LDS_PEAK_TBS = 157.3  # gfx950 LDS reads: 256 B/clk/CU x 256 CUs x 2.4 GHz
FLOP_PER_SITE_EVAL = 156  # 136 (gradient + Hessian) + 20 (objective, log counted as 1)
# HBM bytes per sweep launch on the default workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by a plain 256 MiB copy; counters cannot be
# read inside this process, so the line carries the profiled figure and says so
SWEEP_TRAFFIC_BYTES_C2 = {"sweep_class_fp64": 1.657e9, "sweep_fp64": 1.563e8}
SWEEP_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of {k} on this workload, scaled by a 256 MiB "
                        "copy_kernel (profiles/{k}_c2.md); profiled, not measured in this run")
CLASS_NOTE = ("achieved counts 156 FLOP per site evaluation, the direct per-site form's work (9 coefficients, 9 "
              "gradient and 45 Hessian terms); the class-sorted sweep obtains the same f, g, H from 2-11 sums a site "
              "in each genotype-pair pattern's basis (about 31 VALU instructions a site, VALU issue 0.62 of SIMD "
              "cycles in profiles/sweep_class_fp64_c2.md), so frac can exceed 1")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_flops(out) -> float:
    """W_pair = 156 * n_xy * (iterations + 1); at the iteration cap iterations + 1 = cap."""
    it = out["iters"].astype(np.float64)
    return float((FLOP_PER_SITE_EVAL * out["n_sites"].astype(np.float64) * (it + 1)).sum())


def cpu_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit():
        n = min(n, int(cap))
    return max(1, n)


def cpu_model() -> str:
    """The host CPU's model name (as lscpu reports it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(codes, af, budget_s: float, gl=None, buckets=None):
    """The oracle in the reference's order on a seeded random sample of
    off-diagonal pairs, work stealing over pairs on this host's threads.  Hard
    calls run in the reference's own AVX-512 loop shape (ORDER_FAST: c_s rows
    per 8-site block, getexp/getmant log, three 15-accumulator Hessian passes
    per tile; bit-identical to the 8-lane loops); likelihood rows and
    multi-allelic cohorts in the 8-lane loops (buckets: the 2-, 3-, 4-allele
    likelihood rows, site-major; frequencies and re-bucketing untimed).
    Returns (baseline dict, the sample's pairs, its records)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = cpu_threads()
    parts = None
    if buckets is not None:
        parts, _ = O.multi_prepare({A: b for A, b in zip((2, 3, 4), buckets)})
        n = buckets[0].shape[1]
    else:
        n = codes.shape[0] if gl is None else gl.shape[0]
    fast = parts is None and gl is None and O.lib().ko_has_avx512()
    order = O.ORDER_FAST if fast else O.ORDER_REF

    def fit(pairs):
        if parts is not None:
            return O.multi_fit_pairs_threads(parts, pairs, nthreads=threads)
        if gl is None:
            return O.fit_pairs(codes, af, pairs=pairs, order=order, nthreads=threads)
        return O.fit_pairs_gl(gl, af, pairs=pairs, order=O.ORDER_REF, nthreads=threads)
    rng = np.random.default_rng(12345)

    def sample(m):
        x = rng.integers(0, n, m)
        y = rng.integers(0, n, m)
        y = np.where(y == x, (y + 1) % n, y)
        return np.stack([np.minimum(x, y), np.maximum(x, y)], 1)

    fit(sample(threads))  # warm-up: library load, first-touch of the cohort
    # the sample's size from two probes: a small one sizes a second of about
    # one second, whose rate (thread start-up and uneven pairs amortised) sizes
    # the sample to the budget
    probe = sample(8 * threads)
    for target in (1.0, budget_s):
        t = time.perf_counter()
        fit(probe)
        dt = time.perf_counter() - t
        m = int(max(2 * threads, min(400_000, min(target, budget_s) / max(dt, 1e-3) * len(probe))))
        if target < budget_s:
            probe = sample(m)
    pairs = sample(m)
    t = time.perf_counter()
    recs = fit(pairs)
    dt = time.perf_counter() - t
    shape = ("the reference's AVX-512 loop shape (c_s rows per 8-site block, getexp/getmant log, three "
             "15-accumulator Hessian passes per 32-block tile), bit-identical to the oracle's 8-lane loops"
             if fast else "the oracle's 8-lane loops (one accumulator chain per Hessian entry)")
    return ({"value": m / dt, "unit": "pair-fits/s", "cores": threads, "kind": "port", "isa": O.isa(),
             "cpu_model": cpu_model(), "loop_shape": shape,
             "sample": f"{m} random off-diagonal pairs of the same cohort, {dt:.1f} s"}, pairs, recs)


def record_errors(out, ref, tol_log10_cond=12.0):
    """Per-output errors of records `out` against `ref` on the pairs ref puts
    below log10 condition 12 (both converged): max |dDelta_k|, |d kinship|,
    relative objective, |d log10 condition|."""
    from kestrel_amd import kinship
    lc, lo = np.log10(ref["cond"]), np.log10(out["cond"])
    well = np.isfinite(lc) & (lc < tol_log10_cond) & np.isfinite(lo)
    if not well.any():
        return {"pairs": int(len(out)), "well_conditioned": 0}
    d = np.abs(out["jacquard"] - ref["jacquard"]).max(1)[well]
    return {"pairs": int(len(out)), "well_conditioned": int(well.sum()),
            "max_abs_delta": float(d.max()),
            "max_abs_kinship": float(np.abs(kinship(out["jacquard"]) - kinship(ref["jacquard"]))[well].max()),
            "max_rel_objective": float((np.abs(out["obj"] - ref["obj"]) / np.abs(ref["obj"]))[well].max()),
            "max_abs_log10_cond": float(np.abs(lo[well] - lc[well]).max()),
            "same_iterations": float((out["iters"] == ref["iters"]).mean())}


def tri_index(n, pairs):
    """Record positions of pairs (x <= y) in the row-major upper triangle with the diagonal."""
    x = pairs[:, 0].astype(np.int64)
    y = pairs[:, 1].astype(np.int64)
    return x * n - x * (x - 1) // 2 + (y - x)


def percentile(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


def serving_run(config="c5", precision="fp64", policy="srpt", slots=None, sat_requests=12, requests=100,
                load=0.7, single_probes=20):
    """Serving workload: a device-resident cohort (default 10,000 x 100,000,
    the FP32-mixed form) answers requests of 1-256 new query samples against
    the whole cohort or a random subset of 500-5,000 samples (even odds).
    Phase 1 submits sat_requests requests at once (saturated: pair-fits/s
    capacity); phase 2 submits `requests` fresh requests at Poisson times with
    a mean rate of `load` x the phase-1 capacity (request latency p50/p99),
    with `single_probes` 1-query requests against the whole cohort at uniform
    times in the same window (their latency is reported apart).  Returns the
    summary dict."""
    import queue
    import threading

    from kestrel_amd import Cohort, synth

    t = time.perf_counter()
    n_query_pool = 256
    kw = dict(synth.CONFIGS[config])
    kw["n_samples"] += n_query_pool
    c = synth.generate(**kw)
    n = c.codes.shape[0] - n_query_pool
    v = c.codes.shape[1]
    packed = synth.pack2(c.codes)
    cohort_packed, pool = packed[:n], packed[n:]
    del c
    log(f"cohort {config}: {n} x {v} (+{n_query_pool} query samples) generated in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    co = Cohort(packed=cohort_packed, n_sites=v, precision=precision)
    if slots:
        co.set_option("slots", slots)
    co.set_option("serve_policy", {"rr": 0, "srpt": 1}[policy])
    co.serve_start(max_query_rows=1024, max_pairs=1 << 24)
    st0 = co.serve_stats()
    log(f"device cohort, allele frequencies and engine in {time.perf_counter() - t:.2f}s, kept {co.n_kept} sites, "
        f"{st0['slots']} slots, engine holds {st0['device_bytes'] / 2**30:.2f} GiB")
    rng = np.random.default_rng(0xC5)

    def make_request():
        nq = int(rng.integers(1, 257))
        q0 = int(rng.integers(0, n_query_pool - nq + 1))
        sub = None if rng.random() < 0.5 else np.sort(rng.choice(n, int(rng.integers(500, 5001)), replace=False))
        return pool[q0:q0 + nq], sub

    def run(reqs, arrivals, kinds):
        """Submit reqs at arrival offsets (s), in arrival order; returns (wall s,
        pairs, latencies ms by kind)."""
        lat = {}
        ids_q = queue.Queue()

        def waiter():
            done = 0
            for i, kind in iter(ids_q.get, None):
                out, ms = co.wait(i)
                lat.setdefault(kind, []).append(ms)
                del out
                done += 1
                if done % 10 == 0:  # progress of a phase that runs for minutes
                    log(f"  {done} of {len(reqs)} requests done at {time.perf_counter() - t0:.1f}s")

        t0 = time.perf_counter()
        th = threading.Thread(target=waiter)
        th.start()
        pairs = 0
        for j in np.argsort(arrivals, kind="stable"):
            (q, sub), at = reqs[j], arrivals[j]
            d = t0 + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids_q.put((co.submit(query_packed=q, subset=sub), kinds[j]))
            pairs += q.shape[0] * (n if sub is None else len(sub))
        ids_q.put(None)
        th.join()
        return time.perf_counter() - t0, pairs, lat

    # warm up: one mid-size request, then one 1-query request on the idle engine
    t = time.perf_counter()
    co.wait(co.submit(query_packed=pool[:16]))
    log(f"warmup request 16 x {n}: {time.perf_counter() - t:.2f}s")
    idle = [co.wait(co.submit(query_packed=pool[i:i + 1]))[1] for i in range(5)]
    k = sat_requests
    sat = [make_request() for _ in range(k)]
    dt, pairs, lat_sat = run(sat, [0.0] * k, ["request"] * k)
    capacity = pairs / dt
    log(f"saturated: {k} requests, {pairs} pairs in {dt:.2f}s = {capacity:.0f} pair-fits/s")
    reqs = [make_request() for _ in range(requests)]
    mean_pairs = np.mean([q.shape[0] * (n if s is None else len(s)) for q, s in reqs])
    rate = load * capacity / mean_pairs
    arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
    arrivals -= arrivals[0]
    window = float(arrivals[-1]) if requests > 1 else 1.0
    probes = [(pool[int(i):int(i) + 1], None) for i in rng.integers(0, n_query_pool, single_probes)]
    probe_at = list(np.sort(rng.uniform(0.0, window, single_probes)))
    dt2, pairs2, lat = run(reqs + probes, list(arrivals) + probe_at,
                           ["request"] * requests + ["single_query"] * single_probes)
    st = co.serve_stats()
    occupancy = st["busy_slot_sweeps"] / max(1, st["sweeps"]) / st["slots"]
    n_kept = co.n_kept
    lr, l1 = lat.get("request", []), lat.get("single_query", [])
    log(f"poisson: {requests} requests at {rate:.3f}/s (+{single_probes} 1-query), {pairs2} pairs in {dt2:.2f}s, "
        f"p50 {percentile(lr, 50):.0f} ms p99 {percentile(lr, 99):.0f} ms, 1-query p50 {percentile(l1, 50):.0f} ms")
    co.close()

    def r1(x):
        return None if x is None else round(x, 1)
    return {
        "workload": f"{config} serving: cohort {n} samples x {v} sites ({n_kept} segregating) device-resident, "
                    f"{precision}; requests of 1-256 queries vs all or 500-5000 samples",
        "samples": n, "sites": v, "precision": precision, "policy": policy, "slots": st0["slots"],
        "saturated": {"requests": k, "pairs": int(pairs), "seconds": round(dt, 3),
                      "pair_fits_per_s": round(capacity, 1),
                      "latency_ms_p50": r1(percentile(lat_sat.get("request", []), 50)),
                      "latency_ms_p99": r1(percentile(lat_sat.get("request", []), 99))},
        "poisson": {"requests": requests, "load": load, "rate_per_s": round(rate, 4), "pairs": int(pairs2),
                    "seconds": round(dt2, 3), "pair_fits_per_s": round(pairs2 / dt2, 1),
                    "latency_ms_p50": r1(percentile(lr, 50)), "latency_ms_p99": r1(percentile(lr, 99)),
                    "latency_ms_max": r1(max(lr) if lr else None),
                    "p99_is_max": len(lr) < 100},
        "single_query": {"requests_under_load": single_probes, "latency_ms_p50": r1(percentile(l1, 50)),
                         "latency_ms_p99": r1(percentile(l1, 99)),
                         "idle_latency_ms_p50": r1(percentile(idle, 50)),
                         "pairs_each": n},
        "slot_occupancy": round(occupancy, 4),
    }


def serving(args):
    """The serving workload alone as the JSON line (--serving)."""
    res = serving_run(args.config, args.precision, args.policy, args.slots, args.requests, args.poisson_requests,
                      args.load, args.single_probes)
    k = res["saturated"]["requests"]
    line = {
        "metric": "pair-fits/sec (serving)", "value": res["saturated"]["pair_fits_per_s"], "unit": "pair-fits/s",
        "n_gpus": 1, "steps": k + res["poisson"]["requests"], "warmup": 1,
        "ms_per_step": round(res["saturated"]["seconds"] / k * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic", "config": {"workload": res["workload"], "samples": res["samples"],
                                        "sites": res["sites"], "precision": args.precision, "policy": args.policy},
        **{k2: res[k2] for k2 in ("saturated", "poisson", "single_query", "slot_occupancy")},
    }
    print(json.dumps(line), flush=True)


def cli_end_to_end(args):
    """The drop-in CLI on the workload's cohort written as a VCF: one-hot GL
    records (the ML path, -f GL), BGZF, in a temporary directory; times the
    kestrel binary (parse with the reader's threads, device frequencies and
    fit, TSV write) and reports its phase line."""
    import re
    import shutil
    import subprocess
    import tempfile

    from kestrel_amd import synth

    c = synth.config(args.config)
    n, v = c.codes.shape
    root = os.path.dirname(os.path.abspath(__file__))
    tmp = tempfile.mkdtemp(prefix="kestrel_cli_")
    try:
        path = os.path.join(tmp, "cohort.vcf.gz")
        t = time.perf_counter()
        table = np.array(["0,-1000,-1000", "-1000,0,-1000", "-1000,-1000,0", ".,.,."])
        with synth._BgzfWriter(path) as f:
            f.write("##fileformat=VCFv4.2\n")
            f.write("#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(c.names) + "\n")
            buf = []
            for s in range(v):
                buf.append(f"1\t{s + 1}\t.\tA\tC\t.\tPASS\t.\tGL\t" + "\t".join(table[c.codes[:, s]]) + "\n")
                if len(buf) == 2000:
                    f.write("".join(buf))
                    buf = []
                if (s + 1) % 25000 == 0:
                    log(f"VCF: {s + 1} of {v} sites written ({time.perf_counter() - t:.1f}s)")
            f.write("".join(buf))
        vcf_bytes = os.path.getsize(path)
        tsv = os.path.join(tmp, "out.tsv")
        threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
        t = time.perf_counter()
        r = subprocess.run([os.path.join(root, "kestrel_amd", "kestrel"), path, tsv, "-f", "GL", "-t", str(threads)],
                           capture_output=True, text=True)
        wall = time.perf_counter() - t
        if r.returncode:
            log(r.stderr)
            sys.exit(r.returncode)
        m = re.search(r"Phases: parse ([0-9.]+) s, device setup and fit ([0-9.]+) s, write ([0-9.]+) s", r.stderr)
        rows = sum(1 for _ in open(tsv)) - 1
        line = {"metric": "end-to-end CLI seconds (VCF in, TSV out)", "value": round(wall, 3), "unit": "s",
                "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                "config": {"workload": f"{args.config}: {n} samples x {v} sites as a one-hot GL VCF (BGZF), "
                                       f"kestrel -f GL -t {threads}", "vcf_bytes": vcf_bytes, "pairs": rows},
                "phases_s": ({"parse": float(m.group(1)), "frequencies_and_fit": float(m.group(2)),
                              "write": float(m.group(3))} if m else None),
                "pair_fits_per_s_end_to_end": round((rows - n) / wall, 1)}
        print(json.dumps(line), flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cls_bench(args):
    """The CLS / Brier estimator of -f GT (kestrel_fit_pairs_gt) on the
    workload's sizes without missing calls (the GT path drops every site with
    a missing call, so the workload's 1 % missingness would leave no site).
    One step = one fit of every pair from the device-resident cohort; the
    9 x 9 Gram matrix is built once per cohort (in the warmup)."""

    from kestrel_amd import Cohort, synth

    t = time.perf_counter()
    c = synth.config(args.config, missing=0.0)
    n, v = c.codes.shape
    log(f"cohort {args.config} without missing calls: {n} x {v} in {time.perf_counter() - t:.1f}s")
    co = Cohort.gt(c.codes)
    for i in range(args.warmup):
        t = time.perf_counter()
        co.fit_pairs_gt()
        log(f"warmup {i}: {time.perf_counter() - t:.3f}s")
    times, kms = [], 0.0
    out = None
    for i in range(args.steps):
        t = time.perf_counter()
        out, _ = co.fit_pairs_gt()
        times.append(time.perf_counter() - t)
        kms += co.last_fit_stats()["sweep_ms"]
        log(f"step {i}: {times[-1]:.3f}s, per-pair kernels {co.last_fit_stats()['sweep_ms']:.1f} ms")
    total = float(np.sum(times))
    offdiag = int((out["x"] != out["y"]).sum())
    value = offdiag * args.steps / total
    # c_xy = sum over sites of m_s[gx][gy]: 9 FP64 adds and a 72-byte LDS row per pair and site
    pair_sites = float(len(out)) * v * args.steps
    lds_tbs = pair_sites * 72 / (kms * 1e-3) / 1e12
    import ctypes

    from kestrel_amd import _lib
    measured = ctypes.c_double()
    _lib.check(_lib.load().kestrel_measure_lds(ctypes.byref(measured)))
    cpu = None
    if args.cpu_budget > 0:  # the oracle's CLS (reference order) on a sample of pairs
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        rng = np.random.default_rng(0xC15)
        threads = cpu_threads()

        def sample(m):
            p = np.stack([rng.integers(0, n, m), rng.integers(0, n, m)], 1).astype(np.uint32)
            return p[p[:, 0] != p[:, 1]]
        probe = sample(4 * threads)
        t = time.perf_counter()
        O.cls_fit_pairs(c.codes, pairs=probe, nthreads=threads)
        dt = time.perf_counter() - t
        pairs = sample(int(max(2 * threads, min(200_000, args.cpu_budget / max(dt, 1e-3) * len(probe)))))
        t = time.perf_counter()
        O.cls_fit_pairs(c.codes, pairs=pairs, nthreads=threads)
        dt = time.perf_counter() - t
        cpu = {"value": len(pairs) / dt, "unit": "pair-fits/s", "cores": threads, "kind": "port",
               "sample": f"{len(pairs)} random off-diagonal pairs of the same cohort, {dt:.1f} s"}
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "pair-fits/sec (CLS, -f GT)", "value": round(value, 1), "unit": "pair-fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(total / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config} without missing calls: {n} samples x {v} sites, {offdiag} "
                               f"off-diagonal pairs (+{len(out) - offdiag} self pairs), CLS / Brier estimator",
                   "samples": n, "sites": v, "pairs": offdiag},
        "roofline": {"bound": "lds", "achieved": round(lds_tbs, 3), "peak": LDS_PEAK_TBS, "unit": "TB/s",
                     "frac": round(lds_tbs / LDS_PEAK_TBS, 4), "traffic": None, "kernel": "cls_sum_kernel",
                     "kernel_ms_per_step": round(kms / args.steps, 2), "measured_peak": round(measured.value, 2),
                     "frac_of_measured": round(lds_tbs / measured.value, 4),
                     "note": "LDS bytes: one 72-byte m row per pair and site (kernel time: the sum and QP "
                             "kernels); peak 256 B/clk/CU x 256 CUs x 2.4 GHz; "
                             "measured: conflict-free 16-byte reads (kestrel_measure_lds)"},
        "cpu_baseline": cpu, "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    co.close()
    print(json.dumps(line), flush=True)


def mixed_errors(codes, co_mixed, config, n_pairs=20000, seed=0xE44, gl=None, buckets=None):
    """Errors of the mixed form against the FP64 form on a seeded sample of
    off-diagonal pairs, by the bar of the FP32-mixed form: max |dDelta_k|,
    |d kinship|, relative objective and log10 condition on the pairs whose
    FP64 log10 condition is below 12 (and max |dDelta_k| below 6)."""
    from kestrel_amd import Cohort, kinship
    rng = np.random.default_rng(seed)
    n = codes.shape[0]
    x = rng.integers(0, n, n_pairs)
    y = rng.integers(0, n, n_pairs)
    keep = x != y
    pairs = np.stack([np.minimum(x, y), np.maximum(x, y)], 1)[keep]
    a = co_mixed.fit_pairs(pairs)
    with (Cohort.gl_multi(*buckets) if buckets else Cohort(codes) if gl is None else Cohort.gl(gl)) as co64:
        b = co64.fit_pairs(pairs)
    lc = np.log10(b["cond"])
    lca = np.log10(a["cond"])
    capped = ~np.isfinite(lca)  # the mixed form reached the iteration cap (condition NaN)
    well = np.isfinite(lc) & (lc < 12) & ~capped
    good = well & (lc < 6)
    d = np.abs(a["jacquard"] - b["jacquard"]).max(1)
    return {"pairs": int(len(pairs)), "well_conditioned": int(well.sum()),
            "mixed_at_iteration_cap": int(capped.sum()), "fp64_at_iteration_cap": int((~np.isfinite(lc)).sum()),
            "max_abs_delta": float(d[well].max()), "max_abs_delta_cond_below_1e6": float(d[good].max()),
            "pairs_over_1e-6": int((d[well] > 1e-6).sum()),
            "max_abs_kinship": float(np.abs(kinship(a["jacquard"]) - kinship(b["jacquard"]))[well].max()),
            "max_rel_objective": float((np.abs(a["obj"] - b["obj"]) / np.abs(b["obj"]))[well].max()),
            "max_abs_log10_cond": float(np.abs(lca[well] - lc[well]).max()),
            "iterations_mean": [round(float(a["iters"].mean()), 3), round(float(b["iters"].mean()), 3)],
            "reference": "the FP64 form of this library (bit-exact with the oracle in its summation order)"}


def timed_fits(co, steps, warmup):
    """warmup untimed fits, then `steps` timed fits of every pair: the
    throughput, the algorithmic rate over the time with a sweep running, and
    the sweep launch statistics."""
    for i in range(warmup):
        t = time.perf_counter()
        co.fit_pairs()
        log(f"warmup {i}: {time.perf_counter() - t:.3f}s")
    times, flops, sweep_ms, sweeps, busy_ms, groups, rec_ms, rec_pairs = [], 0.0, 0.0, 0, 0.0, 1, 0.0, 0
    out, form = None, None
    for i in range(steps):
        t = time.perf_counter()
        out = co.fit_pairs()
        dt = time.perf_counter() - t
        st = co.last_fit_stats()
        times.append(dt)
        flops += algorithmic_flops(out)
        sweep_ms += st["sweep_ms"]
        sweeps += st["sweeps"]
        busy_ms += st["busy_ms"]
        groups = st["groups"]
        rec_ms += st["recondition_ms"]
        rec_pairs += st["recondition_pairs"]
        form = st["form"]
        log(f"step {i}: {dt:.3f}s, {st['sweeps']} sweeps in {st['groups']} slot groups, sweep launches "
            f"{st['sweep_ms']:.1f} ms summed, {st['busy_ms']:.1f} ms with a sweep running"
            + (f", FP64 condition pass {st['recondition_ms']:.1f} ms for {st['recondition_pairs']} pairs"
               if st["recondition_pairs"] else ""))
    total = float(np.sum(times))
    offdiag = int((out["x"] != out["y"]).sum())
    # the slot groups' launches overlap: the device rate is the algorithmic
    # work over the time with at least one sweep running
    return {"out": out, "seconds": total, "value": offdiag * steps / total, "form": form,
            "achieved": flops / (busy_ms / 1e3) / 1e12,
            "recondition": {"ms_per_step": round(rec_ms / steps, 2), "pairs_per_step": rec_pairs // max(1, steps)},
            "launch": {"avg_launch_ms": round(sweep_ms / sweeps, 4), "concurrent_launches": groups,
                       "sweep_busy_ms_per_step": round(busy_ms / steps, 2),
                       "algorithmic_flop_per_launch": flops / sweeps}}


def hessian_forms_bench(args):
    """The per-site Hessian accumulation in its vector and matrix-core forms
    (--hessian-forms): time per pair and site of each form on the same
    synthetic FP32 rows, the Hessian's share (each form against the same form
    without it) and each form's error against FP64 sums of the same rows."""
    import ctypes
    from kestrel_amd import _lib
    L = _lib.load()
    n_pairs, n_sites = 30720, args.hessian_sites
    rng = np.random.default_rng(0x4E55)
    table = np.zeros((256, 12), np.float32)
    table[:, :9] = rng.uniform(0.0, 1.0, (256, 9)).astype(np.float32)
    table[:, 9] = (1.0 / rng.uniform(0.01, 1.0, 256)).astype(np.float32)  # 1/L of a site
    names = {0: "lane_fp64", 1: "lane_fp32", 2: "mfma_bf16x2", 3: "mfma_bf16x3", 4: "lane_fp32_no_hessian",
             5: "sites_across_lanes_no_hessian", 6: "sites_across_lanes_fp32"}
    # FP64 sums of the FP32 rows of the first 256 pairs (every per-pair scale)
    w = (np.float32(1.0) + (np.arange(256) % 64).astype(np.float32) / np.float32(64.0)).astype(np.float32)
    r = (table[None, :, 9] * w[:, None]).astype(np.float32)            # 256 pairs x 256 table sites
    cp = (table[None, :, :9] * r[:, :, None]).astype(np.float32)      # c' rows
    reps = n_sites // 256
    href = np.einsum("psi,psj->pij", cp.astype(np.float64), cp.astype(np.float64)) * reps
    gref = cp.astype(np.float64).sum(1) * reps
    il, jl = np.tril_indices(9)
    order = np.lexsort((jl, il))  # row by row: (0,0), (1,0), (1,1), ...
    il, jl = il[order], jl[order]
    scale = np.sqrt(href[:, il, il] * href[:, jl, jl])
    res = {}
    for form in range(7):
        nseg = n_sites // 4096
        out = np.zeros((nseg, n_pairs, 54), np.float64 if form == 0 else np.float32)
        ms = ctypes.c_double()
        _lib.check(L.kestrel_measure_hessian_form(form, n_pairs, n_sites, table.ctypes.data_as(ctypes.c_void_p),
                                                  out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms)))
        e = {"ms": round(ms.value, 4), "ns_per_pair_site": round(ms.value * 1e6 / (n_pairs * n_sites), 5)}
        tot = out[:, :256].astype(np.float64).sum(0)  # segments folded in FP64, as the sweeps' fold
        if form not in (4, 5):
            e["max_rel_error_hessian"] = float(np.max(np.abs(tot[:, :45] - href[:, il, jl]) / scale))
        e["max_rel_error_gradient"] = float(np.max(np.abs(tot[:, 45:] - gref) / gref))
        res[names[form]] = e
        log(f"form {form} {names[form]}: {e}")
    lane_h = res["lane_fp32"]["ns_per_pair_site"] - res["lane_fp32_no_hessian"]["ns_per_pair_site"]
    line = {"metric": "hessian accumulation ns per pair and site", "n_pairs": n_pairs, "n_sites": n_sites,
            "forms": res,
            "hessian_share_ns": {"lane_fp32": round(lane_h, 5)}}
    base = res["sites_across_lanes_no_hessian"]["ns_per_pair_site"]
    for k in ("mfma_bf16x2", "mfma_bf16x3", "sites_across_lanes_fp32"):
        line["hessian_share_ns"][k] = round(res[k]["ns_per_pair_site"] - base, 5)
    print(json.dumps(line), flush=True)


def small_jobs(co, n, reps=5):
    """Fits of few pairs (one query against part or all of the cohort, a few
    queries against all) in both layouts: one pair per lane, and one pair per
    wavefront with its sites across the lanes (option "pair_lanes" 64).  Wall
    milliseconds per fit (median of reps, after one untimed fit) and the
    wavefront form's errors against the per-lane form."""
    rng = np.random.default_rng(0x5A11)
    jobs = []
    for q, m in ((1, 100), (1, n - 1), (4, n - 1), (32, n - 1)):
        queries = rng.choice(n, q, replace=False)
        pairs = []
        for x in queries:
            others = rng.choice(np.setdiff1d(np.arange(n), [x]), m, replace=False)
            pairs.append(np.stack([np.full(m, x), others], 1))
        jobs.append((f"{q} x {m}", np.concatenate(pairs).astype(np.uint32)))
    res = []
    for name, pairs in jobs:
        row = {"job": name, "pairs": int(len(pairs))}
        outs = {}
        for lanes in (1, 64):
            co.set_option("pair_lanes", lanes)
            co.fit_pairs(pairs=pairs)
            ts = []
            for _ in range(reps):
                t = time.perf_counter()
                outs[lanes] = co.fit_pairs(pairs=pairs)
                ts.append((time.perf_counter() - t) * 1e3)
            row[f"ms_pair_lanes_{lanes}"] = round(float(np.median(ts)), 2)
        co.set_option("pair_lanes", 1)
        row["speedup"] = round(row["ms_pair_lanes_1"] / row["ms_pair_lanes_64"], 2)
        row["errors_vs_per_lane"] = record_errors(outs[64], outs[1])
        log(f"small job {row}")
        res.append(row)
    return res


def measured_peak(mixed):
    """FP32 (mixed) or FP64 FMA-chain rate of this device in TFLOP/s, or None."""
    try:
        import ctypes
        from kestrel_amd import _lib
        pk = ctypes.c_double()
        _lib.check((_lib.load().kestrel_measure_fp32_peak if mixed else _lib.load().kestrel_measure_fp64_peak)(
            ctypes.byref(pk)))
        return pk.value
    except Exception as e:  # the measurement is informative only
        log(f"peak measurement failed: {e}")
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work (0 = skip)")
    ap.add_argument("--segment-sites", type=int, default=None)
    ap.add_argument("--slots", type=int, default=None)
    ap.add_argument("--streams", type=int, default=None, help="slot groups on separate streams (1 or 2)")
    ap.add_argument("--batch", type=int, default=None, help="iterations queued per host synchronisation")
    ap.add_argument("--graph", action="store_true", help="relaunch each batch from a captured HIP graph")
    ap.add_argument("--pair-tile", type=int, default=None, help="claim order in tiles of this many samples")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="mixed: per-site terms in FP32 (KESTREL_MIXED), with its errors against the FP64 form")
    ap.add_argument("--skip-masked", type=int, choices=(-1, 0, 1), default=None,
                    help="claims in runs + sweeps passing over masked sites (default: the library's automatic choice)")
    ap.add_argument("--class-sweep", type=int, choices=(0, 1), default=None,
                    help="FP64 fits of hard calls by the class-sorted sweep (1) or the site-table sweep (0)")
    ap.add_argument("--autotune", type=int, default=0, metavar="PAIRS",
                    help="before the warmup, tune slots/streams/batch on this many random pairs (0 = off)")
    ap.add_argument("--serving", action="store_true", help="serving workload (default config c5)")
    ap.add_argument("--soft", action="store_true", help="soft genotype likelihoods instead of hard calls")
    ap.add_argument("--depth", type=float, default=4.0, help="mean read depth of --soft likelihoods")
    ap.add_argument("--requests", type=int, default=12, help="serving: requests of the saturated phase")
    ap.add_argument("--poisson-requests", type=int, default=100, help="serving: requests of the Poisson phase")
    ap.add_argument("--single-probes", type=int, default=20,
                    help="serving: 1-query requests at uniform times in the Poisson phase")
    ap.add_argument("--serving-precision", choices=("fp64", "mixed"), default="fp64",
                    help="form of the serving workload: FP64 (the class-sorted sweep, faster on this device than "
                         "the FP32-mixed kernels the configuration names) or mixed")
    ap.add_argument("--no-mixed-leg", action="store_true",
                    help="skip the FP32-mixed fit of the same workload in the default run")
    ap.add_argument("--no-small-jobs", action="store_true",
                    help="skip the few-pair fits in both layouts (one pair per lane / per wavefront)")
    ap.add_argument("--no-serving-leg", action="store_true",
                    help="skip the serving workload (C5) in the default run")
    ap.add_argument("--load", type=float, default=0.7, help="Poisson offered load / saturated capacity")
    ap.add_argument("--policy", choices=("srpt", "rr"), default="srpt",
                    help="serving feed order: shortest remaining first or round robin")
    ap.add_argument("--cli", action="store_true", help="end-to-end kestrel CLI run on the workload as a VCF")
    ap.add_argument("--cls", action="store_true", help="the CLS / Brier estimator of -f GT")
    ap.add_argument("--multi", type=float, default=0.0, metavar="FRACTION",
                    help="this fraction of the sites with 3 or 4 alleles (4:1), as one-hot likelihood rows")
    ap.add_argument("--hessian-forms", action="store_true",
                    help="time the per-site Hessian in its vector and matrix-core forms (synthetic rows)")
    ap.add_argument("--hessian-sites", type=int, default=32768, help="sites per pair of --hessian-forms")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if args.serving:
        if args.config == "c2":
            args.config = "c5"
        args.precision = args.serving_precision
        serving(args)
        return
    if args.cli:
        cli_end_to_end(args)
        return
    if args.cls:
        cls_bench(args)
        return
    if args.hessian_forms:
        hessian_forms_bench(args)
        return

    from kestrel_amd import Cohort, synth

    t = time.perf_counter()
    c = synth.config(args.config)
    n, v = c.codes.shape
    gl = None
    buckets = None
    if args.soft:
        gl = synth.soft_likelihoods(c.codes, seed=0x50F7, mean_depth=args.depth)
    if args.multi > 0:
        # the workload's sizes with a fraction of the sites tri/tetra-allelic
        # (hard calls of the multi-allelic generator, 1 % missing) as one-hot
        # likelihood rows: (V_A, N, A(A+1)/2), site-major as vcf.rs builds them
        v3 = int(round(v * args.multi * 0.8))
        v4 = int(round(v * args.multi * 0.2))
        idx = synth.multi_genotypes(n, (v - v3 - v4, v3, v4), seed=0xA4, missing=0.01)
        buckets = list(synth.multi_onehot(idx))
    log(f"cohort {args.config}: {n} x {v}{' soft likelihoods' if args.soft else ''}"
        f"{f' with {args.multi:.0%} tri/tetra-allelic sites' if buckets else ''} generated in "
        f"{time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    co = (Cohort.gl(gl, precision=args.precision) if args.soft
          else Cohort.gl_multi(*buckets, precision=args.precision) if buckets
          else Cohort(c.codes, precision=args.precision))
    if args.segment_sites:
        co.set_option("segment_sites", args.segment_sites)
    if args.slots:
        co.set_option("slots", args.slots)
    if args.streams:
        co.set_option("streams", args.streams)
    if args.batch:
        co.set_option("batch", args.batch)
    if args.graph:
        co.set_option("graph", 1)
    if args.pair_tile is not None:
        co.set_option("pair_tile", args.pair_tile)
    if args.skip_masked is not None:
        co.set_option("skip_masked", args.skip_masked)
    if args.class_sweep is not None:
        co.set_option("class_sweep", args.class_sweep)
    if buckets:
        parts, _ = co.multi_info()
        kept = None
        log(f"device cohort + allele frequencies in {time.perf_counter() - t:.2f}s, sites with 2 / 3 / 4 alleles "
            f"after zero-frequency removal: {parts}")
    else:
        af, kept = co.allele_freqs()
        log(f"device cohort + allele frequencies in {time.perf_counter() - t:.2f}s, kept {co.n_kept} sites")
    tuned = None
    if args.autotune:
        t = time.perf_counter()
        tuned = co.autotune(sample=args.autotune)
        log(f"autotune ({time.perf_counter() - t:.1f}s, untimed): {tuned}")

    fit64 = timed_fits(co, args.steps, args.warmup)
    out, value, total, achieved = fit64["out"], fit64["value"], fit64["seconds"], fit64["achieved"]
    offdiag = int((out["x"] != out["y"]).sum())
    # FP64 fits of hard calls run the class-sorted sweep unless --class-sweep
    # 0 (the --multi rows are one-hot: their 2-allele part is held as 2-bit
    # codes and takes it too when no called row lost its 1.0 to an allele
    # below the minimum frequency); the library reports the form it ran
    class_form = fit64["form"] == "class"
    main = ("sweep_mixed" if args.precision == "mixed" else "sweep_class_fp64" if class_form else "sweep_fp64")
    kernel = ((main + (" + sweep_glm_mixed" if args.precision == "mixed" else " + sweep_glm")) if buckets
              else ("sweep_gl_mixed" if args.precision == "mixed" else "sweep_gl") if args.soft
              else main)
    profiled = args.config == "c2" and not args.segment_sites and kernel in SWEEP_TRAFFIC_BYTES_C2
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_VALU_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_VALU_PEAK_TFLOPS, 4),
                "traffic": SWEEP_TRAFFIC_BYTES_C2[kernel] if profiled else None,
                "traffic_source": SWEEP_TRAFFIC_SOURCE.format(k=kernel) if profiled else None,
                "kernel": kernel, **fit64["launch"]}
    if class_form:
        roofline["note"] = CLASS_NOTE
    mixed = args.precision == "mixed"
    if mixed:  # the mixed form is priced against FP32 (vendor figure: packed FP32)
        roofline["peak"] = FP32_VALU_PEAK_TFLOPS
        roofline["frac"] = round(achieved / FP32_VALU_PEAK_TFLOPS, 4)
    pk = measured_peak(mixed)
    if pk:
        roofline["measured_peak"] = round(pk, 2)
        roofline["frac_of_measured"] = round(achieved / pk, 4)
    precision_error = None
    if mixed:
        precision_error = mixed_errors(c.codes, co, args.config, gl=gl, buckets=buckets)
        log(f"mixed vs FP64 form: {precision_error}")
    # the FP32-mixed form on the same workload, beside the FP64 form
    mixed_leg = None
    if not mixed and not args.soft and not buckets and not args.no_mixed_leg:
        from kestrel_amd import _lib
        co.set_option("precision", _lib.MIXED)
        fm = timed_fits(co, args.steps, min(args.warmup, 1))
        co.set_option("precision", _lib.FP64)
        pkm = measured_peak(True)
        mixed_leg = {
            "value": round(fm["value"], 1), "unit": "pair-fits/s", "ms_per_step": round(fm["seconds"] / args.steps * 1e3, 2),
            "vs_fp64_form": round(fm["value"] / value, 4),
            "forms": {"mixed": "sweep_mixed (site-table sweep, FP32 Hessian terms)", "fp64": kernel},
            "roofline": {"bound": "valu", "achieved": round(fm["achieved"], 3), "peak": FP32_VALU_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(fm["achieved"] / FP32_VALU_PEAK_TFLOPS, 4),
                         "kernel": "sweep_mixed", **fm["launch"],
                         **({"measured_peak": round(pkm, 2), "frac_of_measured": round(fm["achieved"] / pkm, 4)}
                            if pkm else {})},
            "fp64_condition_pass": fm["recondition"],
            "iterations_mean": round(float(fm["out"]["iters"].mean()), 3),
            "errors_vs_fp64_form": record_errors(fm["out"], out),
        }
        log(f"mixed form: {mixed_leg}")
    small = None
    if not mixed and not args.soft and not buckets and not args.no_small_jobs:
        small = small_jobs(co, n)
    cpu, errors = None, None
    if buckets and args.cpu_budget > 0:
        cpu, _, _ = cpu_baseline(None, None, args.cpu_budget, buckets=buckets)
        log(f"cpu baseline: {cpu}")
    elif args.cpu_budget > 0:
        glk = None if gl is None else np.ascontiguousarray(gl.transpose(1, 0, 2)[:, kept])
        cpu, cpairs, crecs = cpu_baseline(c.codes[:, kept], af[kept], args.cpu_budget, gl=glk)
        log(f"cpu baseline: {cpu}")
        # every output's error against the oracle in the reference's order on
        # the pairs the CPU leg fitted
        idx = tri_index(n, cpairs)
        errors = {"reference": "the oracle in the reference's 8-lane order (" + cpu["sample"] + ")",
                  **record_errors(out[idx], crecs)}
        log(f"errors vs reference order: {errors}")
        if mixed_leg:
            mixed_leg["errors_vs_reference_order"] = record_errors(fm["out"][idx], crecs)
    serving_leg = None
    if not args.no_serving_leg and args.config == "c2" and not args.soft and not buckets and not mixed:
        serving_leg = serving_run("c5", args.serving_precision, args.policy, None, args.requests,
                                  args.poisson_requests, args.load, args.single_probes)
    line = {
        "metric": "pair-fits/sec", "value": round(value, 1), "unit": "pair-fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(total / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {n} samples x {v} sites ({co.n_kept} segregating"
                               + (", 2-allele part" if buckets else "") + f"), "
                               f"{offdiag} off-diagonal pairs (+{len(out) - offdiag} self pairs), "
                               + ("FP32-mixed" if mixed else "FP64")
                               + (f", soft likelihoods (mean depth {args.depth})" if args.soft else "")
                               + (f", {args.multi:.0%} of sites tri/tetra-allelic" if buckets else ""),
                   "samples": n, "sites": v, "pairs": offdiag, "precision": args.precision,
                   "input": ("soft genotype likelihoods" if args.soft else
                             "one-hot likelihoods, 2-4 alleles" if buckets else "hard calls"),
                   "launch_shape": ({k: tuned[k] for k in ("slots", "streams", "batch")} if tuned else "default"),
                   "hip_graph": bool(args.graph),
                   "skip_masked": "auto" if args.skip_masked is None else args.skip_masked},
        "roofline": roofline, "cpu_baseline": cpu,
        **({"precision_error": precision_error} if mixed else {}),
        **({"errors": errors} if errors else {}),
        **({"mixed": mixed_leg} if mixed_leg else {}),
        **({"serving": serving_leg} if serving_leg else {}),
        **({"small_jobs": small} if small else {}),
        "iterations_mean": round(float(out["iters"].mean()), 3),
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    co.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
