"""Benchmark of the DEM compile hot path on one MI355X.

Workload (a "step"): one device-resident compile of BB [[144,12,12]], 12
rounds, circuit-level depolarizing p=1e-3, correlation level L2 -- the
reference's data/circuits/bb12.stim -- with its STEPG already uploaded to HBM.
The timed region is K back-to-back compiles bracketed by HIP events on the
pipeline's own stream.  `value` is error (fault) locations compiled per
second: 43,632 per compile x K / elapsed.

Also reported: per-stage device time, the roofline of the dominant kernel,
the reference-comparable compile time (STEPG in host memory -> DEM arrays on
host, the span the reference times at Cli.hpp:293-302) and a CPU baseline
(the oracle's forward Pauli-frame compiler, single thread, on this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Other workloads (--workload, one JSON line each): batch (256 uneven
variants batched vs one by one), grad (forward + backward), serving
(Poisson request streams), long (40-100-round memories, whole vs
time-chunked sweep).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
REFERENCE_BB12_L2_US = 5902.7  # ae/results/arxiv060826/single_level_compile.csv:8 (sm_120)
METRIC = "error-locations/s & DEM compile µs, batched vs single on 1 GPU; % HBM peak"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_circuit(name: str) -> str:
    with open(os.path.join(ROOT, "greenpeas_amd", "data", "circuits", f"{name}.stim")) as f:
        return f.read()


def algorithmic_bytes(N: int, Wd: int, Nprime: int, C: int, nnz_bar: float) -> dict:
    """Per-compile algorithmic byte model of each stage (dense 64-bit
    sensitivity words, fixed 32-slot classes, 8-bit LSD sort digits)."""
    return {
        "sweep": 8 * N + 8 * N * Wd,
        "classify": 8 * N * Wd + 128 * N + 12 * N,
        "sort": 8 * 24 * Nprime,
        "reduce_emit": 20 * Nprime + C * (8 + 4 * nnz_bar + 12),
    }


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "bb12_plain_compile_hbm_traffic.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes over this workload, each scaled
    by the counter's reading on a copy of known size); None without one."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
    except OSError:
        return None
    start = text.find("```json")
    if start < 0:
        return None
    end = text.find("```", start + 7)
    try:
        data = json.loads(text[start + 7:end])
        return data["kernels"][kernel]["hbm_bytes_per_launch"]
    except (ValueError, KeyError):
        return None


def cpu_baseline(text: str, locations: int, budget_s: float = 8.0) -> dict:
    from oracle import oracle as O

    O.compile_frames(text)  # warm
    runs, t0 = 0, time.perf_counter()
    while True:
        O.compile_frames(text)
        runs += 1
        el = time.perf_counter() - t0
        if el >= budget_s or runs >= 200:
            break
    return {
        "value": locations * runs / el,
        "unit": "error-locations/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{runs} single-threaded compiles of bb12 L2 with the oracle's forward "
                   f"Pauli-frame compiler (stim-semantics port, not stim) in {el:.1f} s on "
                   f"{platform.processor() or platform.machine()}"),
        "us_per_compile": el / runs * 1e6,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--workload", default="plain", choices=("plain", "batch", "grad", "serving", "long"),
                    help="plain: the headline bb12 compile; the others print their own line")
    ap.add_argument("--variants", type=int, default=256, help="batch workload size")
    ap.add_argument("--sweep-rpt", type=int, default=0,
                    help="tuning: fewest sweep records per thread (1, 2, 4, 8; 0 = library default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 5 device copies of BYTES bytes (scales rocprof byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.calibrate_copy, "copies": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the DEM "
              "compile path shards across GPUs by circuit with no communication", file=sys.stderr)
        return 2

    if args.sweep_rpt:
        from greenpeas_amd import autotune

        autotune.set_tuning("sweep_rpt", args.sweep_rpt)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "grad":
        return bench_grad(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "long":
        return bench_long(args)

    import numpy as np

    from greenpeas_amd import get_driver

    text = read_circuit("bb12")
    log("building driver for bb12 (L2)")
    driver = get_driver(text)
    params = driver.params
    locations = params["num_locations"]
    N = params["num_nodes"]
    Wd = (params["num_detectors"] + params["num_observables"] + 63) // 64

    # reference-comparable compile (host STEPG -> host DEM arrays)
    dem = driver.compile_detector_error_model()
    compile_us = []
    for _ in range(20):
        t0 = time.perf_counter()
        dem = driver.compile_detector_error_model()
        compile_us.append((time.perf_counter() - t0) * 1e6)
    compile_us.sort()
    trimmed = compile_us[1:-1]
    log(f"compile (upload+device+download): {statistics.mean(trimmed):.1f} us")
    h2h = driver.timings()  # the last of those compiles, by stage

    # device-resident throughput
    driver.bench_device(max(args.warmup, 1))
    log(f"warmup done ({args.warmup}); timing {args.steps} steps")
    dev_us = driver.bench_device(args.steps)
    stages = driver.timings()
    value = locations / (dev_us * 1e-6)
    log(f"device pipeline {dev_us:.1f} us/compile -> {value:.3e} error-locations/s")

    nnz = np.diff(dem.offsets)
    C = dem.num_errors
    nnz_bar = float(nnz.mean()) if C else 0.0
    # N' = non-empty p>0 nodes: from the sensitivity stage of this compile
    idx, sens = driver.debug_sensitivities()
    Nprime = int(np.count_nonzero(sens.any(axis=1)))
    model = algorithmic_bytes(N, Wd, Nprime, C, nnz_bar)
    # per-stage device time (HIP events on the pipeline's stream): sweep,
    # classify (+ scatter into the sort buckets), bucket sort + merge, offset
    # scans + emit
    stage_us = {k: stages[k] for k in ("sweep", "classify", "buckets", "emit")}
    dom = max(stage_us, key=stage_us.get)
    dom_bytes = {"sweep": model["sweep"], "classify": model["classify"],
                 "buckets": model["sort"], "emit": model["reduce_emit"]}[dom]
    kernel_name = {"sweep": "sweepKernel", "classify": "classifyKernel", "buckets": "bucketKernel",
                   "emit": "emitKernel"}[dom]
    achieved = dom_bytes / (stage_us[dom] * 1e-6) / 1e9
    sortmerge_us = stage_us["buckets"] + stage_us["emit"]
    sortmerge_gbs = (model["sort"] + model["reduce_emit"]) / (sortmerge_us * 1e-6) / 1e9

    cpu = None if args.no_cpu else cpu_baseline(text, locations)
    if cpu:
        log(f"cpu baseline {cpu['value']:.3e} error-locations/s ({cpu['us_per_compile']:.0f} us/compile)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "error-locations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dev_us / 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 bitsets + f64 merge",
        "data": "reference circuit file data/circuits/bb12.stim (deterministic, no random data)",
        "config": {
            "workload": "plain DEM compile, BB [[144,12,12]] 12-round memory, p=1e-3, L2, device-resident",
            "error_locations": locations,
            "stepg_nodes": N,
            "detectors": params["num_detectors"],
            "observables": params["num_observables"],
            "dem_errors": C,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(kernel_name),
            "algorithmic_bytes": dom_bytes,
            "kernel_us": stage_us[dom],
        },
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "stage_us": stage_us,
        "sort_merge": {"us": sortmerge_us, "algorithmic_GBs": sortmerge_gbs,
                       "frac_hbm": sortmerge_gbs / HBM_PEAK_GBS},
        "compile_us_host_to_host": statistics.mean(trimmed),
        "host_to_host_last_us": {k: h2h[k] for k in ("upload", "download", "total")},
        "reference_compile_us_bb12_L2": REFERENCE_BB12_L2_US,
        "speedup_vs_reference_compile": REFERENCE_BB12_L2_US / statistics.mean(trimmed),
        "speedup_vs_cpu_baseline": None if cpu is None else value / cpu["value"],
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_batch(args) -> int:
    """Batched compile of seeded uneven variants vs the same variants one by
    one through the same kernels (device time; inputs resident)."""
    from greenpeas_amd.batch import BatchCompiler, variant_circuits
    from greenpeas_amd.codes import count_locations

    t0 = time.perf_counter()
    texts = variant_circuits(args.variants, seed=1234)
    locs = sum(count_locations(t) for t in texts)
    log(f"{len(texts)} variants, {locs} error locations, generated in {time.perf_counter() - t0:.1f} s")
    bc = BatchCompiler()
    t0 = time.perf_counter()
    bc.set_circuits(texts)
    log(f"host parse + STEPG build: {time.perf_counter() - t0:.2f} s")
    batched = [bc.run() for _ in range(max(args.warmup, 1) + 3)][-3:]
    log(f"batched device us: {batched}")
    # stage times of the batched pipeline and the sort/merge stages' share of
    # HBM peak at this size (byte model as for the plain workload)
    runs = [bc.stage_us() for _ in range(3)]
    stages, items = min(runs, key=lambda r: r[0]["buckets"] + r[0]["emit"])
    shapes = [bc.shape(i) for i in range(len(texts))]
    C = sum(s[0] for s in shapes)
    nnz_bar = sum(s[1] for s in shapes) / C if C else 0.0
    model = algorithmic_bytes(0, 0, items, C, nnz_bar)
    sm_us = stages["buckets"] + stages["emit"]
    sm_gbs = (model["sort"] + model["reduce_emit"]) / (sm_us * 1e-6) / 1e9
    log(f"batched stages us: {stages}; sort/merge {sm_gbs:.0f} GB/s over {items} items")
    one = bc.run(one_by_one=True)
    log(f"one-by-one device us: {one:.0f}")
    b = min(batched)
    out = {
        "metric": METRIC, "workload": "batch", "variants": len(texts), "error_locations": locs,
        "batched_device_us": b, "one_by_one_device_us": one, "speedup": one / b,
        "batched_error_locations_per_s": locs / (b * 1e-6),
        "device_bytes": bc.device_bytes(),
        "stage_us": stages,
        "sort_merge": {"us": sm_us, "items": items, "dem_errors": C,
                       "algorithmic_bytes": model["sort"] + model["reduce_emit"],
                       "algorithmic_GBs": sm_gbs, "frac_hbm": sm_gbs / HBM_PEAK_GBS},
        "config": "rotated surface memories, d in {5..15} odd, rounds in [d, 2d], per-instruction "
                  "rates logU[1e-4, 1e-2], 5% of mid-circuit measurements followed by a "
                  "reset-conditioned branch, seed 1234",
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_grad(args) -> int:
    """Forward + backward of L = sum w logit(p(theta)) on bb12."""
    import numpy as np
    import torch

    from greenpeas_amd.autograd import DifferentiableDem, logit_loss

    dd = DifferentiableDem(read_circuit("bb12"))
    rng = np.random.default_rng(2024)
    theta = torch.tensor(np.exp(rng.uniform(np.log(1e-4), np.log(5e-3), dd.num_locations)),
                         device="cuda", requires_grad=True)
    w = torch.tensor(rng.standard_normal(dd.num_errors), device="cuda")
    for _ in range(max(args.warmup, 1)):
        theta.grad = None
        logit_loss(dd.probabilities(theta), w).backward()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        theta.grad = None
        logit_loss(dd.probabilities(theta), w).backward()
    b.record()
    torch.cuda.synchronize()
    us = a.elapsed_time(b) * 1e3 / args.steps
    out = {"metric": METRIC, "workload": "grad", "circuit": "bb12 L2", "locations": dd.num_locations,
           "errors": dd.num_errors, "forward_backward_us": us,
           "gradients_per_s": dd.num_locations / (us * 1e-6)}
    print(json.dumps(out), flush=True)
    return 0


def bench_long(args) -> int:
    """Long memory experiments: device time of the compile with every
    column swept whole and with the planned time-chunked sweep (same DEM)."""
    from greenpeas_amd import autotune, get_driver
    from greenpeas_amd.codes import BBCode, SurfaceCode

    cases = [("surface d=7 x 100 rounds", SurfaceCode(7).memory_text(100, 1e-3)),
             ("surface d=11 x 200 rounds", SurfaceCode(11).memory_text(200, 1e-3)),
             ("BB [[72,12,6]] x 60 rounds", BBCode(6).memory_text(60, 1e-3)),
             ("BB [[144,12,12]] x 40 rounds", BBCode(12).memory_text(40, 1e-3))]
    results = []
    for name, text in cases:
        row = {"circuit": name}
        dems = []
        for mode, key in ((0, "whole"), (1, "planned")):
            autotune.set_tuning("sweep_split", mode)
            d = get_driver(text)
            d.compile_detector_error_model()
            dems.append(d.compile_detector_error_model())  # the second compile runs the plan
            d.bench_device(max(args.warmup, 1))
            us = d.bench_device(max(args.steps // 4, 10))
            row[key] = {"device_us": us, "sweep_us": d.timings()["sweep"], "plan": d.debug_sweep_plan()}
            d.close()
        autotune.set_tuning("sweep_split", 1)
        row["same_dem"] = dems[0] == dems[1]
        row["sweep_speedup"] = row["whole"]["sweep_us"] / row["planned"]["sweep_us"]
        log(f"{name}: sweep {row['whole']['sweep_us']:.1f} -> {row['planned']['sweep_us']:.1f} us, "
            f"device {row['whole']['device_us']:.1f} -> {row['planned']['device_us']:.1f} us")
        results.append(row)
    print(json.dumps({"metric": METRIC, "workload": "long", "results": results}), flush=True)
    return 0


def bench_serving(args) -> int:
    """Poisson request streams against the continuous-batching server."""
    import threading

    from greenpeas_amd.serving import Server, poisson_arrivals, request_pool

    pool = request_pool(48, seed=77)
    log(f"request pool of {len(pool)} circuits ready")
    results = []
    for rate in (100, 500, 1000, 2000, 5000):
        n = min(400, max(60, rate // 5))
        arrivals = poisson_arrivals(rate, n, seed=rate)
        # warmed with the pool's largest circuit (buffers sized, kernels loaded)
        srv = Server(slots=2, max_batch=64, warmup=[max(pool, key=len)])
        futs = []
        t0 = time.perf_counter()
        for i, at in enumerate(arrivals):
            now = time.perf_counter() - t0
            if at > now:
                time.sleep(at - now)
            futs.append(srv.submit(pool[i % len(pool)]))
        for f in futs:
            f.result(timeout=600)
        srv.close()
        st = srv.stats()
        st["offered_rps"] = rate
        results.append(st)
        log(f"rate {rate}: p50 {st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms, "
            f"{st['throughput_rps']:.0f} req/s, mean batch {st['mean_batch']:.1f}")
    print(json.dumps({"metric": METRIC, "workload": "serving", "results": results}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
