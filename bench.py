"""Benchmark of the DEM compile hot path on one MI355X.

Workload (a "step"): one device-resident compile of BB [[144,12,12]], 12
rounds, circuit-level depolarizing p=1e-3, correlation level L2 -- the
reference's data/circuits/bb12.stim -- with its STEPG already uploaded to HBM.
The timed region is K back-to-back compiles bracketed by HIP events on the
pipeline's own stream.  `value` is error (fault) locations compiled per
second: 43,632 per compile x K / elapsed.

The same JSON line also carries:
  * per-stage device time and the roofline of the dominant kernel (model
    bytes, and the HBM traffic the committed rocprofv3 counters measured);
  * three spans timed with the reference's protocol (one warm-up, n calls,
    min and max dropped, mean +- sample std; Perf/Timing.hpp:97-116):
    (i) device pipeline, (ii) STEPG in host memory -> DEM arrays on host (the
    span the reference times, Cli.hpp:293-302), (iii) end to end from circuit
    text (parse + STEPG build + (ii));
  * the batched workload (256 seeded uneven variants, batched vs one by one),
    the differentiable workload (forward + backward on bb12), the serving
    workload (Poisson streams at 100 and 1000 requests/s) and the adaptive
    Iceberg x surface workload (64 seeded branches of d = 6, the reference's
    recompile-per-branch loop against one batch);
  * a CPU baseline: the oracle's forward Pauli-frame compiler (a port with
    stim's semantics, not stim) on one core and on every core of this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
Other workloads (--workload, one JSON line each): batch, grad, serving (the
full rate sweep 100-5000 requests/s), long (40-200-round memories, whole vs
time-chunked sweep), adaptive (Iceberg x surface d = 4 and 6).
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


def trimmed_stats(samples) -> dict:
    """The reference's statistic (Perf/Timing.hpp:97-116): drop the minimum
    and the maximum, then mean and sample standard deviation."""
    xs = sorted(samples)
    if len(xs) > 2:
        xs = xs[1:-1]
    return {"mean": statistics.mean(xs), "std": statistics.stdev(xs) if len(xs) > 1 else 0.0,
            "n": len(samples)}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_cores() -> int:
    """Cores this process may run on (the box's share, not the machine's)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(text: str, locations: int, budget_s: float = 6.0) -> dict:
    """The oracle's forward Pauli-frame compiler (stim's semantics, not stim)
    on bb12 L2: single-threaded per compile with the reference's statistic,
    then every core of this host running independent compiles (ctypes
    releases the GIL, the oracle keeps per-thread state)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    O.compile_frames(text)  # warm
    per = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and len(per) < 200:
        a = time.perf_counter()
        O.compile_frames(text)
        per.append((time.perf_counter() - a) * 1e6)
    one = trimmed_stats(per)
    cores = host_cores()

    def worker(deadline):
        n = 0
        while time.perf_counter() < deadline:
            O.compile_frames(text)
            n += 1
        return n

    t0 = time.perf_counter()
    deadline = t0 + budget_s
    with ThreadPoolExecutor(cores) as ex:
        total = sum(ex.map(worker, [deadline] * cores))
    el = time.perf_counter() - t0
    model = cpu_model()
    return {
        "value": locations / (one["mean"] * 1e-6),
        "unit": "error-locations/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{len(per)} single-threaded compiles of bb12 L2 ({budget_s:.0f} s) with the oracle's "
                   f"forward Pauli-frame compiler (stim-semantics port, not stim) on {model}"),
        "us_per_compile": one["mean"],
        "us_per_compile_std": one["std"],
        "cpu_model": model,
        "all_cores": {"value": locations * total / el, "unit": "error-locations/s", "cores": cores,
                      "compiles": total, "seconds": el,
                      "sample": f"{cores} threads of independent bb12 L2 compiles for {el:.1f} s"},
    }


def measure_plain(args) -> dict:
    """The headline workload; returns the JSON line's core keys."""
    import numpy as np

    from greenpeas_amd import get_driver

    text = read_circuit("bb12")
    log("building driver for bb12 (L2)")
    driver = get_driver(text)
    params = driver.params
    locations = params["num_locations"]
    N = params["num_nodes"]
    Wd = (params["num_detectors"] + params["num_observables"] + 63) // 64

    # (ii) STEPG in host memory -> DEM arrays on host (the reference's span)
    dem = driver.compile_detector_error_model()
    h2h_samples = []
    for _ in range(args.protocol_n):
        t0 = time.perf_counter()
        dem = driver.compile_detector_error_model()
        h2h_samples.append((time.perf_counter() - t0) * 1e6)
    h2h = trimmed_stats(h2h_samples)
    h2h_split = driver.timings()  # the last of those compiles, by stage
    log(f"(ii) compile STEPG host -> DEM host: {h2h['mean']:.1f} +- {h2h['std']:.1f} us")

    # (iii) end to end from circuit text
    e2e_samples = []
    for _ in range(max(args.protocol_n // 5, 5)):
        t0 = time.perf_counter()
        driver.compile_detector_error_model(text)
        e2e_samples.append((time.perf_counter() - t0) * 1e6)
    e2e = trimmed_stats(e2e_samples)
    log(f"(iii) end to end from text: {e2e['mean']:.1f} +- {e2e['std']:.1f} us")

    # (i) device-resident: the headline K back-to-back steps, then the
    # reference's statistic over single timed compiles
    driver.bench_device(max(args.warmup, 1))
    log(f"warmup done ({args.warmup}); timing {args.steps} steps")
    dev_us = driver.bench_device(args.steps)
    stages = driver.timings()
    value = locations / (dev_us * 1e-6)
    log(f"device pipeline {dev_us:.1f} us/compile -> {value:.3e} error-locations/s")
    dev = trimmed_stats([driver.bench_device(1) for _ in range(args.protocol_n)])

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
    traffic = measured_traffic(kernel_name)
    traffic_gbs = None if traffic is None else traffic / (stage_us[dom] * 1e-6) / 1e9
    # what limits the kernel: its measured bytes against the bandwidth peak
    if traffic_gbs is None:
        limiter = "unmeasured"
    elif traffic_gbs >= 0.5 * HBM_PEAK_GBS:
        limiter = "hbm"
    else:
        limiter = ("latency: measured traffic is {:.2f} of HBM peak; the kernel waits on its "
                   "dependency chain, not on bandwidth".format(traffic_gbs / HBM_PEAK_GBS))
    sortmerge_us = stage_us["buckets"] + stage_us["emit"]
    sortmerge_gbs = (model["sort"] + model["reduce_emit"]) / (sortmerge_us * 1e-6) / 1e9
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
            "traffic": traffic,
            "algorithmic_bytes": dom_bytes,
            "kernel_us": stage_us[dom],
            "traffic_GBs": traffic_gbs,
            "traffic_frac": None if traffic_gbs is None else traffic_gbs / HBM_PEAK_GBS,
            "limiter": limiter,
        },
        "stage_us": stage_us,
        "sort_merge": {"us": sortmerge_us, "algorithmic_GBs": sortmerge_gbs,
                       "frac_hbm": sortmerge_gbs / HBM_PEAK_GBS},
        "spans_us": {
            "protocol": "1 warm-up, n calls, min and max dropped, mean +- sample std (Timing.hpp:97-116)",
            "device": dev,
            "stepg_host_to_dem_host": h2h,
            "end_to_end_from_text": e2e,
        },
        "compile_us_host_to_host": h2h["mean"],
        "host_to_host_last_us": {k: h2h_split[k] for k in ("upload", "download", "total")},
        "reference_compile_us_bb12_L2": REFERENCE_BB12_L2_US,
        "speedup_vs_reference_compile": REFERENCE_BB12_L2_US / h2h["mean"],
    }
    driver.close()
    return out


def measure_batch(variants: int, warmup: int) -> dict:
    """Batched compile of seeded uneven variants vs the same variants one by
    one through the same kernels (device time; inputs resident)."""
    from greenpeas_amd.batch import BatchCompiler, variant_circuits
    from greenpeas_amd.codes import count_locations

    t0 = time.perf_counter()
    texts = variant_circuits(variants, seed=1234)
    locs = sum(count_locations(t) for t in texts)
    log(f"{len(texts)} variants, {locs} error locations, generated in {time.perf_counter() - t0:.1f} s")
    bc = BatchCompiler()
    t0 = time.perf_counter()
    bc.set_circuits(texts)
    host_s = time.perf_counter() - t0
    log(f"host parse + STEPG build: {host_s:.2f} s")
    batched = [bc.run() for _ in range(max(warmup, 1) + 3)][-3:]
    log(f"batched device us: {[round(b, 1) for b in batched]}")
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
    batched_bytes = bc.device_bytes()
    one = bc.run(one_by_one=True)
    log(f"one-by-one device us: {one:.0f}")
    b = min(batched)
    out = {
        "variants": len(texts), "error_locations": locs,
        "batched_device_us": b, "one_by_one_device_us": one, "speedup": one / b,
        "batched_error_locations_per_s": locs / (b * 1e-6),
        "host_build_s": host_s,
        "device_bytes_batched": batched_bytes,
        "stage_us": stages,
        "sort_merge": {"us": sm_us, "items": items, "dem_errors": C,
                       "algorithmic_bytes": model["sort"] + model["reduce_emit"],
                       "algorithmic_GBs": sm_gbs, "frac_hbm": sm_gbs / HBM_PEAK_GBS},
        "config": "rotated surface memories, d in {5..15} odd, rounds in [d, 2d], per-instruction "
                  "rates logU[1e-4, 1e-2], 5% of mid-circuit measurements followed by a "
                  "reset-conditioned branch, seed 1234",
    }
    bc.close()
    return out


def measure_grad(steps: int, warmup: int) -> dict:
    """Forward + backward of L = sum w logit(p(theta)) on bb12."""
    import numpy as np
    import torch

    from greenpeas_amd.autograd import DifferentiableDem, logit_loss

    dd = DifferentiableDem(read_circuit("bb12"))
    rng = np.random.default_rng(2024)
    theta = torch.tensor(np.exp(rng.uniform(np.log(1e-4), np.log(5e-3), dd.num_locations)),
                         device="cuda", requires_grad=True)
    w = torch.tensor(rng.standard_normal(dd.num_errors), device="cuda")
    for _ in range(max(warmup, 1)):
        theta.grad = None
        logit_loss(dd.probabilities(theta), w).backward()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        theta.grad = None
        logit_loss(dd.probabilities(theta), w).backward()
    b.record()
    torch.cuda.synchronize()
    us = a.elapsed_time(b) * 1e3 / steps
    log(f"grad: forward + backward {us:.1f} us")
    return {"circuit": "bb12 L2", "locations": dd.num_locations, "errors": dd.num_errors,
            "forward_backward_us": us, "gradients_per_s": dd.num_locations / (us * 1e-6)}


def measure_serving(rates, max_requests: int = 400) -> list:
    """Poisson request streams against the continuous-batching server."""
    from greenpeas_amd.serving import Server, poisson_arrivals, request_pool

    pool = request_pool(48, seed=77)
    log(f"request pool of {len(pool)} circuits ready")
    results = []
    for rate in rates:
        n = min(max_requests, max(60, rate // 5))
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
    return results


def measure_adaptive(d: int = 6, count: int = 64, p: float = 1e-3) -> dict:
    """The adaptive Iceberg x surface-code workload (Cli.hpp:333-432): `count`
    seeded branches of the d-distance memory (rounds = d), compiled (i) as
    the reference does, one driver sized from the static circuit recompiling
    each branch (host parse + STEPG + device + DEM download per branch), and
    (ii) all branches in one batch (host builds on the prepare threads, one
    set of launches)."""
    from concurrent.futures import ThreadPoolExecutor

    from greenpeas_amd import get_driver
    from greenpeas_amd.batch import BatchCompiler, prepare
    from greenpeas_amd.codes import count_locations
    from greenpeas_amd.iceberg import IcebergSurfaceCode, adaptive_branches

    static, _ = IcebergSurfaceCode(d).memory(d, p)
    branches = adaptive_branches(d, count, p=p, seed=2026)
    locs = sum(count_locations(t) for t in branches)
    drv = get_driver(static)
    drv.compile_detector_error_model()
    for t in branches[:4]:  # warm
        drv.compile_detector_error_model(t)
    t0 = time.perf_counter()
    for t in branches:
        drv.compile_detector_error_model(t)
    loop_s = time.perf_counter() - t0
    drv.close()
    bc = BatchCompiler()
    bc.set_circuits(branches)
    bc.run()  # warm (buffers sized)
    with ThreadPoolExecutor(8) as pool:
        t0 = time.perf_counter()
        prepared = list(pool.map(prepare, branches))
        bc.set_prepared(prepared)
        dev_us = bc.run()
        dems = [bc.dem(i) for i in range(len(branches))]
        batch_s = time.perf_counter() - t0
    bc.close()
    out = {"code": f"Iceberg x surface d={d}, {d} rounds, adaptive, p={p}", "branches": count,
           "error_locations": locs, "distinct_circuits": len(set(branches)),
           "loop_us_per_branch": loop_s / count * 1e6, "batch_us_per_branch": batch_s / count * 1e6,
           "batch_device_us": dev_us, "speedup": loop_s / batch_s, "errors_first": dems[0].num_errors}
    log(f"adaptive d={d}: {count} branches, loop {out['loop_us_per_branch']:.0f} us/branch, "
        f"batch {out['batch_us_per_branch']:.0f} us/branch (device {dev_us:.0f} us total)")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--quick", action="store_true",
                    help="plain workload only (no batch, grad and serving keys)")
    ap.add_argument("--protocol-n", type=int, default=100,
                    help="timed calls per span for the reference's statistic")
    ap.add_argument("--workload", default="plain",
                    choices=("plain", "batch", "grad", "serving", "long", "adaptive"),
                    help="plain: the headline bb12 compile; the others print their own line")
    ap.add_argument("--variants", type=int, default=256, help="batch workload size")
    ap.add_argument("--sweep-rpt", type=int, default=0,
                    help="tuning: fewest sweep records per thread (1, 2, 4, 8; 0 = library default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 5 device copies of BYTES bytes (scales rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the DEM compile path "
              "shards across GPUs by circuit with no communication, and this build is measured "
              "on one GPU", file=sys.stderr)
        return 2
    # torch first: the library then binds to torch's HIP runtime, which the
    # differentiable workload shares
    import torch  # noqa: F401

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.calibrate_copy, "copies": 5}), flush=True)
        return 0

    if args.sweep_rpt:
        from greenpeas_amd import autotune

        autotune.set_tuning("sweep_rpt", args.sweep_rpt)
    if args.workload == "batch":
        out = {"metric": METRIC, "workload": "batch", **measure_batch(args.variants, args.warmup)}
        print(json.dumps(out), flush=True)
        return 0
    if args.workload == "grad":
        out = {"metric": METRIC, "workload": "grad", **measure_grad(args.steps, args.warmup)}
        print(json.dumps(out), flush=True)
        return 0
    if args.workload == "serving":
        res = measure_serving((100, 500, 1000, 2000, 5000))
        print(json.dumps({"metric": METRIC, "workload": "serving", "results": res}), flush=True)
        return 0
    if args.workload == "long":
        return bench_long(args)
    if args.workload == "adaptive":
        print(json.dumps({"metric": METRIC, "workload": "adaptive", "results": [
            measure_adaptive(4, 64), measure_adaptive(6, 64)]}), flush=True)
        return 0

    out = measure_plain(args)
    locations = out["config"]["error_locations"]
    if not args.quick:
        out["batch"] = measure_batch(args.variants, 2)
        out["grad"] = measure_grad(50, 5)
        out["serving"] = measure_serving((100, 1000))
        out["adaptive"] = measure_adaptive(6, 64)
    cpu = None if args.no_cpu else cpu_baseline(read_circuit("bb12"), locations)
    if cpu:
        log(f"cpu baseline {cpu['value']:.3e} error-locations/s on 1 core "
            f"({cpu['us_per_compile']:.0f} us/compile), {cpu['all_cores']['value']:.3e} on "
            f"{cpu['all_cores']['cores']} cores")
    out["cpu_baseline"] = None if cpu is None else {
        k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "us_per_compile",
                            "us_per_compile_std", "cpu_model", "all_cores")}
    out["speedup_vs_cpu_baseline"] = None if cpu is None else out["value"] / cpu["value"]
    out["speedup_vs_cpu_all_cores"] = None if cpu is None else out["value"] / cpu["all_cores"]["value"]
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



if __name__ == "__main__":
    sys.exit(main())
