"""Autotuner for the pipeline's two launch-shape knobs.

* ``bucket_fill`` -- the sort's first radix digit: the pipeline partitions
  each circuit's (hash, node) items into 2^b buckets by the top b hash bits
  and sorts each bucket in one wave64 (up to 256 items in registers); b is
  the smallest width with ``bucket_fill * 2^b >= nodes``.  Small buckets mean
  a short register sort but more buckets to scan and emit; large ones the
  opposite.
* ``sweep_rpt`` -- the fewest STEPG records per thread in the sensitivity
  sweep: more records per thread means fewer waves meeting at each layer's
  barrier and more work per wave between barriers.

``tune_driver`` / ``tune_batch`` time candidate values on the GPU for a
workload (one knob after the other) and keep the fastest; results are cached
per (level, size class) in a JSON file so later drivers and batches pick
them up with ``apply_cached``.
"""

from __future__ import annotations

import ctypes
import json
import math
import os

from . import _native

CANDIDATES = (24, 32, 48, 64, 96, 128)  # bucket_fill
RPT_CANDIDATES = (1, 2, 4, 8)            # sweep_rpt
CACHE = os.environ.get("GREENPEAS_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "greenpeas_amd_tune.json"))


def set_tuning(name: str, value: int) -> None:
    _native.check(_native.load().gp_set_tuning(name.encode(), int(value)))


def get_tuning(name: str) -> int:
    v = ctypes.c_uint32()
    _native.check(_native.load().gp_get_tuning(name.encode(), ctypes.byref(v)))
    return int(v.value)


def set_bucket_fill(fill: int) -> None:
    set_tuning("bucket_fill", fill)


def get_bucket_fill() -> int:
    return get_tuning("bucket_fill")


def size_class(num_nodes: int) -> int:
    return int(math.log2(max(num_nodes, 1)))


def _load_cache() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: dict) -> None:
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def _tune_knob(name: str, candidates, measure) -> dict:
    """Times `measure()` with each candidate value of knob `name`; leaves the
    fastest set and returns {"best": value, "us": {value: us}}."""
    results = {}
    old = get_tuning(name)
    try:
        for v in candidates:
            set_tuning(name, v)
            results[v] = measure()
    finally:
        set_tuning(name, old)
    best = min(results, key=results.get)
    set_tuning(name, best)
    return {"best": best, "us": results}


def tune_driver(circuit, level: int = 2, iterations: int = 20, candidates=CANDIDATES,
                rpt_candidates=RPT_CANDIDATES) -> dict:
    """Times the candidate sweep_rpt, then bucket_fill values on one
    circuit's device pipeline; keeps the fastest in the process and the cache."""
    from .error_analysis import Driver

    nodes = 0

    def measure():
        nonlocal nodes
        drv = Driver(circuit, level)
        try:
            drv.compile_detector_error_model()
            nodes = drv.params["num_nodes"]
            return drv.bench_device(iterations)
        finally:
            drv.close()

    rpt = _tune_knob("sweep_rpt", rpt_candidates, measure)
    fill = _tune_knob("bucket_fill", candidates, measure)
    cache = _load_cache()
    cache[f"L{level}:{size_class(nodes)}"] = {"bucket_fill": fill["best"], "sweep_rpt": rpt["best"]}
    _store_cache(cache)
    return {"best": fill["best"], "us": fill["us"], "sweep_rpt": rpt}


def tune_batch(circuits, level: int = 2, candidates=CANDIDATES, rpt_candidates=RPT_CANDIDATES) -> dict:
    """Same, for a batch of circuits compiled in one pipeline."""
    from .batch import BatchCompiler

    def measure():
        bc = BatchCompiler(level)
        try:
            bc.set_circuits(circuits)
            return min(bc.run() for _ in range(3))
        finally:
            bc.close()

    rpt = _tune_knob("sweep_rpt", rpt_candidates, measure)
    fill = _tune_knob("bucket_fill", candidates, measure)
    return {"best": fill["best"], "us": fill["us"], "sweep_rpt": rpt}


def apply_cached(num_nodes: int, level: int = 2) -> int | None:
    """Sets the cached best knobs for circuits of this size, if any; returns
    the bucket fill (None when nothing is cached)."""
    best = _load_cache().get(f"L{level}:{size_class(num_nodes)}")
    if best is None:
        return None
    if isinstance(best, int):  # older cache entries hold the fill only
        best = {"bucket_fill": best}
    set_tuning("bucket_fill", int(best["bucket_fill"]))
    if "sweep_rpt" in best:
        set_tuning("sweep_rpt", int(best["sweep_rpt"]))
    return int(best["bucket_fill"])
