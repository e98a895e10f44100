"""Autotuner for the sort's first radix digit (bucket width).

The pipeline partitions each circuit's (hash, node) items into 2^b buckets
by the top b hash bits and sorts each bucket in one wave64 (up to 256 items
in registers); b is the smallest width with ``bucket_fill * 2^b >= nodes``.
Small buckets mean a short register sort but more buckets to scan and emit;
large ones the opposite.  ``tune`` times candidate fills on the GPU for a
workload and keeps the fastest; results are cached per (level, size class)
in a JSON file so later drivers and batches pick them up with
``apply_cached``.
"""

from __future__ import annotations

import ctypes
import json
import math
import os

from . import _native

CANDIDATES = (24, 32, 48, 64, 96, 128)
CACHE = os.environ.get("GREENPEAS_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "greenpeas_amd_tune.json"))


def set_bucket_fill(fill: int) -> None:
    _native.check(_native.load().gp_set_tuning(b"bucket_fill", int(fill)))


def get_bucket_fill() -> int:
    v = ctypes.c_uint32()
    _native.check(_native.load().gp_get_tuning(b"bucket_fill", ctypes.byref(v)))
    return int(v.value)


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


def tune_driver(circuit, level: int = 2, iterations: int = 20, candidates=CANDIDATES) -> dict:
    """Times each candidate fill on one circuit's device pipeline; keeps the
    fastest in the process and the cache."""
    from .error_analysis import Driver

    results = {}
    old = get_bucket_fill()
    try:
        for fill in candidates:
            set_bucket_fill(fill)
            drv = Driver(circuit, level)
            drv.compile_detector_error_model()
            results[fill] = drv.bench_device(iterations)
            nodes = drv.params["num_nodes"]
            drv.close()
    finally:
        set_bucket_fill(old)
    best = min(results, key=results.get)
    set_bucket_fill(best)
    cache = _load_cache()
    cache[f"L{level}:{size_class(nodes)}"] = best
    _store_cache(cache)
    return {"best": best, "us": results}


def tune_batch(circuits, level: int = 2, candidates=CANDIDATES) -> dict:
    """Same, for a batch of circuits compiled in one pipeline."""
    from .batch import BatchCompiler

    results = {}
    old = get_bucket_fill()
    try:
        for fill in candidates:
            set_bucket_fill(fill)
            bc = BatchCompiler(level)
            bc.set_circuits(circuits)
            results[fill] = min(bc.run() for _ in range(3))
            bc.close()
    finally:
        set_bucket_fill(old)
    best = min(results, key=results.get)
    set_bucket_fill(best)
    return {"best": best, "us": results}


def apply_cached(num_nodes: int, level: int = 2) -> int | None:
    """Sets the cached best fill for circuits of this size, if any."""
    best = _load_cache().get(f"L{level}:{size_class(num_nodes)}")
    if best is not None:
        set_bucket_fill(int(best))
    return best
