"""Autotuner for the pipeline's launch-shape knobs.

* ``bucket_fill`` -- the sort's first radix digit: the pipeline partitions
  each circuit's (hash, node) items into 2^b buckets by the top b hash bits
  and sorts each bucket in one wave64 (up to 256 items in registers); b is
  the smallest width with ``bucket_fill * 2^b >= nodes``.  Small buckets mean
  a short register sort but more buckets to scan and emit; large ones the
  opposite.
* ``bucket_cap`` -- slots per bucket (64, 128, 256): with the digit width it
  is the shape of the hash table; nodes past a bucket's slots go to an
  overflow list sorted by the bucket's whole workgroup.  Fewer slots mean
  less memory and a denser table, more overflow.
* ``sweep_rpt`` -- the fewest STEPG records per thread in the sensitivity
  sweep (the wave tile): more records per thread means fewer waves meeting
  at each layer's barrier and more work per wave between barriers.

``tune_driver`` times candidate values on the GPU for one circuit (one knob
after the other, each kept for its size class while the next is timed) and
keeps the fastest for circuits of its size class (level, floor(log2 STEPG
nodes)); results are cached in a JSON file.  The library applies a size
class's tuned values itself when it parses or loads a circuit of that class
(gp_set_tuning_class); ``install_cached`` hands it the cache, once per
process, before the first driver or batch is made.  ``tune_batch`` tunes the
process-wide values for a batch.  Every knob leaves the DEMs bit-identical.
"""

from __future__ import annotations

import ctypes
import json
import os
import threading

from . import _native

CANDIDATES = (24, 32, 48, 64, 96, 128)  # bucket_fill
CAP_CANDIDATES = (64, 128, 256)         # bucket_cap
RPT_CANDIDATES = (1, 2, 4, 8)            # sweep_rpt
KNOBS = ("sweep_rpt", "bucket_fill", "bucket_cap")
CACHE = os.environ.get("GREENPEAS_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "greenpeas_amd_tune.json"))


def set_tuning(name: str, value: int) -> None:
    _native.check(_native.load().gp_set_tuning(name.encode(), int(value)))


def get_tuning(name: str) -> int:
    v = ctypes.c_uint32()
    _native.check(_native.load().gp_get_tuning(name.encode(), ctypes.byref(v)))
    return int(v.value)


def set_tuning_class(level: int, size_class: int, name: str, value: int) -> None:
    """Knob `name` = `value` for circuits of this level and size class (0
    removes the entry)."""
    _native.check(_native.load().gp_set_tuning_class(int(level), int(size_class), name.encode(), int(value)))


def get_tuning_class(level: int, size_class: int, name: str) -> int:
    v = ctypes.c_uint32()
    _native.check(_native.load().gp_get_tuning_class(int(level), int(size_class), name.encode(),
                                                     ctypes.byref(v)))
    return int(v.value)


def clear_tuning_classes() -> None:
    _native.check(_native.load().gp_clear_tuning_classes())


def set_bucket_fill(fill: int) -> None:
    set_tuning("bucket_fill", fill)


def get_bucket_fill() -> int:
    return get_tuning("bucket_fill")


def size_class(num_nodes: int) -> int:
    """floor(log2 nodes), the library's size class of a circuit."""
    return max(int(num_nodes), 1).bit_length() - 1


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


_installed = False
_install_lock = threading.Lock()


def install_cached(force: bool = False) -> int:
    """Hands every cached size class's knobs to the library (once per
    process unless `force`); returns the number of classes installed."""
    global _installed
    with _install_lock:
        if _installed and not force:
            return 0
        _installed = True
        n = 0
        for key, best in _load_cache().items():
            try:
                level, cls = key[1:].split(":")
                level, cls = int(level), int(cls)
            except ValueError:
                continue
            if isinstance(best, int):  # older cache entries hold the fill only
                best = {"bucket_fill": best}
            for name in KNOBS:
                if name in best:
                    set_tuning_class(level, cls, name, int(best[name]))
            n += 1
        return n


def _tune_class_knob(level: int, cls: int, name: str, candidates, measure) -> dict:
    """Times `measure()` with each candidate value of knob `name` for the
    size class; keeps the fastest there and returns {"best", "us"}."""
    results = {}
    for v in candidates:
        set_tuning_class(level, cls, name, v)
        results[v] = measure()
    best = min(results, key=results.get)
    set_tuning_class(level, cls, name, best)
    return {"best": best, "us": results}


def _tune_knob(name: str, candidates, measure) -> dict:
    """Times `measure()` with each candidate value of the process-wide knob
    `name`; leaves the fastest set and returns {"best", "us"}."""
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
                rpt_candidates=RPT_CANDIDATES, cap_candidates=CAP_CANDIDATES, cache: bool = True) -> dict:
    """Times the candidate sweep_rpt, bucket_fill and bucket_cap values on
    one circuit's device pipeline, one after the other, for its size class;
    keeps the fastest for that class in the library (and the cache)."""
    from .error_analysis import Driver

    def measure():
        drv = Driver(circuit, level)
        try:
            drv.compile_detector_error_model()
            return drv.bench_device(iterations)
        finally:
            drv.close()

    drv = Driver(circuit, level)
    nodes = drv.params["num_nodes"]
    drv.close()
    cls = size_class(nodes)
    out = {"size_class": cls}
    for name, cands in (("sweep_rpt", rpt_candidates), ("bucket_fill", candidates),
                        ("bucket_cap", cap_candidates)):
        out[name] = _tune_class_knob(level, cls, name, cands, measure)
    if cache:
        c = _load_cache()
        c[f"L{level}:{cls}"] = {name: out[name]["best"] for name in KNOBS}
        _store_cache(c)
    out["best"] = out["bucket_fill"]["best"]
    return out


def tune_batch(circuits, level: int = 2, candidates=CANDIDATES, rpt_candidates=RPT_CANDIDATES,
               cap_candidates=CAP_CANDIDATES) -> dict:
    """The process-wide knobs for a batch compiled in one pipeline (size
    class entries are set aside while timing and handed back after)."""
    from .batch import BatchCompiler

    def measure():
        bc = BatchCompiler(level)
        try:
            bc.set_circuits(circuits)
            return min(bc.run() for _ in range(3))
        finally:
            bc.close()

    clear_tuning_classes()
    try:
        out = {name: _tune_knob(name, cands, measure)
               for name, cands in (("sweep_rpt", rpt_candidates), ("bucket_fill", candidates),
                                   ("bucket_cap", cap_candidates))}
    finally:
        install_cached(force=True)
    out["best"] = out["bucket_fill"]["best"]
    return out


def apply_cached(num_nodes: int, level: int = 2) -> int | None:
    """Installs the cached classes (see install_cached) and returns the
    tuned bucket fill of circuits of this size (None when none is cached)."""
    install_cached()
    v = get_tuning_class(level, size_class(num_nodes), "bucket_fill")
    return v or None
