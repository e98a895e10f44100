"""Batched compilation: many circuits (noise or branch variants of uneven
size) through one set of GPU launches.

The reference compiles one circuit per ``Driver::compile`` call
(Driver.hpp:138-162) and recompiles per branch in its adaptive workload
(Cli.hpp:408-416).  Here all circuits of a batch share every launch: the
sweep has one workgroup per (circuit, detector column), longest time bands
first; classify runs over the concatenated p>0 nodes; each circuit
owns its own range of sort buckets, so the output stays grouped by circuit
and ordered by class hash within each.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native, autotune
from .dem import DetectorErrorModel
from .error_analysis import CorrelationLevel, _read_dem, circuit_text


class PreparedCircuit:
    """A circuit parsed and built up to its device image on the host (the
    host half of a compile: gp_prepare_circuit), ready to be loaded into a
    batch without parsing again."""

    def __init__(self, circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2):
        self._lib = _native.load()
        self._h = ctypes.c_void_p()
        self.level = int(correlation_level)
        _native.check(self._lib.gp_prepare_circuit(circuit_text(circuit).encode(), self.level,
                                                   ctypes.byref(self._h)))

    def device_bytes(self) -> int:
        """Device bytes a new batch context allocates to compile this circuit
        alone (gp_prepared_device_bytes)."""
        b = ctypes.c_uint64()
        _native.check(self._lib.gp_prepared_device_bytes(self._h, ctypes.byref(b)))
        return int(b.value)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.gp_prepared_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


def prepare(circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2) -> PreparedCircuit:
    return PreparedCircuit(circuit, correlation_level)


def _handles(prepared) -> ctypes.Array:
    hs = [p._h.value for p in prepared]
    if not all(hs):
        raise ValueError("a prepared circuit was closed")
    return (ctypes.c_void_p * len(hs))(*hs)


def device_memory(device: int = 0) -> tuple[int, int]:
    """(free, total) bytes of device memory (hipMemGetInfo)."""
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    _native.check(_native.load().gp_device_memory(device, ctypes.byref(f), ctypes.byref(t)))
    return int(f.value), int(t.value)


class BatchCompiler:
    """Compiles a list of circuits on one GPU; device buffers grow on demand
    and are reused across batches."""

    def __init__(self, correlation_level: CorrelationLevel = CorrelationLevel.L2, device: int = 0):
        self._lib = _native.load()
        autotune.install_cached()  # tuned knobs per size class, once per process
        self._h = ctypes.c_void_p()
        _native.check(self._lib.gp_batch_create(int(correlation_level), device, ctypes.byref(self._h)))
        self.count = 0

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.gp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def set_circuits(self, circuits, threads: int = 0) -> None:
        """Parses and builds the STEPGs of `circuits` (host threads)."""
        texts = [circuit_text(c).encode() for c in circuits]
        arr = (ctypes.c_char_p * len(texts))(*texts)
        _native.check(self._lib.gp_batch_set_circuits(self._h, arr, len(texts), threads))
        self.count = len(texts)

    def set_prepared(self, prepared) -> None:
        """Makes prepared circuits (PreparedCircuit, at this batch's level)
        the batch's circuits; nothing is parsed."""
        prepared = list(prepared)
        _native.check(self._lib.gp_batch_set_prepared(self._h, _handles(prepared), len(prepared)))
        self.count = len(prepared)

    def bytes_after(self, prepared) -> int:
        """Device bytes this compiler holds after compiling `prepared` (its
        buffers where they fit, grown where not; gp_batch_bytes_after)."""
        prepared = list(prepared)
        b = ctypes.c_uint64()
        _native.check(self._lib.gp_batch_bytes_after(self._h, _handles(prepared), len(prepared),
                                                     ctypes.byref(b)))
        return int(b.value)

    def compile(self, one_by_one: bool = False) -> list[DetectorErrorModel]:
        """Compiles every circuit of the batch; returns their DEMs in order.
        `one_by_one` runs the same kernels once per circuit instead (the
        baseline the batched launch is measured against)."""
        _native.check(self._lib.gp_batch_compile(self._h, 1 if one_by_one else 0))
        return [self.dem(i) for i in range(self.count)]

    def run(self, one_by_one: bool = False) -> float:
        """Compiles without reading the DEMs back into Python; returns the
        device microseconds of the compile."""
        _native.check(self._lib.gp_batch_compile(self._h, 1 if one_by_one else 0))
        return self.device_us()

    def dem(self, index: int) -> DetectorErrorModel:
        return _read_dem(self._lib, self._lib.gp_batch_dem_shape, self._lib.gp_batch_dem_copy, self._h,
                         ctypes.c_uint32(index))

    def shape(self, index: int) -> tuple[int, int]:
        """(errors, targets) of circuit `index`'s DEM without copying it."""
        ne, nt = ctypes.c_uint32(), ctypes.c_uint64()
        _native.check(self._lib.gp_batch_dem_shape(self._h, ctypes.c_uint32(index), ctypes.byref(ne),
                                                   ctypes.byref(nt), None, None))
        return ne.value, nt.value

    def stage_us(self) -> tuple[dict, int]:
        """Runs the compiled batch once more with events between the stages:
        ({"sweep", "classify", "buckets", "emit"} device microseconds, the
        number of items the sort stages handled).  Call after compile()/run()."""
        st = (ctypes.c_double * 4)()
        items = ctypes.c_uint64()
        _native.check(self._lib.gp_batch_stage_us(self._h, st, ctypes.byref(items)))
        return dict(zip(("sweep", "classify", "buckets", "emit"), list(st))), int(items.value)

    def device_us(self) -> float:
        us = ctypes.c_double()
        _native.check(self._lib.gp_batch_device_us(self._h, ctypes.byref(us)))
        return us.value

    def layout_stats(self) -> dict:
        """{"s_words": words of the sensitivity matrix, "band_relayouts":
        compiles that found a column band below its stored layers and ran
        again with more, "buckets": sort buckets, "bucket_cap": slots per
        bucket} (gp_batch_layout_stats)."""
        st = (ctypes.c_uint64 * 4)()
        _native.check(self._lib.gp_batch_layout_stats(self._h, st))
        return {"s_words": int(st[0]), "band_relayouts": int(st[1]), "buckets": int(st[2]),
                "bucket_cap": int(st[3])}

    def device_bytes(self) -> int:
        b = ctypes.c_uint64()
        _native.check(self._lib.gp_batch_device_bytes(self._h, ctypes.byref(b)))
        return int(b.value)


def compile_batch(circuits, correlation_level: CorrelationLevel = CorrelationLevel.L2,
                  device: int = 0) -> list[DetectorErrorModel]:
    bc = BatchCompiler(correlation_level, device)
    bc.set_circuits(circuits)
    out = bc.compile()
    bc.close()
    return out


def variant_circuits(count: int, seed: int, distances=(5, 7, 9, 11, 13, 15),
                     branch_fraction: float = 0.05, p_range=(1e-4, 1e-2)) -> list[str]:
    """The batched workload: `count` rotated-surface-code memories with a
    seeded distance d from `distances`, rounds in [d, 2d], every noise
    instruction's rate log-uniform in `p_range`, and a seeded fraction of
    mid-circuit measurements followed by a reset-conditioned branch."""
    from .codes import Noise, SurfaceCode

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = int(rng.choice(distances))
        rounds = int(rng.integers(d, 2 * d + 1))
        code = SurfaceCode(d)
        nchecks = len(code.checks)
        branches = set()
        for r in range(rounds - 1):  # mid-circuit measurement rounds
            for a in range(nchecks):
                if rng.random() < branch_fraction:
                    branches.add((r, a))
        noise = Noise(rng=np.random.default_rng(int(rng.integers(1 << 62))), lo=p_range[0], hi=p_range[1])
        out.append(code.memory_text(rounds, noise=noise, branches=branches))
    return out
