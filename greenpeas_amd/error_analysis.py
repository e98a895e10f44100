"""Drop-in error-analysis API: ``get_driver`` / ``Driver`` / ``CorrelationLevel``.

Mirrors the reference's pybind module ``_gppy.error_analysis``
(source/QEC/ErrorAnalysis/ErrorAnalysis.pybind.cu:47-79), re-exported by
``greenpeas.error_analysis`` (python/greenpeas/error_analysis/__init__.py:3):

* ``get_driver(circuit, correlation_level=CorrelationLevel.L2)`` builds a
  driver whose device buffers are sized for ``circuit``;
* ``Driver.compile_detector_error_model(circuit)`` compiles any circuit that
  fits in those buffers; ``Driver.compile_detector_error_model()`` recompiles
  the circuit parsed last;
* errors: unsupported gates and bad arguments raise ``ValueError`` (the
  reference's ``std::invalid_argument``), capacity overflows and flow-rule
  violations raise ``RuntimeError`` (``std::runtime_error``).

A circuit is stim circuit text, or any object whose ``str()`` is stim circuit
text (e.g. a ``stim.Circuit``), or a :class:`pathlib.Path` to a ``.stim`` file.
"""

from __future__ import annotations

import ctypes
import enum
import sys
import os

import numpy as np

from . import _native, autotune
from .dem import DetectorErrorModel


class CorrelationLevel(enum.IntEnum):
    """Pauli components modelled per noise channel (Mixer.hpp:13-20):
    L0 X/Z and XX/ZZ pairs, L1 adds Y/XZ/ZX, L2 every component (the stim-
    equivalent level)."""

    L0 = 0
    L1 = 1
    L2 = 2


def circuit_text(circuit) -> str:
    """stim circuit text of `circuit` (text, stim.Circuit-like object or path)."""
    if isinstance(circuit, os.PathLike):
        with open(circuit) as f:
            return f.read()
    if isinstance(circuit, bytes):
        return circuit.decode()
    if isinstance(circuit, str):
        return circuit
    return str(circuit)


def _read_dem(lib, shape_fn, copy_fn, *handle) -> DetectorErrorModel:
    ne, nt, nd, no = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
    _native.check(shape_fn(*handle, ctypes.byref(ne), ctypes.byref(nt), ctypes.byref(nd), ctypes.byref(no)))
    n, t = ne.value, nt.value
    probs = np.empty(n, np.float64)
    offsets = np.empty(n + 1, np.uint32)
    targets = np.empty(max(t, 1), np.uint32)
    hashes = np.empty(max(n, 1), np.uint64)
    _native.check(copy_fn(*handle, probs.ctypes.data, offsets.ctypes.data, targets.ctypes.data,
                          hashes.ctypes.data))
    return DetectorErrorModel(probs, offsets, targets[:t], nd.value, no.value, hashes[:n])


class _PinnedBlock:
    """Holds a gp_dem_view token; lets go of it when the last array over
    its memory is gone."""

    __slots__ = ("_lib", "_token")

    def __init__(self, lib, token):
        self._lib = lib
        self._token = token

    def __del__(self):
        if self._token and not sys.is_finalizing():
            self._lib.gp_dem_release(self._token)
            self._token = None


def _pinned_array(addr: int, n: int, ctype, dtype, owner: _PinnedBlock) -> np.ndarray:
    if n == 0 or not addr:
        return np.empty(0, dtype)
    buf = (ctype * n).from_address(addr)
    buf._owner = owner  # the array's base holds the block
    return np.frombuffer(buf, dtype)


def _view_dem(lib, handle) -> DetectorErrorModel:
    v = _native.DemView()
    _native.check(lib.gp_dem_acquire(handle, ctypes.byref(v)))
    owner = _PinnedBlock(lib, v.token)
    n, t = v.num_errors, v.num_targets
    probs = _pinned_array(v.probs, n, ctypes.c_double, np.float64, owner)
    offsets = _pinned_array(v.offsets, n + 1, ctypes.c_uint32, np.uint32, owner)
    if v.target_bytes == 2:
        targets = _pinned_array(v.targets, t, ctypes.c_uint16, np.uint16, owner)
    else:
        targets = _pinned_array(v.targets, t, ctypes.c_uint32, np.uint32, owner)
    hashes = _pinned_array(v.hashes, n, ctypes.c_uint64, np.uint64, owner) if v.hashes else None
    return DetectorErrorModel(probs, offsets, targets, v.num_detectors, v.num_observables, hashes)


class Driver:
    """A DEM compiler whose GPU buffers are sized once and reused
    (Driver.hpp:37-219)."""

    def __init__(self, circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2,
                 device: int = 0):
        self._lib = _native.load()
        autotune.install_cached()  # tuned knobs per size class, once per process
        self._h = ctypes.c_void_p()
        self.correlation_level = CorrelationLevel(correlation_level)
        self.device = device
        _native.check(self._lib.gp_get_driver(circuit_text(circuit).encode(), int(correlation_level),
                                              device, ctypes.byref(self._h)))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.gp_driver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def compile_detector_error_model(self, circuit=None) -> DetectorErrorModel:
        """Compiles `circuit` (or the circuit parsed last) into its DEM."""
        text = None if circuit is None else circuit_text(circuit).encode()
        _native.check(self._lib.gp_compile_detector_error_model(self._h, text))
        return self.last_dem()

    def last_dem(self) -> DetectorErrorModel:
        """The last compile's DEM; its arrays are the pinned host memory the
        driver downloaded into (no copy), kept alive by the arrays."""
        return _view_dem(self._lib, self._h)

    def last_hashes(self) -> np.ndarray:
        """The 64-bit class hash of each error of the last compile, in output
        order (the sort key), read from the device."""
        ne, nt, nd, no = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        _native.check(self._lib.gp_dem_shape(self._h, ctypes.byref(ne), ctypes.byref(nt), ctypes.byref(nd),
                                             ctypes.byref(no)))
        h = np.empty(max(ne.value, 1), np.uint64)
        _native.check(self._lib.gp_dem_copy(self._h, None, None, None, h.ctypes.data))
        return h[: ne.value]

    def last_dem_copy(self) -> DetectorErrorModel:
        """The last compile's DEM copied into fresh numpy arrays."""
        return _read_dem(self._lib, self._lib.gp_dem_shape, self._lib.gp_dem_copy, self._h)

    def dem_text(self) -> str:
        n = ctypes.c_size_t()
        _native.check(self._lib.gp_dem_text(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        _native.check(self._lib.gp_dem_text(self._h, buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def parse(self, circuit) -> None:
        _native.check(self._lib.gp_driver_parse(self._h, circuit_text(circuit).encode()))

    @property
    def params(self) -> dict:
        a = np.zeros(7, np.uint32)
        _native.check(self._lib.gp_driver_params(self._h, a.ctypes.data))
        keys = ("num_qubits", "num_layers", "num_measurements", "num_detectors",
                "num_observables", "num_nodes", "num_locations")
        return {k: int(v) for k, v in zip(keys, a)}

    def timings(self) -> dict:
        """Per-stage microseconds of the last compile (HIP events)."""
        a = np.zeros(8, np.float64)
        _native.check(self._lib.gp_driver_timings(self._h, a.ctypes.data))
        keys = ("upload", "sweep", "classify", "partition", "buckets", "emit", "download", "total")
        return {k: float(v) for k, v in zip(keys, a)}

    def bench_device(self, iterations: int = 10) -> float:
        """Mean device microseconds per compile of the loaded circuit."""
        us = ctypes.c_double()
        _native.check(self._lib.gp_driver_bench_device(self._h, iterations, ctypes.byref(us)))
        return us.value

    def debug_sweep_plan(self) -> tuple[int, int, int]:
        """(columns swept in time chunks, extra first-launch sweep blocks,
        second-launch blocks) of this driver's current circuit."""
        plan = (ctypes.c_uint32 * 3)()
        _native.check(self._lib.gp_debug_sweep_plan(self._h, plan))
        return tuple(int(x) for x in plan)

    def debug_sensitivities(self) -> tuple[np.ndarray, np.ndarray]:
        """(node index of each p>0 node, its sensitivity words) from the GPU."""
        n, w = ctypes.c_uint32(), ctypes.c_uint32()
        _native.check(self._lib.gp_debug_pnodes(self._h, ctypes.byref(n), ctypes.byref(w)))
        idx = np.empty(max(n.value, 1), np.uint32)
        sens = np.empty((max(n.value, 1), max(w.value, 1)), np.uint64)
        _native.check(self._lib.gp_debug_sensitivities(self._h, idx.ctypes.data, sens.ctypes.data))
        return idx[: n.value], sens[: n.value, : w.value]

    def debug_stepg(self) -> tuple[np.ndarray, np.ndarray]:
        nodes = self.params["num_nodes"]
        adj = np.empty(max(nodes, 1), np.uint64)
        probs = np.empty(max(nodes, 1), np.float64)
        _native.check(self._lib.gp_debug_stepg(self._h, adj.ctypes.data, probs.ctypes.data))
        return adj[:nodes], probs[:nodes]


def get_driver(circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2,
               device: int = 0) -> Driver:
    """Driver sized for `circuit` (ErrorAnalysis.pybind.cu:72-78)."""
    return Driver(circuit, correlation_level, device)


def build_stepg_host(circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2):
    """Host-only STEPG of `circuit`: (adjacency, probs, dims) -- touches no GPU."""
    lib = _native.load()
    text = circuit_text(circuit).encode()
    dims = np.zeros(4, np.uint32)
    _native.check(lib.gp_build_stepg_host(text, int(correlation_level), dims.ctypes.data, None, None))
    n = int(dims[2])
    adj = np.empty(max(n, 1), np.uint64)
    probs = np.empty(max(n, 1), np.float64)
    _native.check(lib.gp_build_stepg_host(text, int(correlation_level), dims.ctypes.data,
                                          adj.ctypes.data, probs.ctypes.data))
    return adj[:n], probs[:n], dims


def estimate_device_bytes(circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2) -> int:
    lib = _native.load()
    b = ctypes.c_uint64()
    _native.check(lib.gp_estimate_device_bytes(circuit_text(circuit).encode(), int(correlation_level),
                                               ctypes.byref(b)))
    return int(b.value)


__all__ = ["CorrelationLevel", "Driver", "get_driver", "build_stepg_host", "estimate_device_bytes",
           "circuit_text"]
