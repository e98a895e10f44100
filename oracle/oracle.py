"""ctypes view of the CPU oracle (oracle/gp_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the greenpeas_amd package.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libgp_oracle.so")
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: run `make -C {_HERE}`")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.gpo_last_error.restype = ctypes.c_char_p
        L.gpo_compile_stepg.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(P)]
        L.gpo_compile_stepg_theta.argtypes = [ctypes.c_char_p, ctypes.c_int, P, ctypes.c_uint32, ctypes.POINTER(P)]
        L.gpo_compile_frames.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
        L.gpo_dem_free.argtypes = [P]
        for f in ("gpo_dem_num_errors", "gpo_dem_num_detectors", "gpo_dem_num_observables"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = ctypes.c_uint32
        L.gpo_dem_num_targets.argtypes = [P]
        L.gpo_dem_num_targets.restype = ctypes.c_uint64
        L.gpo_dem_copy.argtypes = [P, P, P, P]
        L.gpo_dem_hashes.argtypes = [P, P]
        L.gpo_build_stepg.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(P)]
        L.gpo_stepg_free.argtypes = [P]
        L.gpo_stepg_dims.argtypes = [P, P]
        L.gpo_stepg_copy.argtypes = [P, P, P, P, P]
        L.gpo_stepg_sweep.argtypes = [P, P]
        L.gpo_fnv1a64_slots.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32]
        L.gpo_fnv1a64_slots.restype = ctypes.c_uint64
        for f in ("gpo_split_dep1", "gpo_split_dep2"):
            getattr(L, f).argtypes = [ctypes.c_double]
            getattr(L, f).restype = ctypes.c_double
        L.gpo_xor_merge.argtypes = [ctypes.c_double, ctypes.c_double]
        L.gpo_xor_merge.restype = ctypes.c_double
        L.gpo_location_count.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)]
        L.gpo_logit_loss_grad.argtypes = [ctypes.c_char_p, ctypes.c_int, P, ctypes.c_uint32, P,
                                          ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), P,
                                          ctypes.c_uint32]
        _lib = L
    return _lib


class OracleError(Exception):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def _check(rc: int) -> None:
    if rc:
        raise OracleError(rc, lib().gpo_last_error().decode())


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class OracleDem:
    """An oracle DEM as CSR arrays, in the compiler's output order."""

    probs: np.ndarray
    offsets: np.ndarray
    targets: np.ndarray
    num_detectors: int
    num_observables: int
    hashes: np.ndarray | None = None

    @property
    def num_errors(self) -> int:
        return len(self.probs)

    def errors(self):
        for i in range(self.num_errors):
            yield float(self.probs[i]), tuple(int(t) for t in self.targets[self.offsets[i]:self.offsets[i + 1]])


def _take_dem(h: ctypes.c_void_p, with_hash: bool) -> OracleDem:
    L = lib()
    n = L.gpo_dem_num_errors(h)
    nt = L.gpo_dem_num_targets(h)
    p = np.empty(n, np.float64)
    off = np.empty(n + 1, np.uint32)
    tg = np.empty(max(nt, 1), np.uint32)
    L.gpo_dem_copy(h, _ptr(p), _ptr(off), _ptr(tg))
    hashes = None
    if with_hash:
        hashes = np.empty(max(n, 1), np.uint64)
        L.gpo_dem_hashes(h, _ptr(hashes))
        hashes = hashes[:n]
    d = OracleDem(p, off, tg[:nt], L.gpo_dem_num_detectors(h), L.gpo_dem_num_observables(h), hashes)
    L.gpo_dem_free(h)
    return d


def compile_stepg(text: str, level: int = 2, theta: np.ndarray | None = None) -> OracleDem:
    h = ctypes.c_void_p()
    if theta is None:
        _check(lib().gpo_compile_stepg(text.encode(), level, ctypes.byref(h)))
    else:
        th = np.ascontiguousarray(theta, np.float64)
        _check(lib().gpo_compile_stepg_theta(text.encode(), level, _ptr(th), len(th), ctypes.byref(h)))
    return _take_dem(h, True)


def compile_frames(text: str) -> OracleDem:
    h = ctypes.c_void_p()
    _check(lib().gpo_compile_frames(text.encode(), ctypes.byref(h)))
    return _take_dem(h, False)


@dataclass
class OracleStepg:
    layers: int
    qubits: int
    nodes_per_qubit: int
    nodes: int
    measurements: int
    detectors: int
    observables: int
    words: int
    adjacency: np.ndarray
    probs: np.ndarray
    keys: np.ndarray
    vals: np.ndarray  # measurements x words, row-major
    _handle: object = None

    def sweep(self) -> np.ndarray:
        """Dense sensitivity matrix, nodes x words."""
        S = np.empty((self.nodes, max(self.words, 1)), np.uint64)
        lib().gpo_stepg_sweep(self._handle, _ptr(S))
        return S[:, : self.words]

    def __del__(self):
        if self._handle is not None:
            lib().gpo_stepg_free(self._handle)
            self._handle = None


def build_stepg(text: str, level: int = 2) -> OracleStepg:
    L = lib()
    h = ctypes.c_void_p()
    _check(L.gpo_build_stepg(text.encode(), level, ctypes.byref(h)))
    dims = np.empty(8, np.uint32)
    L.gpo_stepg_dims(h, _ptr(dims))
    nl, nq, k, nn, M, D, O, W = (int(x) for x in dims)
    adj = np.empty(nn, np.uint64)
    probs = np.empty(nn, np.float64)
    keys = np.empty(max(M, 1), np.uint32)
    vals = np.empty((max(M, 1), max(W, 1)), np.uint64)
    L.gpo_stepg_copy(h, _ptr(adj), _ptr(probs), _ptr(keys), _ptr(vals))
    return OracleStepg(nl, nq, k, nn, M, D, O, W, adj, probs, keys[:M], vals[:M, :W], h)


def fnv1a64_slots(vals, pad_to: int = 32) -> int:
    a = np.ascontiguousarray(vals, np.uint32)
    return int(lib().gpo_fnv1a64_slots(_ptr(a) if len(a) else None, len(a), pad_to))


def split_dep1(p: float) -> float:
    return lib().gpo_split_dep1(p)


def split_dep2(p: float) -> float:
    return lib().gpo_split_dep2(p)


def xor_merge(p_old: float, p_new: float) -> float:
    return lib().gpo_xor_merge(p_old, p_new)


def location_count(text: str) -> int:
    n = ctypes.c_uint32()
    _check(lib().gpo_location_count(text.encode(), ctypes.byref(n)))
    return n.value


def logit_loss_grad(text: str, theta: np.ndarray, w: np.ndarray, level: int = 2):
    th = np.ascontiguousarray(theta, np.float64)
    ww = np.ascontiguousarray(w, np.float64)
    g = np.empty(len(th), np.float64)
    loss = ctypes.c_double()
    _check(lib().gpo_logit_loss_grad(text.encode(), level, _ptr(th), len(th), _ptr(ww), len(ww),
                                     ctypes.byref(loss), _ptr(g), len(g)))
    return loss.value, g
