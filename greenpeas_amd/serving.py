"""Serving loop: continuous batching of DEM compile requests on one GPU.

No counterpart in the reference (which compiles synchronously, one circuit
per call).  Requests arrive at any time; `slots` worker threads each own a
batched compiler with its own HIP stream and device buffers, so batches on
different slots overlap on the device.  Whenever a slot is idle it admits
the queued requests, oldest first, while

* the batch holds fewer than `max_batch` requests, and
* the estimated device footprint of everything admitted and still running,
  across all slots, stays within `hbm_budget` bytes,

then compiles them in one batched pipeline and completes their futures.
Requests that arrive while every slot is busy wait for the next free slot
(continuous batching, not a fixed batch queue).  A request larger than the
whole budget is rejected with `MemoryError` on submission.

Per-request results are independent of what they are batched with: each
circuit owns its node range, column work items and sort buckets, and the
merge order within a circuit does not depend on the batch.

`warmup` circuits (e.g. the largest request expected) are compiled by every
slot before the server takes requests, so device buffers are sized and the
kernels loaded before the first request's latency is counted.
"""

from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .error_analysis import CorrelationLevel, circuit_text


def circuit_summary(text: str) -> tuple[int, int, int, int]:
    """(layers, qubits, detectors, observables) read off circuit text in one
    native pass (nothing is parsed or built)."""
    out = np.zeros(4, np.uint64)
    _native.check(_native.load().gp_circuit_summary(text.encode(), out.ctypes.data))
    return tuple(int(x) for x in out)


def estimate_bytes(text: str, level: int = 2) -> int:
    """Cheap upper estimate of the device bytes one circuit needs in the
    batched pipeline (from circuit_summary; builds nothing)."""
    layers, qubits, dets, obs = circuit_summary(text)
    k = (3, 5, 8)[level]
    nodes = layers * qubits * k
    ranks = nodes * 3 // 5 + 64
    cols = (dets + obs + 31) // 32
    # sort buckets (64 nodes each at the default fill, 256 slots per bucket,
    # plus an arena of one slot per node), 40 bytes per slot
    buckets = 1
    while buckets * 64 < ranks:
        buckets <<= 1
    slots = buckets * 256 + ranks
    return int(ranks * (cols * 4 + 232) + slots * 40 + cols * 4096 + nodes * 24
               + layers * cols * 4 + 65536)


@dataclass
class Request:
    text: str
    nbytes: int
    t_submit: float
    future: Future = field(default_factory=Future)
    t_start: float = 0.0
    t_done: float = 0.0


class Server:
    """Continuous-batching DEM compile server (see the module docstring)."""

    def __init__(self, correlation_level: CorrelationLevel = CorrelationLevel.L2, device: int = 0,
                 slots: int = 2, max_batch: int = 64, hbm_budget: int = 64 << 30,
                 parse_threads: int = 8, compiler_factory=None, warmup=()):
        self.level = int(correlation_level)
        self.slots = slots
        self.max_batch = max_batch
        self.hbm_budget = hbm_budget
        self.parse_threads = parse_threads
        if compiler_factory is None:
            from .batch import BatchCompiler

            def compiler_factory():
                return BatchCompiler(CorrelationLevel(self.level), device)
        self._factory = compiler_factory
        self._queue: collections.deque[Request] = collections.deque()
        self._cv = threading.Condition()
        self._in_use = 0  # estimated bytes admitted and not yet finished
        self._closed = False
        self.completed: list[Request] = []
        self.batches: list[tuple[int, int, float]] = []  # (slot, size, device us)
        self._warmup = [circuit_text(c) for c in warmup]
        self._warmup_errors: list[BaseException] = []
        self._ready = threading.Barrier(slots + 1)
        self._workers = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                         for i in range(slots)]
        for w in self._workers:
            w.start()
        self._ready.wait()  # every slot warmed up (or failed to)
        if self._warmup_errors:
            self.close()
            raise RuntimeError("server warm-up failed") from self._warmup_errors[0]

    # ------------------------------------------------------------------
    def submit(self, circuit) -> Future:
        text = circuit_text(circuit)
        nbytes = estimate_bytes(text, self.level)
        if nbytes > self.hbm_budget:
            raise MemoryError(f"request needs ~{nbytes} device bytes, budget is {self.hbm_budget}")
        req = Request(text, nbytes, time.perf_counter())
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            self._queue.append(req)
            self._cv.notify_all()
        return req.future

    def _admit(self) -> list[Request]:
        """Oldest-first admission under the batch-size and HBM limits."""
        batch: list[Request] = []
        while self._queue and len(batch) < self.max_batch:
            r = self._queue[0]
            if self._in_use + r.nbytes > self.hbm_budget and (batch or self._in_use):
                break
            self._queue.popleft()
            self._in_use += r.nbytes
            batch.append(r)
        return batch

    def _worker(self, slot: int) -> None:
        try:
            compiler = self._factory()
            if self._warmup:
                compiler.set_circuits(self._warmup, threads=self.parse_threads)
                compiler.compile()
        except Exception as e:  # noqa: BLE001 -- reported by the constructor
            with self._cv:
                self._warmup_errors.append(e)
            self._ready.wait()
            return
        self._ready.wait()
        while True:
            with self._cv:
                while not self._closed and not self._queue:
                    self._cv.wait()
                if self._closed and not self._queue:
                    return
                batch = self._admit()
                if not batch:
                    self._cv.wait(0.001)
                    continue
            t0 = time.perf_counter()
            for r in batch:
                r.t_start = t0
            try:
                compiler.set_circuits([r.text for r in batch], threads=self.parse_threads)
                dems = compiler.compile()
                dev_us = compiler.device_us()
                err = None
            except Exception as e:  # noqa: BLE001 -- delivered to every future
                dems, dev_us, err = None, 0.0, e
            t1 = time.perf_counter()
            with self._cv:
                self._in_use -= sum(r.nbytes for r in batch)
                self.batches.append((slot, len(batch), dev_us))
                for i, r in enumerate(batch):
                    r.t_done = t1
                    self.completed.append(r)
                self._cv.notify_all()
            for i, r in enumerate(batch):
                if err is None:
                    r.future.set_result(dems[i])
                else:
                    r.future.set_exception(err)

    # ------------------------------------------------------------------
    def close(self, wait: bool = True) -> None:
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for w in self._workers:
                w.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self) -> dict:
        lat = np.array([r.t_done - r.t_submit for r in self.completed]) * 1e3
        if not len(lat):
            return {"requests": 0}
        span = max(r.t_done for r in self.completed) - min(r.t_submit for r in self.completed)
        return {
            "requests": int(len(lat)),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()),
            "throughput_rps": float(len(lat) / span) if span > 0 else float("inf"),
            "batches": len(self.batches),
            "mean_batch": float(np.mean([b[1] for b in self.batches])),
        }


def poisson_arrivals(rate: float, count: int, seed: int) -> np.ndarray:
    """Arrival offsets (seconds) of a seeded Poisson process."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))


def request_pool(count: int, seed: int, surface=(5, 21), bb=(6, 12), bb_rounds=(6, 18),
                 p_range=(1e-4, 1e-2)) -> list[str]:
    """The serving workload's request mix: rotated surface memories with
    d in [surface], rounds = d; BB [[72,12,6]] or [[144,12,12]] memories with
    rounds in bb_rounds; one noise rate per request, log-uniform in p_range."""
    from .codes import BBCode, SurfaceCode

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        p = float(np.exp(rng.uniform(np.log(p_range[0]), np.log(p_range[1]))))
        if rng.random() < 0.5:
            d = int(rng.integers(surface[0], surface[1] + 1))
            out.append(SurfaceCode(d).memory_text(d, p))
        else:
            d = int(rng.choice(bb))
            out.append(BBCode(d).memory_text(int(rng.integers(bb_rounds[0], bb_rounds[1] + 1)), p))
    return out
