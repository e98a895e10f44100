"""Serving loop: continuous batching of DEM compile requests on one GPU.

No counterpart in the reference (which compiles synchronously, one circuit
per call).  A request passes two stages:

1. prepare -- on a pool of `prepare_threads` host threads the circuit text
   is parsed and built into its device image (gp_prepare_circuit), and the
   device bytes a new batch needs for it alone are known exactly
   (gp_prepared_device_bytes).  A request needing more than the whole
   budget fails here with MemoryError.
2. compile -- `slots` worker threads each own a batched compiler (its own
   HIP stream and device buffers).  An idle slot admits prepared requests,
   oldest first, while the batch holds fewer than `max_batch` requests and
   the bytes every other slot holds (or has reserved for its running batch)
   plus the bytes this slot will hold after compiling the batch
   (gp_batch_bytes_after: its buffers where they fit, grown where not; an
   upper bound of its peak too) stay within the budget.  Then it compiles
   them in one batched pipeline and completes their futures.

So host builds of later requests overlap the device work of earlier batches,
and batches on different slots overlap on the device.  The budget is
`hbm_budget` bytes or, by default, the device's free memory once the slots
have warmed up plus what they hold, less a margin of 5 % of the device (at
least 1 GiB).  A slot whose own buffers keep the oldest request out while
no other slot holds anything frees them (a new compiler) and admits again.

Per-request results are independent of what they are batched with: each
circuit owns its node range, column work items and sort buckets, and the
merge order within a circuit does not depend on the batch.

`warmup` circuits (e.g. the largest request expected) are compiled by every
slot before the server takes requests, so device buffers are sized and the
kernels loaded before the first request's latency is counted.
"""

from __future__ import annotations

import heapq
import itertools
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
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


@dataclass
class Request:
    text: str
    seq: int
    t_submit: float
    future: Future = field(default_factory=Future)
    prepared: object = None
    nbytes: int = 0  # device bytes a new batch needs for this request alone
    t_ready: float = 0.0
    t_start: float = 0.0
    t_done: float = 0.0

    def __lt__(self, other: "Request") -> bool:
        return self.seq < other.seq


class Server:
    """Continuous-batching DEM compile server (see the module docstring).

    compiler_factory() makes a slot's compiler (default: a BatchCompiler on
    `device`): set_prepared(list), compile() -> results in order,
    device_us(), device_bytes(), bytes_after(list), close().
    preparer(text) makes a prepared request (default: batch.prepare at the
    server's level): device_bytes(), close()."""

    def __init__(self, correlation_level: CorrelationLevel = CorrelationLevel.L2, device: int = 0,
                 slots: int = 2, max_batch: int = 64, hbm_budget: int | None = None,
                 prepare_threads: int = 8, compiler_factory=None, preparer=None, warmup=()):
        self.level = int(correlation_level)
        self.slots = slots
        self.max_batch = max_batch
        on_device = compiler_factory is None
        if compiler_factory is None:
            from .batch import BatchCompiler

            def compiler_factory():
                return BatchCompiler(CorrelationLevel(self.level), device)
        if preparer is None:
            from .batch import prepare

            def preparer(text):
                return prepare(text, CorrelationLevel(self.level))
        self._factory = compiler_factory
        self._preparer = preparer
        self._cv = threading.Condition()
        self._ready: list[Request] = []  # prepared, not yet admitted (a heap by submission order)
        self._held = [0] * slots  # device bytes each slot holds or has reserved
        self._preparing = 0
        self._seq = itertools.count()
        self._closed = False
        self.hbm_budget = hbm_budget if hbm_budget is not None else 1 << 62
        self.completed: list[Request] = []
        self.batches: list[tuple[int, int, float]] = []  # (slot, size, device us)
        self._warmup = [circuit_text(c) for c in warmup]
        self._warmup_errors: list[BaseException] = []
        self._started = threading.Barrier(slots + 1)
        self._go = threading.Event()
        self._pool = ThreadPoolExecutor(max(1, prepare_threads), thread_name_prefix="gp-prepare")
        self._workers = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                         for i in range(slots)]
        for w in self._workers:
            w.start()
        self._started.wait()  # every slot warmed up (or failed to)
        if self._warmup_errors:
            self.close()
            raise RuntimeError("server warm-up failed") from self._warmup_errors[0]
        if hbm_budget is None and on_device:
            from .batch import device_memory

            free, total = device_memory(device)
            self.hbm_budget = free + sum(self._held) - max(1 << 30, total // 20)
        self._go.set()

    # ------------------------------------------------------------------
    def submit(self, circuit) -> Future:
        """Queues a circuit; its future gets the DEM, or the error that its
        build or compile raised (MemoryError: larger than the budget)."""
        text = circuit_text(circuit)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            req = Request(text, next(self._seq), time.perf_counter())
            self._preparing += 1
        self._pool.submit(self._prepare, req)
        return req.future

    def _prepare(self, req: Request) -> None:
        err = None
        try:
            req.prepared = self._preparer(req.text)
            req.nbytes = int(req.prepared.device_bytes())
            if req.nbytes > self.hbm_budget:
                err = MemoryError(f"request needs {req.nbytes} device bytes, the budget is {self.hbm_budget}")
        except BaseException as e:  # noqa: BLE001 -- delivered to the future
            err = e
        req.t_ready = time.perf_counter()
        with self._cv:
            self._preparing -= 1
            if err is None:
                heapq.heappush(self._ready, req)
            self._cv.notify_all()
        if err is not None:
            if req.prepared is not None:
                req.prepared.close()
            req.future.set_exception(err)

    def _admit(self, slot: int, compiler) -> list[Request] | None:
        """Oldest-first admission under the batch-size and memory limits
        (under the lock).  None: this slot's own buffers keep the oldest
        request out while no other slot holds anything."""
        others = sum(h for t, h in enumerate(self._held) if t != slot)
        batch: list[Request] = []
        after = self._held[slot]
        while self._ready and len(batch) < self.max_batch:
            r = self._ready[0]
            need = compiler.bytes_after([x.prepared for x in batch] + [r.prepared])
            if others + need > self.hbm_budget:
                if not batch and others == 0 and self._held[slot] > 0:
                    return None
                break
            heapq.heappop(self._ready)
            batch.append(r)
            after = need
        if batch:
            self._held[slot] = max(self._held[slot], after)
        return batch

    def _worker(self, slot: int) -> None:
        compiler = None
        try:
            compiler = self._factory()
            if self._warmup:
                prepared = [self._preparer(t) for t in self._warmup]
                compiler.set_prepared(prepared)
                compiler.compile()
                for p in prepared:
                    p.close()
            self._held[slot] = int(compiler.device_bytes())
        except Exception as e:  # noqa: BLE001 -- reported by the constructor
            with self._cv:
                self._warmup_errors.append(e)
            self._started.wait()
            return
        self._started.wait()
        self._go.wait()
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._ready and not self._preparing:
                        compiler.close()
                        return
                    batch = self._admit(slot, compiler)
                    if batch is None or batch:
                        break
                    self._cv.wait()
                if batch is None:
                    self._held[slot] = 0
            if batch is None:
                compiler.close()  # frees this slot's buffers
                compiler = self._factory()
                continue
            t0 = time.perf_counter()
            for r in batch:
                r.t_start = t0
            try:
                compiler.set_prepared([r.prepared for r in batch])
                dems = compiler.compile()
                dev_us = compiler.device_us()
                err = None
            except Exception as e:  # noqa: BLE001 -- delivered to every future
                dems, dev_us, err = None, 0.0, e
            held = int(compiler.device_bytes())
            t1 = time.perf_counter()
            with self._cv:
                self._held[slot] = held
                self.batches.append((slot, len(batch), dev_us))
                for r in batch:
                    r.t_done = t1
                    self.completed.append(r)
                self._cv.notify_all()
            for i, r in enumerate(batch):
                r.prepared.close()
                r.prepared = None
                if err is None:
                    r.future.set_result(dems[i])
                else:
                    r.future.set_exception(err)

    # ------------------------------------------------------------------
    def close(self, wait: bool = True) -> None:
        """Stops taking requests; the queued ones are still compiled."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        self._go.set()
        if wait:
            for w in self._workers:
                w.join()
        self._pool.shutdown(wait=wait)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def held_bytes(self) -> list[int]:
        """Device bytes each slot holds (or has reserved for its batch)."""
        with self._cv:
            return list(self._held)

    def stats(self) -> dict:
        done = list(self.completed)
        lat = np.array([r.t_done - r.t_submit for r in done]) * 1e3
        if not len(lat):
            return {"requests": 0}
        prep = np.array([r.t_ready - r.t_submit for r in done]) * 1e3
        queue = np.array([r.t_start - r.t_ready for r in done]) * 1e3
        span = max(r.t_done for r in done) - min(r.t_submit for r in done)
        return {
            "requests": int(len(lat)),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()),
            "prepare_p50_ms": float(np.percentile(prep, 50)),
            "queue_p50_ms": float(np.percentile(queue, 50)),
            "throughput_rps": float(len(lat) / span) if span > 0 else float("inf"),
            "batches": len(self.batches),
            "mean_batch": float(np.mean([b[1] for b in self.batches])),
            "budget_bytes": int(self.hbm_budget),
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
