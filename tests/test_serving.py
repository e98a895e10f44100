"""Serving loop host logic on the CPU: admission by device footprint,
continuous batching across slots, rejection of oversized requests, error
delivery and warm-up, with a fake preparer and compiler standing in for the
library and the GPU."""

from __future__ import annotations

import threading
import time

import pytest

from conftest import read_circuit
from greenpeas_amd import serving


class FakePrepared:
    """A 'prepared' request: its text; it needs len(text) bytes alone."""

    closed = 0

    def __init__(self, text):
        if "BUILD_FAIL" in text:
            raise ValueError("planted build failure")
        self.text = text

    def device_bytes(self):
        return len(self.text)

    def close(self):
        FakePrepared.closed += 1


class FakeCompiler:
    """Records batches; 'compiles' by returning the text lengths.  A batch
    needs the sum of its requests' bytes; one that does not fit what the
    compiler holds adds that much to it (old buffers are kept)."""

    log: list = []
    lock = threading.Lock()
    delay = 0.02
    made = 0

    def __init__(self):
        self.held = 0
        self.closed = False
        with FakeCompiler.lock:
            FakeCompiler.made += 1

    def bytes_after(self, prepared):
        need = sum(p.device_bytes() for p in prepared)
        return self.held + need if need > self.held else self.held

    def set_prepared(self, prepared):
        self.texts = [p.text for p in prepared]
        self.need = sum(len(t) for t in self.texts)

    def compile(self):
        with FakeCompiler.lock:
            FakeCompiler.log.append(list(self.texts))
        time.sleep(FakeCompiler.delay)
        if self.need > self.held:
            self.held += self.need
        if any("FAIL" in t for t in self.texts):
            raise RuntimeError("planted failure")
        return [len(t) for t in self.texts]

    def device_us(self):
        return 1.0

    def device_bytes(self):
        return self.held

    def close(self):
        self.closed = True


def server(**kw):
    return serving.Server(compiler_factory=FakeCompiler, preparer=FakePrepared, **kw)


def test_continuous_batching_and_admission():
    FakeCompiler.log = []
    texts = [f"R 0\nTICK\nM {i}\n" for i in range(40)]
    one = len(texts[0])
    srv = server(slots=2, max_batch=8, hbm_budget=one * 10)
    held_seen = []
    futs = [srv.submit(t) for t in texts]
    while not all(f.done() for f in futs):
        held_seen.append(sum(srv.held_bytes()))
        time.sleep(0.002)
    assert [f.result(timeout=10) for f in futs] == [len(t) for t in texts]
    srv.close()
    sizes = [len(b) for b in FakeCompiler.log]
    assert sum(sizes) == 40
    assert max(sizes) <= 8
    # what the slots hold and reserve together never exceeds the budget
    assert max(held_seen) <= one * 10
    st = srv.stats()
    assert st["requests"] == 40 and st["p99_ms"] >= st["p50_ms"] > 0
    assert st["budget_bytes"] == one * 10


def test_oversized_request_fails_its_future():
    srv = server(slots=1, hbm_budget=1000)
    fut = srv.submit(read_circuit("bb12"))
    with pytest.raises(MemoryError):
        fut.result(timeout=10)
    assert srv.submit("R 0\nTICK\nM 0\n").result(timeout=10) == len("R 0\nTICK\nM 0\n")
    srv.close()


def test_build_errors_reach_only_their_own_future():
    srv = server(slots=1, max_batch=4)
    bad = srv.submit("R 0\n# BUILD_FAIL\n")
    good = srv.submit("R 1\nTICK\nM 1\n")
    with pytest.raises(ValueError, match="planted build"):
        bad.result(timeout=5)
    assert good.result(timeout=5) == len("R 1\nTICK\nM 1\n")
    srv.close()


def test_errors_reach_every_future_of_the_batch():
    FakeCompiler.log = []
    srv = server(slots=1, max_batch=4)
    FakeCompiler.delay = 0.2
    try:
        first = srv.submit("R 0\nTICK\nM 0\n")  # occupies the slot
        time.sleep(0.05)
        f1 = srv.submit("R 0\nTICK\nM 0\n# FAIL\n")
        f2 = srv.submit("R 1\nTICK\nM 1\n")
        assert first.result(timeout=5) > 0
        with pytest.raises(RuntimeError, match="planted"):
            f1.result(timeout=5)
        with pytest.raises(RuntimeError, match="planted"):
            f2.result(timeout=5)
    finally:
        FakeCompiler.delay = 0.02
        srv.close()


def test_slot_frees_its_buffers_when_they_alone_block_a_request():
    """A slot that holds buffers from an earlier batch and cannot grow them
    within the budget for the oldest request, while no other slot holds
    anything, takes a new compiler (its buffers freed) and admits it."""
    FakeCompiler.made = 0
    srv = server(slots=1, max_batch=8, hbm_budget=100)
    first, second = "A" * 60, "B" * 70  # 60 + 70 > 100: no room beside the first's buffers
    assert srv.submit(first).result(timeout=5) == 60
    assert srv.held_bytes() == [60]
    assert srv.submit(second).result(timeout=5) == 70
    assert srv.held_bytes() == [70]
    srv.close()
    assert FakeCompiler.made == 2


def test_queued_requests_are_compiled_after_close():
    srv = server(slots=2, max_batch=4)
    futs = [srv.submit("R %d\nTICK\nM 0\n" % i) for i in range(12)]
    srv.close()
    assert all(f.done() and not f.exception() for f in futs)
    with pytest.raises(RuntimeError, match="closed"):
        srv.submit("R 0\n")


def test_request_pool_shapes():
    pool = serving.request_pool(6, seed=1, surface=(5, 7), bb=(6,), bb_rounds=(6, 7))
    assert len(pool) == 6 and all(t.startswith("R ") for t in pool)
    arr = serving.poisson_arrivals(1000.0, 100, 3)
    assert arr.shape == (100,) and (arr[1:] > arr[:-1]).all()


def test_circuit_summary_matches_the_parser():
    """The one-pass summary used for admission agrees with the full parse
    (layers, qubits via nodes per layer, detectors, observables), REPEAT
    bodies included."""
    from greenpeas_amd.error_analysis import build_stepg_host

    rep = "R 0 1 2\nTICK\nREPEAT 5 {\n  CX 0 1\n  TICK\n  M 1\n  DETECTOR rec[-1]\n}\nM 0 2\nOBSERVABLE_INCLUDE(3) rec[-1]\n"
    for text in [read_circuit(n) for n in ("surface3", "surface5", "bb6", "bb12")] + [rep]:
        layers, qubits, dets, obs = serving.circuit_summary(text)
        _, _, dims = build_stepg_host(text)
        assert layers == int(dims[0])
        assert qubits * 8 == int(dims[1])
        assert dets == text.count("DETECTOR") if "REPEAT" not in text else dets == 5
        lines = [ln for ln in text.splitlines() if ln.startswith("OBSERVABLE_INCLUDE(")]
        assert obs == max(int(ln.split("(")[1].split(")")[0]) for ln in lines) + 1
    assert serving.circuit_summary("X_ERROR(0.1) 7\nMPP X2*Z9\n# 100\n")[1] == 10


def test_warmup_runs_on_every_slot_before_requests():
    FakeCompiler.log = []
    srv = server(slots=3, warmup=["R 0\nTICK\nM 0\n"])
    # the constructor returns once every slot has compiled the warm-up batch
    assert FakeCompiler.log == [["R 0\nTICK\nM 0\n"]] * 3
    assert srv.submit("R 1\nTICK\nM 1\n").result(timeout=10) == len("R 1\nTICK\nM 1\n")
    srv.close()


def test_warmup_failure_is_raised_by_the_constructor():
    with pytest.raises(RuntimeError, match="warm-up"):
        server(slots=2, warmup=["FAIL"])
