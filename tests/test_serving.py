"""Serving loop host logic on the CPU: admission by estimated HBM footprint,
continuous batching across slots, rejection of oversized requests and error
delivery, with a fake compiler standing in for the GPU."""

from __future__ import annotations

import threading
import time

import pytest

from conftest import read_circuit
from greenpeas_amd import serving
from greenpeas_amd.dem import DetectorErrorModel
from greenpeas_amd.error_analysis import estimate_device_bytes


class FakeCompiler:
    """Records batches; 'compiles' by returning the text length."""

    log: list = []
    lock = threading.Lock()
    delay = 0.02

    def set_circuits(self, texts, threads=0):
        self.texts = list(texts)

    def compile(self):
        with FakeCompiler.lock:
            FakeCompiler.log.append(list(self.texts))
        time.sleep(FakeCompiler.delay)
        if any("FAIL" in t for t in self.texts):
            raise RuntimeError("planted failure")
        return [len(t) for t in self.texts]

    def device_us(self):
        return 1.0


def test_estimate_is_an_upper_bound_of_the_real_footprint():
    for name in ("surface3", "surface5", "bb6", "bb12"):
        text = read_circuit(name)
        assert serving.estimate_bytes(text) >= 0.8 * estimate_device_bytes(text)


def test_continuous_batching_and_admission():
    FakeCompiler.log = []
    texts = [f"R 0\nTICK\nM {i}\n" for i in range(40)]
    one = serving.estimate_bytes(texts[0])
    srv = serving.Server(slots=2, max_batch=8, hbm_budget=one * 10, compiler_factory=FakeCompiler)
    futs = [srv.submit(t) for t in texts]
    assert [f.result(timeout=10) for f in futs] == [len(t) for t in texts]
    srv.close()
    sizes = [len(b) for b in FakeCompiler.log]
    assert sum(sizes) == 40
    assert max(sizes) <= 8
    # the two slots never hold more than the budget together
    assert all(s <= 10 for s in sizes)
    st = srv.stats()
    assert st["requests"] == 40 and st["p99_ms"] >= st["p50_ms"] > 0


def test_oversized_request_rejected():
    srv = serving.Server(slots=1, hbm_budget=1000, compiler_factory=FakeCompiler)
    with pytest.raises(MemoryError):
        srv.submit(read_circuit("bb12"))
    srv.close()


def test_errors_reach_every_future_of_the_batch():
    FakeCompiler.log = []
    srv = serving.Server(slots=1, max_batch=4, compiler_factory=FakeCompiler)
    FakeCompiler.delay = 0.2
    first = srv.submit("R 0\nTICK\nM 0\n")  # occupies the slot
    time.sleep(0.05)
    f1 = srv.submit("R 0\nTICK\nM 0\n# FAIL\n")
    f2 = srv.submit("R 1\nTICK\nM 1\n")
    assert first.result(timeout=5) > 0
    with pytest.raises(RuntimeError, match="planted"):
        f1.result(timeout=5)
    with pytest.raises(RuntimeError, match="planted"):
        f2.result(timeout=5)
    FakeCompiler.delay = 0.02
    srv.close()


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
    srv = serving.Server(slots=3, compiler_factory=FakeCompiler, warmup=["R 0\nTICK\nM 0\n"])
    # the constructor returns once every slot has compiled the warm-up batch
    assert FakeCompiler.log == [["R 0\nTICK\nM 0\n"]] * 3
    assert srv.submit("R 1\nTICK\nM 1\n").result(timeout=10) == len("R 1\nTICK\nM 1\n")
    srv.close()


def test_warmup_failure_is_raised_by_the_constructor():
    with pytest.raises(RuntimeError, match="warm-up"):
        serving.Server(slots=2, compiler_factory=FakeCompiler, warmup=["FAIL"])
