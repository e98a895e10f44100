"""The serving loop on the GPU: requests submitted concurrently from several
threads while batches run; each result equals the request compiled alone."""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from greenpeas_amd import get_driver
from greenpeas_amd.serving import Server, request_pool

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix():
    pool = request_pool(12, seed=5, surface=(5, 9), bb=(6,), bb_rounds=(6, 8))
    alone = [get_driver(t).compile_detector_error_model(t) for t in pool]
    srv = Server(slots=2, max_batch=8)
    futures = [None] * 36
    rng = np.random.default_rng(0)
    order = rng.integers(0, len(pool), len(futures))

    def client(ix):
        for i in ix:
            time.sleep(float(rng.uniform(0, 0.003)))
            futures[i] = srv.submit(pool[order[i]])

    threads = [threading.Thread(target=client, args=(range(k, len(futures), 3),)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i, f in enumerate(futures):
        assert f.result(timeout=60) == alone[order[i]]
    srv.close()
    st = srv.stats()
    assert st["requests"] == len(futures)
    assert st["batches"] < len(futures)  # requests were batched together


def test_config5_poisson_stream_over_the_full_pool():
    """The serving workload at its stated size: the bench's seeded pool
    (surface d in 5..21 with rounds = d, BB [[72,12,6]] and [[144,12,12]]
    with rounds 6..18, one rate per request logU[1e-4, 1e-2]) driven by a
    seeded Poisson stream at 1000 requests/s.  Every result equals the
    request compiled alone, bit for bit."""
    from greenpeas_amd.serving import poisson_arrivals

    pool = request_pool(48, seed=77)
    summ = [t.split("\n", 1)[0] for t in pool]
    assert any(len(s.split()) - 1 == 21 * 21 for s in summ)  # a d=21 surface memory
    assert any(len(s.split()) - 1 == 144 for s in summ)       # a [[144,12,12]] memory
    alone = []
    for t in pool:
        drv = get_driver(t)
        alone.append(drv.compile_detector_error_model(t))
        drv.close()
    arrivals = poisson_arrivals(1000, 200, seed=1000)
    pick = np.random.default_rng(1000).integers(0, len(pool), len(arrivals))
    srv = Server(slots=2, max_batch=64, warmup=[max(pool, key=len)])
    futs = []
    t0 = time.perf_counter()
    for i, at in enumerate(arrivals):
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        futs.append(srv.submit(pool[pick[i]]))
    for i, f in enumerate(futs):
        assert f.result(timeout=300) == alone[pick[i]], f"request {i} (pool entry {pick[i]}) differs"
    srv.close()
    st = srv.stats()
    assert st["requests"] == len(arrivals)
    assert st["batches"] < len(arrivals)


def test_footprint_bounds_what_the_batch_holds():
    """Admission's footprint over the serving pool (surface d 5..21, bb6 and
    bb12): a new batch context's figure for one circuit equals the prepared
    circuit's own, and over a seeded sequence of batches of 1..24 circuits
    the bytes predicted before each compile bound what the context holds
    after it (gp_batch_device_bytes), and equal it when nothing grew."""
    from greenpeas_amd.batch import BatchCompiler, prepare

    pool = request_pool(48, seed=77)
    prepared = [prepare(t) for t in pool]
    bc = BatchCompiler()
    assert bc.bytes_after(prepared[:1]) == prepared[0].device_bytes()
    rng = np.random.default_rng(11)
    grew = 0
    for step in range(12):
        pick = rng.choice(len(pool), int(rng.integers(1, 25)), replace=False)
        batch = [prepared[i] for i in pick]
        before = bc.device_bytes()
        est = bc.bytes_after(batch)
        assert est >= before
        bc.set_prepared(batch)
        bc.run()
        held = bc.device_bytes()
        print(f"batch {step}: {len(batch)} circuits, held {before >> 20} -> {held >> 20} MiB, "
              f"predicted {est >> 20} MiB", flush=True)
        assert held <= est
        if est == before:
            assert held == before
        grew += held > before
    assert grew >= 1
    bc.close()


def test_prepared_circuits_compile_as_their_text():
    from greenpeas_amd.batch import BatchCompiler, prepare

    pool = request_pool(6, seed=3, surface=(5, 7), bb=(6,), bb_rounds=(6, 7))
    a = BatchCompiler()
    a.set_circuits(pool)
    want = a.compile()
    b = BatchCompiler()
    prepared = [prepare(t) for t in pool]
    b.set_prepared(prepared)
    for p in prepared:
        p.close()  # the batch keeps its own reference to each image
    assert b.compile() == want
    with pytest.raises(ValueError, match="correlation level"):
        b.set_prepared([prepare(pool[0], 1)])
    a.close()
    b.close()


def test_default_budget_comes_from_device_memory():
    from greenpeas_amd.batch import device_memory

    free, total = device_memory(0)
    assert 0 < free <= total
    with Server(slots=2, warmup=["R 0\nTICK\nM 0\nDETECTOR rec[-1]\n"]) as srv:
        held = sum(srv.held_bytes())
        assert held > 0
        # free memory at start-up, less the margin: what the slots hold
        # counts inside the budget, not against it
        assert abs(srv.hbm_budget - (free - max(1 << 30, total // 20))) < (256 << 20)
