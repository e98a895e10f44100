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
