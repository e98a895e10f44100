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
