"""The autotuner on the GPU: every value of the tuned knobs (radix digit
width, slots per bucket, sweep records per thread) leaves the DEMs bit for
bit as they are; size-class entries are applied when a circuit is loaded;
a tuning run keeps its best values for the circuit's size class."""

from __future__ import annotations

import pytest

from conftest import read_circuit
from greenpeas_amd import autotune, get_driver
from greenpeas_amd.batch import BatchCompiler, variant_circuits

pytestmark = pytest.mark.gpu

COMBOS = [  # (bucket_fill, bucket_cap, sweep_rpt)
    (24, 64, 1), (128, 64, 4), (64, 128, 2), (96, 256, 8), (128, 128, 1), (32, 256, 2)]


@pytest.fixture(autouse=True)
def _defaults():
    old = {k: autotune.get_tuning(k) for k in ("bucket_fill", "bucket_cap", "sweep_rpt")}
    autotune.clear_tuning_classes()
    yield
    for k, v in old.items():
        autotune.set_tuning(k, v)
    autotune.clear_tuning_classes()


def test_tuned_knobs_leave_dems_identical():
    texts = [read_circuit(n) for n in ("bb12", "surface5", "iceberg_surface_4_adaptive_circl")]
    variants = variant_circuits(8, seed=3, distances=(5, 7))
    want = [get_driver(t).compile_detector_error_model() for t in texts]
    bc = BatchCompiler()
    bc.set_circuits(variants)
    want_batch = bc.compile()
    bc.close()
    for fill, cap, rpt in COMBOS:
        print(f"bucket_fill {fill} bucket_cap {cap} sweep_rpt {rpt}", flush=True)
        autotune.set_tuning("bucket_fill", fill)
        autotune.set_tuning("bucket_cap", cap)
        autotune.set_tuning("sweep_rpt", rpt)
        for t, w in zip(texts, want):
            assert get_driver(t).compile_detector_error_model() == w
        bc = BatchCompiler()
        bc.set_circuits(variants)
        assert bc.compile() == want_batch
        st = bc.layout_stats()
        assert st["bucket_cap"] == cap
        bc.close()


def test_size_class_entries_apply_at_load():
    text = read_circuit("bb12")
    drv = get_driver(text)
    want = drv.compile_detector_error_model()
    cls = autotune.size_class(drv.params["num_nodes"])
    drv.close()
    bc = BatchCompiler()
    bc.set_circuits([text])
    assert bc.compile() == [want]
    base = bc.layout_stats()
    bc.close()
    assert base["bucket_cap"] == 256
    autotune.set_tuning_class(2, cls, "bucket_fill", 24)
    autotune.set_tuning_class(2, cls, "bucket_cap", 128)
    autotune.set_tuning_class(2, cls + 1, "bucket_cap", 64)  # another class: not taken
    bc = BatchCompiler()
    bc.set_circuits([text])
    assert bc.compile() == [want]
    st = bc.layout_stats()
    print(f"default {base}, tuned {st}", flush=True)
    assert st["bucket_cap"] == 128
    assert st["buckets"] > base["buckets"]  # a wider first digit
    bc.close()
    # a driver made now takes them too (same DEM)
    assert get_driver(text).compile_detector_error_model() == want


def test_tune_driver_keeps_the_best_for_the_size_class():
    text = read_circuit("bb6")
    want = get_driver(text).compile_detector_error_model()
    out = autotune.tune_driver(text, iterations=3, candidates=(32, 64), rpt_candidates=(1, 2),
                               cap_candidates=(128, 256), cache=False)
    print(out, flush=True)
    cls = out["size_class"]
    for name in ("sweep_rpt", "bucket_fill", "bucket_cap"):
        assert out[name]["best"] in out[name]["us"]
        assert autotune.get_tuning_class(2, cls, name) == out[name]["best"]
        assert all(us > 0 for us in out[name]["us"].values())
    assert get_driver(text).compile_detector_error_model() == want
