"""The autotuner's host side, on the CPU: per-size-class knob entries in the
library (set, read back, removed, validated) and the cache handed to the
library by install_cached."""

from __future__ import annotations

import json

import pytest

from greenpeas_amd import autotune


@pytest.fixture(autouse=True)
def _clean_classes():
    autotune.clear_tuning_classes()
    yield
    autotune.clear_tuning_classes()


def test_size_class_is_floor_log2():
    assert [autotune.size_class(n) for n in (1, 2, 3, 4, 1023, 1024, 419328)] == [0, 1, 1, 2, 9, 10, 18]


def test_class_entries_roundtrip_and_validation():
    assert autotune.get_tuning_class(2, 18, "bucket_fill") == 0
    autotune.set_tuning_class(2, 18, "bucket_fill", 96)
    autotune.set_tuning_class(2, 18, "bucket_cap", 128)
    autotune.set_tuning_class(2, 18, "sweep_rpt", 4)
    assert autotune.get_tuning_class(2, 18, "bucket_fill") == 96
    assert autotune.get_tuning_class(2, 18, "bucket_cap") == 128
    assert autotune.get_tuning_class(2, 18, "sweep_rpt") == 4
    # other levels and classes are untouched
    assert autotune.get_tuning_class(1, 18, "bucket_fill") == 0
    assert autotune.get_tuning_class(2, 17, "bucket_fill") == 0
    autotune.set_tuning_class(2, 18, "bucket_fill", 0)  # removes that knob only
    assert autotune.get_tuning_class(2, 18, "bucket_fill") == 0
    assert autotune.get_tuning_class(2, 18, "bucket_cap") == 128
    for name, bad in (("bucket_cap", 100), ("sweep_rpt", 3), ("bucket_fill", 5000), ("stash", 1)):
        with pytest.raises(ValueError):
            autotune.set_tuning_class(2, 18, name, bad)
    with pytest.raises(ValueError):
        autotune.set_tuning_class(3, 18, "bucket_fill", 64)
    with pytest.raises(ValueError):
        autotune.set_tuning_class(2, 64, "bucket_fill", 64)
    autotune.clear_tuning_classes()
    assert autotune.get_tuning_class(2, 18, "bucket_cap") == 0


def test_bucket_cap_knob():
    old = autotune.get_tuning("bucket_cap")
    assert old == 256
    try:
        for v in (64, 128, 256):
            autotune.set_tuning("bucket_cap", v)
            assert autotune.get_tuning("bucket_cap") == v
        with pytest.raises(ValueError):
            autotune.set_tuning("bucket_cap", 32)
    finally:
        autotune.set_tuning("bucket_cap", old)


def test_target_fill_knob():
    old = autotune.get_tuning("target_fill")
    assert old == 4
    try:
        for v in (1, 16, 4):
            autotune.set_tuning("target_fill", v)
            assert autotune.get_tuning("target_fill") == v
        for bad in (0, 17):
            with pytest.raises(ValueError):
                autotune.set_tuning("target_fill", bad)
    finally:
        autotune.set_tuning("target_fill", old)


def test_compact_items_knob():
    old = autotune.get_tuning("compact_items")
    assert old == 1
    try:
        for v in (0, 2, 1):
            autotune.set_tuning("compact_items", v)
            assert autotune.get_tuning("compact_items") == v
        with pytest.raises(ValueError):
            autotune.set_tuning("compact_items", 3)
    finally:
        autotune.set_tuning("compact_items", old)


def test_overflow_start_knob():
    old = autotune.get_tuning("overflow_start")
    assert old == 1 << 20
    try:
        for v in (1, 64, 1 << 24):
            autotune.set_tuning("overflow_start", v)
            assert autotune.get_tuning("overflow_start") == v
        with pytest.raises(ValueError):
            autotune.set_tuning("overflow_start", 0)
    finally:
        autotune.set_tuning("overflow_start", old)


def test_install_cached_hands_the_cache_to_the_library(tmp_path, monkeypatch):
    cache = tmp_path / "tune.json"
    cache.write_text(json.dumps({
        "L2:18": {"bucket_fill": 48, "sweep_rpt": 2, "bucket_cap": 128},
        "L1:12": 32,                       # an older entry: the fill only
        "junk": {"bucket_fill": 24},       # not a class key: skipped
    }))
    monkeypatch.setattr(autotune, "CACHE", str(cache))
    assert autotune.install_cached(force=True) == 2
    assert autotune.get_tuning_class(2, 18, "bucket_fill") == 48
    assert autotune.get_tuning_class(2, 18, "sweep_rpt") == 2
    assert autotune.get_tuning_class(2, 18, "bucket_cap") == 128
    assert autotune.get_tuning_class(1, 12, "bucket_fill") == 32
    assert autotune.apply_cached(1 << 18, 2) == 48
    assert autotune.apply_cached(1 << 10, 2) is None
