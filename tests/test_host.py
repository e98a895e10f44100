"""Host-side product logic (no GPU): the C-ABI library loads and exports
every declared symbol, the product's parser/STEPG builder matches the oracle
bit for bit, the DEM text format round-trips the golden files."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden_dem_path, read_circuit
from greenpeas_amd import _native
from greenpeas_amd.dem import DetectorErrorModel
from greenpeas_amd.error_analysis import build_stepg_host, estimate_device_bytes


def declared_symbols() -> list[str]:
    with open(os.path.join(ROOT, "include", "greenpeas_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    decl = declared_symbols()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(_native.EXPORTS)
    assert lib.gp_version().startswith(b"greenpeas_amd")


@pytest.mark.parametrize("name", ["surface3", "surface5", "bb6", "bb12"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_product_stepg_matches_oracle(oracle, name, level):
    text = read_circuit(name)
    adj, probs, dims = build_stepg_host(text, level)
    s = oracle.build_stepg(text, level)
    assert np.array_equal(adj, s.adjacency)
    assert np.array_equal(probs, s.probs)
    assert dims[3] == oracle.location_count(text)


def test_product_parser_errors():
    with pytest.raises(ValueError, match="Unsupported operation: I"):
        build_stepg_host("I 0\n")
    with pytest.raises(ValueError):
        build_stepg_host("DETECTOR rec[-1]\n")  # record before any measurement
    with pytest.raises(RuntimeError, match="DEP1"):
        build_stepg_host("R 0\nTICK\nDEPOLARIZE1(0.8) 0\n")
    with pytest.raises(RuntimeError, match="XX error in layer l < 1"):
        build_stepg_host("DEPOLARIZE2(0.01) 0 1\nTICK\n", 0)
    with pytest.raises(ValueError):
        build_stepg_host("R 0\n", 3)


def test_product_repeat_and_aliases(oracle):
    text = "RZ 0 1\nREPEAT 3 {\nTICK\nDEPOLARIZE1(0.01) 0 1\nCNOT 0 1\n}\nTICK\nMR(0.01) 0 1\n"
    adj, probs, _ = build_stepg_host(text)
    s = oracle.build_stepg(text, 2)
    assert np.array_equal(adj, s.adjacency) and np.array_equal(probs, s.probs)


@pytest.mark.parametrize("name", ["surface3", "surface5"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_dem_text_round_trip(name, level):
    path = golden_dem_path(name, level)
    d = DetectorErrorModel.from_file(path, 24 if name == "surface3" else 120, 1)
    again = DetectorErrorModel.from_text(str(d), d.num_detectors, d.num_observables)
    assert again == d
    assert d.canonical().approx_equals(d)


def test_estimate_device_bytes_grows_with_circuit():
    a = estimate_device_bytes(read_circuit("surface3"))
    b = estimate_device_bytes(read_circuit("bb12"))
    assert 0 < a < b
    # whole 2 MiB allocation units
    assert a % (2 << 20) == 0 and b % (2 << 20) == 0


def test_prepared_circuit_estimate_matches_the_text_estimate():
    """A prepared circuit (parsed and built on the host, no device) reports
    the same footprint as the text estimate, for every level; a level the
    library does not know is rejected."""
    from greenpeas_amd.batch import PreparedCircuit

    for name in ("surface3", "bb6"):
        text = read_circuit(name)
        for level in (0, 1, 2):
            p = PreparedCircuit(text, level)
            assert p.device_bytes() == estimate_device_bytes(text, level)
            p.close()
            p.close()  # a second close is a no-op
    with pytest.raises(ValueError):
        PreparedCircuit(read_circuit("surface3"), 3)
    with pytest.raises(ValueError):
        PreparedCircuit("FOO 0\n", 2)


def test_product_overlapping_channels(oracle):
    text = ("R 0 1 2\nTICK\nTICK\nTICK\nDEPOLARIZE1(0.01) 0 1 2\nDEPOLARIZE2(0.02) 0 1\nCX 0 2\n"
            "TICK\nH 1\nTICK\nM(0.001) 0 1 2\nDETECTOR rec[-1]\n")
    for level in (0, 1, 2):
        adj, probs, _ = build_stepg_host(text, level)
        s = oracle.build_stepg(text, level)
        assert np.array_equal(adj, s.adjacency) and np.array_equal(probs, s.probs)


def test_tuning_knobs_roundtrip():
    """Every knob set and read back through the C-ABI (no device work); bad
    names and out-of-range values are rejected with ValueError."""
    from greenpeas_amd import autotune

    old_fill, old_rpt = autotune.get_tuning("bucket_fill"), autotune.get_tuning("sweep_rpt")
    try:
        autotune.set_tuning("bucket_fill", 48)
        assert autotune.get_tuning("bucket_fill") == 48
        for r in (1, 2, 4, 8):
            autotune.set_tuning("sweep_rpt", r)
            assert autotune.get_tuning("sweep_rpt") == r
        autotune.set_tuning("stash", 0)
        assert autotune.get_tuning("stash") == 0
        autotune.set_tuning("stash", 1)
        assert autotune.get_tuning("stash") == 1
        autotune.set_tuning("mask_table", 0)
        assert autotune.get_tuning("mask_table") == 0
        autotune.set_tuning("mask_table", 1)
        assert autotune.get_tuning("mask_table") == 1
        autotune.set_tuning("prob_table", 0)
        assert autotune.get_tuning("prob_table") == 0
        autotune.set_tuning("prob_table", 1)
        assert autotune.get_tuning("prob_table") == 1
        autotune.set_tuning("packed_sort", 0)
        assert autotune.get_tuning("packed_sort") == 0
        autotune.set_tuning("packed_sort", 1)
        assert autotune.get_tuning("packed_sort") == 1
        autotune.set_tuning("sweep_split", 0)
        assert autotune.get_tuning("sweep_split") == 0
        autotune.set_tuning("sweep_split", 2)
        assert autotune.get_tuning("sweep_split") == 2
        autotune.set_tuning("sweep_split", 1)
        assert autotune.get_tuning("sweep_split") == 1
        with pytest.raises(ValueError):
            autotune.set_tuning("sweep_split", 3)
        for m in (0, 2, 1):
            autotune.set_tuning("s_bands", m)
            assert autotune.get_tuning("s_bands") == m
        with pytest.raises(ValueError):
            autotune.set_tuning("s_bands", 3)
        with pytest.raises(ValueError):
            autotune.set_tuning("sweep_rpt", 3)
        with pytest.raises(ValueError):
            autotune.set_tuning("no_such_knob", 1)
    finally:
        autotune.set_tuning("bucket_fill", old_fill)
        autotune.set_tuning("sweep_rpt", old_rpt)


def _digest(text, level, parallel):
    import numpy as np

    from greenpeas_amd import _native

    d = np.zeros(2, np.uint64)
    _native.check(_native.load().gp_debug_image_digest(text.encode(), level, parallel, d.ctypes.data))
    return tuple(int(x) for x in d)


@pytest.mark.parametrize("name", ["surface3", "bb6", "bb12", "iceberg_surface_6_static_circl"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_pooled_host_build_equals_single_thread(name, level):
    """The parse, STEPG and device image built on the worker pool (layer
    ranges, pair blocks and rank ranges in parallel, joined in order) are
    byte for byte the single-thread build."""
    text = read_circuit(name)
    assert _digest(text, level, 1) == _digest(text, level, 0)


def test_pooled_host_build_generated_circuits():
    """Larger generated circuits: noisy surface d=11 x 40 rounds with a
    rate per instruction, a reset-branch variant, BB [[144,12,12]] x 18."""
    import numpy as np

    from greenpeas_amd.batch import variant_circuits
    from greenpeas_amd.codes import BBCode, Noise, SurfaceCode

    texts = [SurfaceCode(11).memory_text(40, noise=Noise(rng=np.random.default_rng(8))),
             variant_circuits(1, seed=3, distances=(9,))[0], BBCode(12).memory_text(18, 1e-3)]
    # noise only at the start: the active nodes spread from two error nodes
    # through every layer range of the pooled pass (20 qubits x 600 CX layers)
    rng = np.random.default_rng(5)
    lines = ["R " + " ".join(map(str, range(20))), "TICK", "X_ERROR(0.01) 0", "DEPOLARIZE1(0.001) 3"]
    for layer in range(600):
        q = rng.permutation(20)
        lines += ["TICK", "CX " + " ".join(map(str, q[:8]))]
        if layer % 50 == 0:
            lines.append("H " + " ".join(map(str, q[8:11])))
    lines += ["TICK", "M " + " ".join(map(str, range(20)))]
    lines += [f"DETECTOR rec[-{20 - q}]" for q in range(20)] + ["OBSERVABLE_INCLUDE(0) rec[-1] rec[-2]"]
    texts.append("\n".join(lines) + "\n")
    for t in texts:
        assert _digest(t, 2, 1) == _digest(t, 2, 0)


def test_pooled_build_errors_match():
    """A flow-rule violation deep in a long circuit raises the same error
    from the pooled build as from one pass, and the next build (reusing the
    per-thread scratch) is still exact."""
    text = ("R 0 1 2\n" + "TICK\nDEPOLARIZE1(0.001) 0 1 2\nCX 0 1\n" * 3000
            + "TICK\nDEPOLARIZE2(0.01) 0 1\nDEPOLARIZE2(0.01) 1 2\n"
            + "TICK\nDEPOLARIZE1(0.001) 0 1 2\n" * 3000 + "TICK\nM 0 1 2\nDETECTOR rec[-1]\n")
    msgs = []
    for par in (0, 1):
        with pytest.raises(RuntimeError) as e:
            _digest(text, 2, par)
        msgs.append(str(e.value))
    assert msgs[0] == msgs[1]
    good = read_circuit("bb6")
    assert _digest(good, 2, 1) == _digest(good, 2, 0)
