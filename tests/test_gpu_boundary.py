"""The reference-shaped boundary on the GPU:
* the pybind11 module greenpeas_amd._gppy.error_analysis against the
  reference's golden DEMs (python/tests/test_error_analysis.py:27-59:
  every level, buffer reuse from a larger max circuit, both overloads);
* the C++ facade (include/greenpeas_amd.hpp) through a C++ program;
* the STEPG entry fed a reference-layout STEPG (the oracle's restatement of
  STEPG.hpp / Mixer.hpp / Circuit.hpp) compiling to the circuit's own DEM;
* a caller's stream."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, circuit_path, golden_dem_path, read_circuit
from greenpeas_amd import DetectorErrorModel, _native, get_driver

pytestmark = pytest.mark.gpu

REL = 1e-12
FACADE = os.path.join(ROOT, "tests", "cpp", "facade_check")


def gold(name, level, dem):
    return DetectorErrorModel.from_file(golden_dem_path(name, level), dem.num_detectors, dem.num_observables)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_pybind_module_matches_golden(level):
    from greenpeas_amd._gppy import error_analysis as ea

    lv = ea.CorrelationLevel(level)
    drv = ea.get_driver(read_circuit("surface5"), correlation_level=lv)
    big = drv.compile_detector_error_model()  # the max circuit, no argument
    assert big.approx_equals(gold("surface5", level, big), REL)
    small = drv.compile_detector_error_model(read_circuit("surface3"))  # buffer reuse
    assert small.approx_equals(gold("surface3", level, small), REL)
    # a path object works as the circuit too (str() of it is not stim text)
    import pathlib

    again = drv.compile_detector_error_model(pathlib.Path(circuit_path("surface3")))
    assert again == small
    with pytest.raises(RuntimeError, match="capacity"):
        drv.compile_detector_error_model(read_circuit("bb6"))


def test_pybind_default_level_is_l2_and_matches_ctypes_driver():
    from greenpeas_amd._gppy import error_analysis as ea

    text = read_circuit("bb12")
    a = ea.get_driver(text).compile_detector_error_model()
    b = get_driver(text).compile_detector_error_model()
    assert a == b


@pytest.mark.skipif(not os.path.exists(FACADE), reason="facade check program not built")
def test_cpp_facade_matches_golden():
    r = subprocess.run([FACADE, circuit_path("surface5"), "2", circuit_path("surface3")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    parts = r.stdout.split("=== ")[1:]
    assert len(parts) == 2
    for name, part in zip(("surface5", "surface3"), parts):
        body = part.split("\n", 1)[1]
        ref = get_driver(read_circuit(name)).compile_detector_error_model()
        dem = DetectorErrorModel.from_text(body, ref.num_detectors, ref.num_observables)
        assert dem == ref
        assert dem.approx_equals(gold(name, 2, dem), REL)
    r = subprocess.run([FACADE, circuit_path("surface3"), "2", circuit_path("bb6")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 3 and "EXCEPTION runtime_error: circuit exceeds the driver's capacity" in r.stdout


def load_stepg(drv, st, level):
    W = st.words
    M = st.measurements
    sens = np.ascontiguousarray(st.vals.T.reshape(-1)) if M and W else np.zeros(1, np.uint64)
    keep = (st.adjacency, st.probs, st.keys, sens)
    s = _native.Stepg(level, st.qubits, st.layers, M, st.detectors, st.observables,
                      st.adjacency.ctypes.data, st.probs.ctypes.data, st.keys.ctypes.data, sens.ctypes.data)
    _native.check(_native.load().gp_driver_load_stepg(drv._h, ctypes.byref(s)))
    del keep


@pytest.mark.parametrize("name,level", [("surface3", 0), ("surface5", 1), ("bb6", 2), ("bb12", 2)])
def test_stepg_entry_compiles_the_circuits_dem(oracle, name, level):
    """The reference-layout STEPG of a circuit, handed in through
    gp_driver_load_stepg, compiles to the DEM of the circuit text: same
    hashes and order, targets, and probabilities bit for bit."""
    text = read_circuit(name)
    drv = get_driver(text, correlation_level=level)
    ref = drv.compile_detector_error_model()
    ref_h = drv.last_hashes()
    st = oracle.build_stepg(text, level)
    load_stepg(drv, st, level)
    dem = drv.compile_detector_error_model()
    assert np.array_equal(drv.last_hashes(), ref_h)
    assert dem == ref
    assert drv.params["num_nodes"] == st.nodes


def test_stepg_entry_rejects_malformed_graphs(oracle):
    text = read_circuit("surface3")
    drv = get_driver(text)
    st = oracle.build_stepg(text, 2)
    with pytest.raises(ValueError, match="correlation level"):
        load_stepg(drv, st, 1)
    bad = oracle.build_stepg(text, 2)
    bad.adjacency[5] = np.uint64(3) | (np.uint64(0xFFFFFFFF) << np.uint64(32))  # flows backwards
    with pytest.raises(ValueError, match="not one or two layers later"):
        load_stepg(drv, bad, 2)
    bad = oracle.build_stepg(text, 2)
    bad.probs[7] = 1.5
    with pytest.raises(ValueError, match="outside"):
        load_stepg(drv, bad, 2)
    # the driver still compiles its circuit afterwards
    assert drv.compile_detector_error_model(text).num_errors == 229


def test_driver_on_a_caller_stream():
    torch = pytest.importorskip("torch")
    text = read_circuit("bb6")
    ref = get_driver(text).compile_detector_error_model()
    s = torch.cuda.Stream()
    lib = _native.load()
    h = ctypes.c_void_p()
    _native.check(lib.gp_get_driver_on_stream(text.encode(), 2, 0, ctypes.c_void_p(s.cuda_stream), ctypes.byref(h)))
    try:
        _native.check(lib.gp_compile_detector_error_model(h, None))
        from greenpeas_amd.error_analysis import _read_dem

        dem = _read_dem(lib, lib.gp_dem_shape, lib.gp_dem_copy, h)
        assert dem == ref
        _native.check(lib.gp_driver_set_stream(h, None))
        _native.check(lib.gp_compile_detector_error_model(h, None))
        assert _read_dem(lib, lib.gp_dem_shape, lib.gp_dem_copy, h) == ref
    finally:
        lib.gp_driver_destroy(h)
