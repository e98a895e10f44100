"""The reference-shaped boundary on the CPU: the pybind11 module's names,
enum values, defaults and exception classes (greenpeas_amd._gppy mirrors
_gppy.error_analysis, ErrorAnalysis.pybind.cu:47-79), the C++ facade's
exception classes (include/greenpeas_amd.hpp), and the STEPG entry's
argument checks -- all failing before any device work."""

from __future__ import annotations

import ctypes
import inspect
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, circuit_path
from greenpeas_amd import _native

FACADE = os.path.join(ROOT, "tests", "cpp", "facade_check")


def test_pybind_module_names_and_defaults():
    from greenpeas_amd._gppy import error_analysis as ea

    assert {"CorrelationLevel", "Driver", "get_driver"} <= set(dir(ea))
    assert [int(ea.CorrelationLevel.L0), int(ea.CorrelationLevel.L1), int(ea.CorrelationLevel.L2)] == [0, 1, 2]
    assert ea.L2 == ea.CorrelationLevel.L2  # export_values
    doc = ea.get_driver.__doc__
    assert "correlation_level" in doc and "CorrelationLevel.L2" in doc
    assert "compile_detector_error_model" in dir(ea.Driver)


def test_pybind_errors_are_the_reference_classes():
    from greenpeas_amd._gppy import error_analysis as ea

    with pytest.raises(ValueError, match="Unsupported operation: S"):
        ea.get_driver("R 0\nTICK\nS 0\n")
    with pytest.raises(ValueError, match="bad target"):
        ea.get_driver("R 0\nTICK\nM x\n")
    with pytest.raises(TypeError):
        ea.get_driver("R 0\n", correlation_level=7)


@pytest.mark.skipif(not os.path.exists(FACADE), reason="facade check program not built")
def test_facade_throws_invalid_argument_before_device_work(tmp_path):
    bad = tmp_path / "bad.stim"
    bad.write_text("R 0\nTICK\nS 0\n")
    r = subprocess.run([FACADE, str(bad), "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3
    assert "EXCEPTION invalid_argument: Unsupported operation: S" in r.stdout


def test_stepg_struct_matches_the_header():
    """gp_stepg's field order and sizes as the ctypes mirror declares them."""
    with open(os.path.join(ROOT, "include", "greenpeas_amd.h")) as f:
        text = f.read()
    body = text[text.index("typedef struct gp_stepg {"):text.index("} gp_stepg;")]
    for name, _ in _native.Stepg._fields_:
        assert name in body, name
    assert ctypes.sizeof(_native.Stepg) == 4 * 6 + 8 * 4  # 6 ints (24 bytes, 8-aligned), 4 pointers


def test_stepg_entry_rejects_bad_graphs_before_device_work():
    """Argument checks of gp_driver_load_stepg need no driver state beyond
    its level; a NULL driver is refused first."""
    lib = _native.load()
    st = _native.Stepg()
    assert lib.gp_driver_load_stepg(None, ctypes.byref(st)) == _native.GP_INVALID_ARGUMENT
    assert "NULL" in lib.gp_last_error().decode()
