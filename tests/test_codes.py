"""Circuit generators against the reference's generator outputs
(data/circuits/*.stim, pinned by tests/QEC/Codes/CodeTest.cpp:30-52 and
python/tests/test_codes.py:62-73) and code matrices (data/codes/surface)."""

from __future__ import annotations

import os

import pytest

from conftest import GOLDEN, read_circuit
from greenpeas_amd import codes


@pytest.mark.parametrize("d", range(3, 15))
def test_rotated_surface_matrices_match_reference(d):
    base = os.path.join(GOLDEN, "codes", "surface", f"d{d}")
    sc = codes.rotated_surface_code(d)
    nx, n, hx = codes.read_mtx(os.path.join(base, "Hx.mtx"))
    nz, _, hz = codes.read_mtx(os.path.join(base, "Hz.mtx"))
    assert (n, nx, nz) == (sc.num_data, sc.num_x, sc.num_z)
    assert hx == sc.hx and hz == sc.hz
    assert [c for _, c, _ in codes.read_mtx(os.path.join(base, "Lx.mtx"))[2]] == sc.lx[0]
    assert [c for _, c, _ in codes.read_mtx(os.path.join(base, "Lz.mtx"))[2]] == sc.lz[0]


@pytest.mark.parametrize("name,d", [("surface3", 3), ("surface5", 5), ("surface7", 7), ("surface9", 9)])
def test_surface_memory_matches_reference_circuit(name, d):
    assert codes.SurfaceCode(d).get_memory(d, 0.001) == read_circuit(name)


@pytest.mark.parametrize("name,d", [("bb6", 6), ("bb10", 10), ("bb12", 12)])
def test_bb_memory_matches_reference_circuit(name, d):
    assert codes.BBCode(d).get_memory(d, 0.001) == read_circuit(name)


def test_fault_location_counts():
    # counts quoted with the reference's published compile times
    assert codes.count_locations(read_circuit("surface3")) == 539
    assert codes.count_locations(read_circuit("bb12")) == 43632


@pytest.mark.parametrize("d", [15, 17, 21])
def test_large_surface_codes_are_valid_css(d):
    """Beyond the reference's data (d <= 14): every X check commutes with
    every Z check, logicals commute with the other type's checks and
    anticommute with each other."""
    sc = codes.rotated_surface_code(d)
    xs = [set() for _ in range(sc.num_x)]
    zs = [set() for _ in range(sc.num_z)]
    for r, c, _ in sc.hx:
        xs[r].add(c)
    for r, c, _ in sc.hz:
        zs[r].add(c)
    assert sc.num_x == sc.num_z == (d * d - 1) // 2
    assert all(len(x & z) % 2 == 0 for x in xs for z in zs)
    lx, lz = set(sc.lx[0]), set(sc.lz[0])
    assert all(len(lz & x) % 2 == 0 for x in xs)
    assert all(len(lx & z) % 2 == 0 for z in zs)
    assert len(lx & lz) % 2 == 1


def test_memory_with_oracle(oracle):
    """A generated d5 memory with seeded per-instruction rates and two
    reset-conditioned branches compiles identically with both oracles."""
    import numpy as np

    text = codes.SurfaceCode(5).memory_text(7, noise=codes.Noise(rng=np.random.default_rng(3)),
                                            branches={(2, 1), (4, 7)})
    a = oracle.compile_stepg(text, 2)
    b = oracle.compile_frames(text)
    from greenpeas_amd.dem import DetectorErrorModel as DEM

    da = DEM(a.probs, a.offsets, a.targets, a.num_detectors, a.num_observables)
    db = DEM(b.probs, b.offsets, b.targets, b.num_detectors, b.num_observables)
    assert da.approx_equals(db, 1e-12)
