"""GPU parity: the HIP pipeline, called through the C-ABI, against the oracle
and the reference's golden DEMs.

Bars: detector/observable sets and order bit-exact; probabilities within
1e-12 relative error (REL below); the sensitivity stage bit-exact.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_dem_path, read_circuit
from greenpeas_amd import CorrelationLevel, DetectorErrorModel, get_driver

pytestmark = pytest.mark.gpu

REL = 1e-12


def as_dem(od) -> DetectorErrorModel:
    return DetectorErrorModel(od.probs, od.offsets, od.targets, od.num_detectors, od.num_observables)


@pytest.mark.parametrize("name", ["surface3", "surface5"])
@pytest.mark.parametrize("level", [CorrelationLevel.L0, CorrelationLevel.L1, CorrelationLevel.L2])
def test_compile_matches_golden_dem(name, level):
    """python/tests/test_error_analysis.py:27-45, with the 1e-12 bar."""
    text = read_circuit(name)
    driver = get_driver(text, correlation_level=level)
    dem = driver.compile_detector_error_model(text)
    gold = DetectorErrorModel.from_file(golden_dem_path(name, int(level)), dem.num_detectors,
                                        dem.num_observables)
    assert dem.num_errors == gold.num_errors
    # same order (ascending FNV-1a-64 of the 32 padded slots) and targets
    assert np.array_equal(dem.offsets, gold.offsets)
    assert np.array_equal(dem.targets, gold.targets)
    assert np.max(np.abs(dem.probs - gold.probs) / gold.probs) <= REL


@pytest.mark.parametrize("name", ["surface3", "surface5", "bb6", "bb12", "iceberg_surface_4_static_circl",
                                  "iceberg_surface_4_static_pheno", "iceberg_surface_6_static_circl",
                                  "iceberg_surface_6_static_pheno"])
def test_compile_matches_oracles(oracle, name):
    text = read_circuit(name)
    drv = get_driver(text)
    dem = drv.compile_detector_error_model(text)
    ref = oracle.compile_stepg(text, 2)
    # identical order and hashes as the STEPG restatement
    assert dem.hashes is None  # the DEM alone, as the reference's
    assert np.array_equal(drv.last_hashes(), ref.hashes)
    assert np.array_equal(dem.offsets, ref.offsets)
    assert np.array_equal(dem.targets, ref.targets)
    assert np.max(np.abs(dem.probs - ref.probs) / ref.probs) <= REL
    # and the same model as the independent forward Pauli-frame compiler
    assert dem.approx_equals(as_dem(oracle.compile_frames(text)), REL)


@pytest.mark.parametrize("name", ["surface3", "bb6", "bb12"])
@pytest.mark.parametrize("level", [0, 2])
def test_sensitivity_sweep_bit_exact(oracle, name, level):
    text = read_circuit(name)
    driver = get_driver(text, correlation_level=level)
    driver.compile_detector_error_model()
    idx, sens = driver.debug_sensitivities()
    s = oracle.build_stepg(text, level)
    S = s.sweep()
    assert np.array_equal(idx, np.nonzero(s.probs > 0)[0])
    assert np.array_equal(sens, S[idx])


@pytest.mark.parametrize("level", [CorrelationLevel.L0, CorrelationLevel.L1, CorrelationLevel.L2])
def test_driver_reuse(level):
    """A driver sized for d5 compiles d3 (test_error_analysis.py:48-59)."""
    driver = get_driver(read_circuit("surface5"), correlation_level=level)
    dem = driver.compile_detector_error_model(read_circuit("surface3"))
    gold = DetectorErrorModel.from_file(golden_dem_path("surface3", int(level)), dem.num_detectors,
                                        dem.num_observables)
    assert np.array_equal(dem.targets, gold.targets)
    assert np.max(np.abs(dem.probs - gold.probs) / gold.probs) <= REL
    # back up to the max circuit: no stale state from the smaller compile
    big = driver.compile_detector_error_model(read_circuit("surface5"))
    gold5 = DetectorErrorModel.from_file(golden_dem_path("surface5", int(level)), big.num_detectors,
                                         big.num_observables)
    assert np.array_equal(big.targets, gold5.targets)


def test_capacity_overflow_raises():
    driver = get_driver(read_circuit("surface3"))
    with pytest.raises(RuntimeError, match="capacity"):
        driver.compile_detector_error_model(read_circuit("bb6"))
    # the driver still works afterwards
    dem = driver.compile_detector_error_model(read_circuit("surface3"))
    assert dem.num_errors == 229


def test_recompile_without_argument_is_identical():
    text = read_circuit("bb6")
    driver = get_driver(text)
    a = driver.compile_detector_error_model()
    b = driver.compile_detector_error_model()
    c = driver.compile_detector_error_model(text)
    assert a == b == c


def test_dem_text_output():
    text = read_circuit("surface3")
    driver = get_driver(text)
    dem = driver.compile_detector_error_model(text)
    again = DetectorErrorModel.from_text(driver.dem_text(), dem.num_detectors, dem.num_observables)
    assert again == dem


def test_empty_and_noiseless(oracle):
    driver = get_driver(read_circuit("surface3"))
    assert driver.compile_detector_error_model("R 0\nTICK\nM 0\nDETECTOR rec[-1]\n").num_errors == 0
    assert driver.compile_detector_error_model("").num_errors == 0


def test_long_run_takes_fallback_path(oracle):
    """An idle qubit over 600 layers: 1200 nodes share one class, far beyond a
    wave's 256-item sort capacity, so the workgroup fallback merges them."""
    text = "R 0 1\n" + "TICK\nDEPOLARIZE1(0.001) 0 1\n" * 600 + "TICK\nM 0 1\nDETECTOR rec[-2]\nDETECTOR rec[-1]\n"
    dem = get_driver(text).compile_detector_error_model(text)
    ref = as_dem(oracle.compile_stepg(text, 2))
    assert dem.approx_equals(ref, REL)
    assert dem.approx_equals(as_dem(oracle.compile_frames(text)), REL)


def test_long_class_beyond_32_targets(oracle):
    text = "R 0 1\nTICK\nX_ERROR(0.01) 0 1\nM 0 1\n" + "DETECTOR rec[-1]\n" * 40 + "DETECTOR rec[-2]\n" * 3
    drv = get_driver(text)
    dem = drv.compile_detector_error_model(text)
    ref = oracle.compile_stepg(text, 2)
    assert np.array_equal(drv.last_hashes(), ref.hashes)
    assert dem.approx_equals(as_dem(ref), REL)
    assert max(np.diff(dem.offsets)) == 40


def test_errors_are_reference_exception_classes():
    driver = get_driver(read_circuit("surface3"))
    with pytest.raises(ValueError, match="Unsupported operation"):
        driver.compile_detector_error_model("S 0\n")
    with pytest.raises(RuntimeError):
        driver.compile_detector_error_model("R 0\nTICK\nDEPOLARIZE1(0.9) 0\nTICK\nM 0\n")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_circuits(oracle, seed):
    """Random Clifford circuits (R/H/CX/M/noise) against both oracles."""
    rng = np.random.default_rng(seed)
    nq = 12
    lines = [f"R {' '.join(map(str, range(nq)))}"]
    nmeas = 0
    for layer in range(40):
        lines.append("TICK")
        qs = rng.permutation(nq)
        lines.append(f"DEPOLARIZE1({rng.uniform(1e-4, 1e-2):.6g}) {' '.join(map(str, range(nq)))}")
        pairs = qs[: 2 * (nq // 4)].reshape(-1, 2)
        if layer >= 2:
            lines.append(f"DEPOLARIZE2({rng.uniform(1e-4, 1e-2):.6g}) " + " ".join(f"{a} {b}" for a, b in pairs))
        lines.append("CX " + " ".join(f"{a} {b}" for a, b in pairs))
        hq = [int(q) for q in qs[2 * (nq // 4):] if rng.random() < 0.5]
        if hq:
            lines.append("H " + " ".join(map(str, hq)))
        if layer % 8 == 7:
            lines.append("TICK")
            m = [int(q) for q in rng.choice(nq, 4, replace=False)]
            lines.append(f"X_ERROR(0.001) {' '.join(map(str, m))}")
            lines.append(f"M(0.002) {' '.join(map(str, m))}")
            nmeas += 4
            for k in range(1, 5):
                lines.append(f"DETECTOR rec[-{k}]")
            if nmeas > 4:
                lines.append("DETECTOR rec[-1] rec[-5]")
            lines.append(f"R {' '.join(map(str, m))}")
    lines.append("TICK")
    lines.append(f"M {' '.join(map(str, range(nq)))}")
    lines.append("OBSERVABLE_INCLUDE(0) rec[-1] rec[-2]")
    lines.append("DETECTOR rec[-3] rec[-4]")
    text = "\n".join(lines) + "\n"
    drv = get_driver(text)
    dem = drv.compile_detector_error_model(text)
    ref = oracle.compile_stepg(text, 2)
    assert np.array_equal(drv.last_hashes(), ref.hashes)
    assert dem.approx_equals(as_dem(ref), REL)


@pytest.mark.parametrize("rpt", [2, 4, 8])
def test_sweep_records_per_thread_knob(oracle, rpt):
    """The sweep's launch shape (records per thread) changes nothing in the
    result: bit-identical DEMs and sensitivities to the default shape."""
    from greenpeas_amd import autotune

    texts = [read_circuit(n) for n in ("surface5", "bb6", "bb12")]
    base = [get_driver(t) for t in texts]
    ref = [(d.compile_detector_error_model(), d.debug_sensitivities()) for d in base]
    old = autotune.get_tuning("sweep_rpt")
    try:
        autotune.set_tuning("sweep_rpt", rpt)
        for t, (dem0, (idx0, s0)) in zip(texts, ref):
            d = get_driver(t)
            assert d.compile_detector_error_model() == dem0
            idx, s = d.debug_sensitivities()
            assert np.array_equal(idx, idx0) and np.array_equal(s, s0)
    finally:
        autotune.set_tuning("sweep_rpt", old)


def test_emit_without_target_rows(oracle):
    """With the stored target rows switched off, emit gathers every class's
    targets from the sensitivity matrix (the path of circuits above 65536
    targets): same DEMs, bit for bit."""
    from greenpeas_amd import autotune

    texts = [read_circuit(n) for n in ("surface5", "bb12")]
    ref = [get_driver(t).compile_detector_error_model() for t in texts]
    try:
        autotune.set_tuning("stash", 0)
        for t, dem0 in zip(texts, ref):
            assert get_driver(t).compile_detector_error_model() == dem0
    finally:
        autotune.set_tuning("stash", 1)


def test_exact_bucket_sort_equals_packed_key_sort(oracle):
    """The bucket stage's exact (hash, rank) sort -- where the packed 64-bit
    key cannot order a bucket: two distinct hashes sharing the 32 bits below
    the bucket digit, or a circuit with 2^24 - 1 or more error nodes -- gives
    the same DEMs bit for bit as the packed key, alone and in a batch (one
    circuit has 1,200-member classes, whose buckets take the oversized
    path)."""
    from greenpeas_amd import autotune
    from greenpeas_amd.batch import BatchCompiler

    names = ("surface3", "surface5", "bb12", "iceberg_surface_6_static_circl")
    texts = [read_circuit(n) for n in names]
    texts.append("R 0 1\n" + "TICK\nDEPOLARIZE1(0.001) 0 1\n" * 600
                 + "TICK\nM 0 1\nDETECTOR rec[-2]\nDETECTOR rec[-1]\n")
    packed = [get_driver(t).compile_detector_error_model() for t in texts]
    bc = BatchCompiler()
    bc.set_circuits(texts)
    packed_batch = bc.compile()
    try:
        autotune.set_tuning("packed_sort", 0)
        for t, dem0 in zip(texts, packed):
            assert get_driver(t).compile_detector_error_model() == dem0
        bc2 = BatchCompiler()
        bc2.set_circuits(texts)
        assert bc2.compile() == packed_batch
    finally:
        autotune.set_tuning("packed_sort", 1)
    ref = oracle.compile_stepg(texts[2], 2)
    assert np.array_equal(packed[2].targets, ref.targets)


@pytest.mark.parametrize("d,rounds", [(5, 100), (7, 100), (5, 260)])
def test_class_stage_masks_wider_than_64_columns(oracle, d, rounds):
    """Circuits with more than 64 detector columns (2,400 to 6,240
    detectors: 3, 5 and 7 mask words per node window) walk their column
    masks two words at a time, reloading the window: identical hashes,
    order, targets and probabilities to the STEPG restatement."""
    from greenpeas_amd.codes import SurfaceCode

    text = SurfaceCode(d).memory_text(rounds, 1e-3)
    drv = get_driver(text)
    dem = drv.compile_detector_error_model()
    assert (dem.num_detectors + dem.num_observables + 31) // 32 > 64
    ref = oracle.compile_stepg(text, 2)
    assert np.array_equal(drv.last_hashes(), ref.hashes)
    assert np.array_equal(dem.offsets, ref.offsets)
    assert np.array_equal(dem.targets, ref.targets)
    assert np.max(np.abs(dem.probs - ref.probs) / ref.probs) <= REL


@pytest.mark.parametrize("stash", [1, 0])
def test_class_stage_without_mask_tables(oracle, stash):
    """With the class stage's LDS column masks switched off, every node tests
    every column's time band (the path of circuits whose masks do not fit in
    LDS): identical hashes, order, targets and probabilities to the STEPG
    restatement, with and without the stored target rows."""
    from greenpeas_amd import autotune

    try:
        autotune.set_tuning("mask_table", 0)
        autotune.set_tuning("stash", stash)
        for name, level in (("surface3", 0), ("surface5", 2), ("bb6", 1), ("bb12", 2)):
            text = read_circuit(name)
            drv = get_driver(text, correlation_level=level)
            dem = drv.compile_detector_error_model()
            ref = oracle.compile_stepg(text, level)
            assert np.array_equal(drv.last_hashes(), ref.hashes)
            assert np.array_equal(dem.offsets, ref.offsets)
            assert np.array_equal(dem.targets, ref.targets)
            assert np.max(np.abs(dem.probs - ref.probs) / ref.probs) <= REL
    finally:
        autotune.set_tuning("mask_table", 1)
        autotune.set_tuning("stash", 1)


def test_probabilities_sent_as_doubles(oracle):
    """With the probability table switched off every error node's double is
    sent (the path of circuits with more than 65536 distinct rates): the same
    DEMs as the STEPG restatement, bit for bit in order and targets."""
    from greenpeas_amd import autotune

    try:
        autotune.set_tuning("prob_table", 0)
        for name, level in (("surface3", 2), ("bb6", 2)):
            text = read_circuit(name)
            drv = get_driver(text, correlation_level=level)
            dem = drv.compile_detector_error_model()
            ref = oracle.compile_stepg(text, level)
            assert np.array_equal(drv.last_hashes(), ref.hashes)
            assert np.array_equal(dem.offsets, ref.offsets)
            assert np.array_equal(dem.targets, ref.targets)
            assert np.max(np.abs(dem.probs - ref.probs) / ref.probs) <= REL
    finally:
        autotune.set_tuning("prob_table", 1)


def test_dem_views_outlive_later_compiles_and_the_driver():
    """A DEM's arrays are views of the driver's pinned result memory: they
    keep their values through later compiles of other circuits and after the
    driver is closed (gp_dem_acquire / gp_dem_release)."""
    small, big = read_circuit("surface3"), read_circuit("surface5")
    d = get_driver(big)
    dem_small = d.compile_detector_error_model(small)
    copy_small = d.last_dem_copy()
    hashes_small = d.last_hashes()
    dem_big = d.compile_detector_error_model(big)
    dem_small2 = d.compile_detector_error_model(small)
    d.close()
    assert dem_small == copy_small and dem_small2 == copy_small
    assert np.array_equal(dem_small.offsets, copy_small.offsets)
    assert np.array_equal(hashes_small, copy_small.hashes)
    assert dem_big == get_driver(big).compile_detector_error_model()
    del dem_small, dem_small2, dem_big  # releases the blocks after the driver


@pytest.mark.parametrize("name", ["surface5", "bb12"])
def test_repeated_compiles_with_narrowed_target_rows(oracle, name):
    """After a load's first compile the stored target rows shrink to the
    narrowest width that holds its largest class (surface5: 4 words;
    bb12's 18-target classes keep 16) and the class stage may switch to 16
    columns per round trip (bb12: ~15 columns per node); every compile
    gives the same DEM and order as the STEPG restatement."""
    text = read_circuit(name)
    ref = oracle.compile_stepg(text, 2)
    drv = get_driver(text)
    for _ in range(3):
        dem = drv.compile_detector_error_model()
        assert np.array_equal(drv.last_hashes(), ref.hashes)
        assert np.array_equal(dem.offsets, ref.offsets)
        assert np.array_equal(dem.targets, ref.targets)
        assert np.max(np.abs(dem.probs - ref.probs) / ref.probs) <= REL
    drv.bench_device(3)  # graph replays with the narrowed rows
    again = drv.compile_detector_error_model()
    assert np.array_equal(again.targets, ref.targets)


@pytest.fixture
def sweep_split():
    """Sets the sweep_split knob for a test; restores the default after."""
    from greenpeas_amd import autotune

    yield lambda mode: autotune.set_tuning("sweep_split", mode)
    autotune.set_tuning("sweep_split", 1)


def _compile_twice(text):
    """A driver's DEM and sensitivities after its second compile of one load
    (the first compile finds the column bands; the second plans the split
    sweep from them and runs it).  The compiles pass no circuit: a compile
    given circuit text loads it anew, and a new load starts unplanned."""
    drv = get_driver(text)
    drv.compile_detector_error_model()
    dem = drv.compile_detector_error_model()
    idx, sens = drv.debug_sensitivities()
    return dem, idx, sens, drv.last_hashes(), drv.debug_sweep_plan()


@pytest.mark.parametrize("name", ["bb12", "surface5", "surface7", "bb6", "bb10"])
def test_split_sweep_identical_to_whole_sweep(name, oracle, sweep_split):
    """Columns swept in time chunks (transfer maps of the middle chunks,
    lower chunks resumed from their entering state) give bit-identical
    sensitivities, hashes and DEMs to sweeping every column whole (forced:
    every column that reaches the first layers, as many chunks as fit)."""
    text = read_circuit(name)
    sweep_split(2)
    split = _compile_twice(text)
    sweep_split(0)
    whole = _compile_twice(text)
    assert split[4][0] >= 1 and split[4][2] >= 2  # the split path ran
    assert whole[4] == (0, 0, 0)
    assert np.array_equal(split[1], whole[1])
    assert np.array_equal(split[2], whole[2])
    assert np.array_equal(split[3], whole[3])
    assert split[0] == whole[0]
    ref = oracle.compile_stepg(text, 2)
    assert np.array_equal(split[3], ref.hashes)


@pytest.mark.parametrize("d,rounds", [(5, 60), (11, 200)])
def test_split_sweep_planned_for_a_long_memory(sweep_split, d, rounds):
    """As planned (the default): a long memory's observables column, far
    longer than any detector column, is swept in chunks cut where a round
    resets its ancillas (d=11 x 200 rounds: a ~1,650-layer column); so is
    the 12-round bb12's observables column (91 pairs against <= 32 for a
    detector column); results equal the whole sweep's."""
    from greenpeas_amd.codes import SurfaceCode

    text = SurfaceCode(d).memory_text(rounds, 1e-3)
    split = _compile_twice(text)
    assert split[4][0] == 1 and split[4][2] >= 2
    assert _compile_twice(read_circuit("bb12"))[4][0] == 1
    sweep_split(0)
    whole = _compile_twice(text)
    assert whole[4] == (0, 0, 0)
    assert np.array_equal(split[2], whole[2]) and np.array_equal(split[3], whole[3])
    assert split[0] == whole[0]


def test_split_sweep_default_plan_bb12(oracle):
    """As planned by default, bb12's observables column is swept in chunks
    whose lower chunks wait for the maps of the chunks above them in the same
    launch; the DEM, its order (the oracle's class hashes) and every
    sensitivity equal those of the whole sweep."""
    from greenpeas_amd import autotune

    text = read_circuit("bb12")
    split = _compile_twice(text)
    assert split[4][0] == 1 and split[4][2] >= 2
    autotune.set_tuning("sweep_split", 0)
    try:
        whole = _compile_twice(text)
    finally:
        autotune.set_tuning("sweep_split", 1)
    assert whole[4] == (0, 0, 0)
    assert np.array_equal(split[2], whole[2]) and np.array_equal(split[3], whole[3])
    assert split[0] == whole[0]
    ref = oracle.compile_stepg(text, 2)
    assert np.array_equal(split[3], ref.hashes)
    assert np.max(np.abs(split[0].probs - ref.probs) / ref.probs) <= REL


def test_split_sweep_repeated_compiles_and_reload(sweep_split):
    """A driver keeps its split plan over many compiles of a load (the
    producers' counts and row lists are cleared by each run for the next) and
    plans anew after a load of another circuit; every result equals the first
    (the load's first compile sweeps every column whole)."""
    a, b = read_circuit("bb12"), read_circuit("surface5")
    sweep_split(2)
    drv = get_driver(a)
    first = drv.compile_detector_error_model()
    for _ in range(4):
        assert drv.compile_detector_error_model() == first
    assert drv.debug_sweep_plan()[0] >= 1
    drv.bench_device(3)  # graph replays and direct launches of the split plan
    assert drv.compile_detector_error_model() == first
    small = drv.compile_detector_error_model(b)
    assert drv.debug_sweep_plan() == (0, 0, 0)  # a new load: not planned yet
    assert drv.compile_detector_error_model() == small
    assert drv.debug_sweep_plan()[0] >= 1
    assert drv.compile_detector_error_model() == small
    assert drv.compile_detector_error_model(a) == first
    assert drv.compile_detector_error_model() == first
    assert drv.compile_detector_error_model() == first


@pytest.mark.parametrize("name", ["iceberg_surface_4_adaptive_circl", "iceberg_surface_6_adaptive_pheno"])
def test_noiseless_adaptive_iceberg_circuits(name):
    """p = 0 circuits of the reference's concatenated code: an empty DEM with
    the circuit's detector and observable counts."""
    text = read_circuit(name)
    dem = get_driver(text).compile_detector_error_model(text)
    assert dem.num_errors == 0
    assert dem.num_detectors == text.count("DETECTOR") and dem.num_observables == 1
