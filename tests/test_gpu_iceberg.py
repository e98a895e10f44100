"""The adaptive Iceberg x surface-code workload on the GPU: the reference's
recompile-per-branch loop (one driver sized from the static circuit,
Cli.hpp:378-416) and the same branches compiled in one batch give the same
DEMs, which match the oracle."""

from __future__ import annotations

import pytest

from greenpeas_amd import DetectorErrorModel, get_driver
from greenpeas_amd.batch import BatchCompiler
from greenpeas_amd.iceberg import IcebergSurfaceCode, adaptive_branches

pytestmark = pytest.mark.gpu


def as_dem(od) -> DetectorErrorModel:
    return DetectorErrorModel(od.probs, od.offsets, od.targets, od.num_detectors, od.num_observables)


@pytest.mark.parametrize("d, count", [(4, 16), (6, 8)])
def test_adaptive_branches_loop_and_batch_agree(oracle, d, count):
    static, _ = IcebergSurfaceCode(d).memory(d, 1e-3)
    branches = adaptive_branches(d, count, p=1e-3, seed=100 * d)
    assert len(set(branches)) > 1
    drv = get_driver(static)
    assert drv.compile_detector_error_model().num_errors > 0
    loop = [drv.compile_detector_error_model(t) for t in branches]
    drv.close()
    bc = BatchCompiler()
    bc.set_circuits(branches)
    batched = bc.compile()
    bc.close()
    for i, (a, b) in enumerate(zip(loop, batched)):
        assert a == b, f"branch {i} differs between the driver loop and the batch"
    for i in (0, count - 1):
        assert loop[i].approx_equals(as_dem(oracle.compile_stepg(branches[i], 2)), 1e-12)
