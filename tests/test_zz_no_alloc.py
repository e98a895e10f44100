"""Allocation-free backward of the node rates (gp_node_rates_backward_noalloc):
through the C-ABI, on two calls in a row -- BB [[144,12,12]] 12 rounds, then
the smaller surface5 -- its gradient equals gp_node_rates_backward's with a
caller buffer bit for bit (the same kernels; the second call reuses the
fixed array the first one filled).  A component count above the fixed
array's size is rejected before any GPU work."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import read_circuit

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _backward(lib, dd, theta, grad_q, noalloc):
    from greenpeas_amd import _native

    grad = torch.full_like(theta, float("nan"))
    stream = torch.cuda.current_stream(theta.device).cuda_stream
    common = (theta.data_ptr(), dd.comp_loc32.data_ptr(), dd.comp_kind8.data_ptr(), dd.node_start.data_ptr(),
              dd.node_comp.data_ptr(), dd.num_nodes)
    tail = (dd.loc_start.data_ptr(), dd.loc_comp.data_ptr(), dd.num_locations, grad.data_ptr(), stream)
    if noalloc:
        rc = lib.gp_node_rates_backward_noalloc(*common, dd.num_components, grad_q.data_ptr(), *tail)
    else:
        buf = torch.empty(max(dd.num_components, 1), dtype=torch.float64, device=theta.device)
        rc = lib.gp_node_rates_backward(*common, grad_q.data_ptr(), buf.data_ptr(), *tail)
    _native.check(rc)
    torch.cuda.synchronize()
    return grad


def test_noalloc_backward_equals_caller_buffer_on_two_calls():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from greenpeas_amd import _native
    from greenpeas_amd.autograd import DifferentiableDem

    lib = _native.load()
    plans = [DifferentiableDem(read_circuit("bb12")), DifferentiableDem(read_circuit("surface5"))]
    assert plans[0].num_components > plans[1].num_components  # the second call is smaller
    assert plans[0].num_components <= _native.NOALLOC_MAX_COMPONENTS
    for k, dd in enumerate(plans):
        rng = np.random.default_rng(40 + k)
        theta = torch.tensor(np.exp(rng.uniform(np.log(1e-4), np.log(5e-3), dd.num_locations)), device="cuda")
        grad_q = torch.tensor(rng.standard_normal(dd.num_nodes), device="cuda")
        a = _backward(lib, dd, theta, grad_q, noalloc=True)
        b = _backward(lib, dd, theta, grad_q, noalloc=False)
        assert torch.isfinite(a).all()
        assert torch.equal(a, b), f"call {k}: allocation-free gradient differs"
    # beyond the fixed array: a clear error, nothing launched
    dd = plans[1]
    rc = lib.gp_node_rates_backward_noalloc(
        theta.data_ptr(), dd.comp_loc32.data_ptr(), dd.comp_kind8.data_ptr(), dd.node_start.data_ptr(),
        dd.node_comp.data_ptr(), dd.num_nodes, _native.NOALLOC_MAX_COMPONENTS + 1, grad_q.data_ptr(),
        dd.loc_start.data_ptr(), dd.loc_comp.data_ptr(), dd.num_locations, a.data_ptr(), None)
    assert rc == _native.GP_INVALID_ARGUMENT
    assert "exceeds the allocation-free maximum" in lib.gp_last_error().decode()
