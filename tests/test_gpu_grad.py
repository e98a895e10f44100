"""Differentiable DEM probabilities on the GPU (config: per-location FP64
noise vector theta, L = sum_e w_e logit(p_e)): forward against the oracle's
compile at the same rates (1e-12), gradient against the oracle's FP64 chain
rule and central differences (relative step 1e-6)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import read_circuit

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ag():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from greenpeas_amd import autograd

    return autograd


def test_xor_merge_kernel_matches_torch_fp64(ag):
    rng = np.random.default_rng(0)
    n, nseg = 5000, 700
    cuts = np.sort(rng.choice(np.arange(1, n), nseg - 1, replace=False))
    seg = np.concatenate([[0], cuts, [n]]).astype(np.int32)
    perm = rng.permutation(n).astype(np.int32)
    vals = rng.uniform(0, 0.6, n)
    v = torch.tensor(vals, dtype=torch.float64, device="cuda", requires_grad=True)
    out = ag.xor_merge(v, torch.tensor(perm, device="cuda"), torch.tensor(seg, device="cuda"))
    g = torch.tensor(rng.standard_normal(nseg), dtype=torch.float64, device="cuda")
    (out * g).sum().backward()
    # plain torch fp64 reference of the same op on the CPU
    vc = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    ref = torch.stack([(1 - torch.prod(1 - 2 * vc[torch.tensor(perm[seg[s]:seg[s + 1]], dtype=torch.long)])) / 2
                       for s in range(nseg)])
    (ref * g.cpu()).sum().backward()
    assert torch.allclose(out.detach().cpu(), ref.detach(), rtol=1e-12, atol=0)
    assert torch.allclose(v.grad.cpu(), vc.grad, rtol=1e-12, atol=1e-300)


def test_xor_merge_gradcheck(ag):
    rng = np.random.default_rng(1)
    seg = torch.tensor([0, 3, 3, 7, 12], dtype=torch.int32, device="cuda")
    idx = torch.tensor(rng.permutation(12).astype(np.int32), device="cuda")
    v = torch.tensor(rng.uniform(0.01, 0.4, 12), dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: ag.xor_merge(x, idx, seg), (v,), eps=1e-7, atol=1e-9)


def _theta(n, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(1e-4), np.log(5e-3), n))


@pytest.mark.parametrize("name", ["surface3", "bb6"])
def test_forward_matches_oracle_compile(ag, oracle, name):
    text = read_circuit(name)
    dd = ag.DifferentiableDem(text)
    theta = _theta(dd.num_locations, 11)
    p = dd.probabilities(torch.tensor(theta, device="cuda")).cpu().numpy()
    ref = oracle.compile_stepg(text, 2, theta)
    assert np.array_equal(dd.offsets, ref.offsets)
    assert np.array_equal(dd.targets, ref.targets)
    assert np.max(np.abs(p - ref.probs) / ref.probs) <= 1e-12


def test_bb12_gradient(ag, oracle):
    """BB [[144,12,12]], 12 rounds: theta ~ logU[1e-4, 5e-3], w ~ N(0,1)."""
    text = read_circuit("bb12")
    dd = ag.DifferentiableDem(text)
    assert dd.num_locations == 43632
    theta = _theta(dd.num_locations, 2024)
    w_np = np.random.default_rng(2025).standard_normal(dd.num_errors)
    th = torch.tensor(theta, device="cuda", requires_grad=True)
    w = torch.tensor(w_np, device="cuda")
    loss = ag.logit_loss(dd.probabilities(th), w)
    loss.backward()
    g = th.grad.cpu().numpy()
    L_ref, g_ref = oracle.logit_loss_grad(text, theta, w_np, 2)
    assert abs(loss.item() - L_ref) <= 1e-10 * abs(L_ref)
    rel = np.abs(g - g_ref) / np.maximum(np.abs(g_ref), 1e-300)
    assert np.max(rel) <= 1e-9
    # central differences on a seeded sample of coordinates
    rng = np.random.default_rng(5)
    pr = dd.probabilities(th.detach()).cpu().numpy()
    noise = 8 * np.finfo(float).eps * float(np.sum(np.abs(w_np * (np.log(pr) - np.log1p(-pr)))))
    with torch.no_grad():
        for k in rng.choice(dd.num_locations, 12, replace=False):
            h = 1e-6 * theta[k]
            tp = th.detach().clone()
            tm = th.detach().clone()
            tp[k] += h
            tm[k] -= h
            fd = (ag.logit_loss(dd.probabilities(tp), w) - ag.logit_loss(dd.probabilities(tm), w)).item() / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * abs(g[k]) + noise / h, (k, fd, g[k])


@pytest.mark.parametrize("name", ["surface5", "bb12"])
def test_fused_node_rates_match_tensor_formula(ag, name):
    """The fused node-probability kernels (split rates + XOR-merge in one
    forward, per-node products + per-location sums backward) against the
    same function written as plain float64 tensor operations."""
    dd = ag.DifferentiableDem(read_circuit(name))
    rng = np.random.default_rng(5)
    theta0 = np.exp(rng.uniform(np.log(1e-4), np.log(5e-2), dd.num_locations))
    w = torch.tensor(rng.standard_normal(dd.num_nodes), device="cuda")
    t1 = torch.tensor(theta0, device="cuda", requires_grad=True)
    q1 = dd.node_probabilities(t1)
    (q1 * w).sum().backward()
    t2 = torch.tensor(theta0, device="cuda", requires_grad=True)
    q2 = ag.xor_merge(dd.component_rates(t2), dd.node_comp, dd.node_start)
    (q2 * w).sum().backward()
    # the split rate 1/2 - x/2 cancels for small rates: a last-bit difference
    # in the root becomes up to ~1e-11 relative at theta = 1e-4, so that is
    # the bar (the kernels evaluate the same formula without contractions)
    assert torch.allclose(q1, q2, rtol=2e-11, atol=0)
    assert torch.allclose(t1.grad, t2.grad, rtol=1e-10, atol=1e-300)
    # deterministic: the same gradient bits on a second run
    t3 = torch.tensor(theta0, device="cuda", requires_grad=True)
    (dd.node_probabilities(t3) * w).sum().backward()
    assert torch.equal(t1.grad, t3.grad)


def test_xor_merge_with_repeated_indices(ag):
    """A value used at several positions (within and across segments) gets
    the sum of its positions' gradients, as torch's own fp64 graph of the
    same op does."""
    rng = np.random.default_rng(3)
    idx = np.array([4, 1, 4, 2, 0, 1, 3, 4, 2], np.int32)
    seg = np.array([0, 3, 3, 6, 9], np.int32)
    vals = rng.uniform(0.01, 0.45, 5)
    v = torch.tensor(vals, dtype=torch.float64, device="cuda", requires_grad=True)
    out = ag.xor_merge(v, torch.tensor(idx, device="cuda"), torch.tensor(seg, device="cuda"))
    g = torch.tensor(rng.standard_normal(4), dtype=torch.float64, device="cuda")
    (out * g).sum().backward()
    vc = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    ref = torch.stack([(1 - torch.prod(1 - 2 * vc[torch.tensor(idx[seg[s]:seg[s + 1]], dtype=torch.long)])) / 2
                       for s in range(4)])
    (ref * g.cpu()).sum().backward()
    assert torch.allclose(out.detach().cpu(), ref.detach(), rtol=1e-13, atol=1e-300)
    assert torch.allclose(v.grad.cpu(), vc.grad, rtol=1e-12, atol=1e-300)
