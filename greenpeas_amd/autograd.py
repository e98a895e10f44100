"""Differentiable DEM probabilities: gradients with respect to per-location
noise rates (no counterpart in the reference).

The error classes of a circuit depend only on its structure; the
probabilities follow from the per-location rates theta by

    r_c  = split(theta[loc(c)])      per noise component c: the rate itself
                                     (X_ERROR, M), 1/2 - 1/2 sqrt(1 - 4t/3)
                                     (DEPOLARIZE1, Mixer.hpp:164) or
                                     1/2 - 1/2 (1 - 16t/15)^(1/8)
                                     (DEPOLARIZE2, Mixer.hpp:184)
    q_n  = XOR-merge of r_c over the components placed on STEPG node n
    p_e  = XOR-merge of q_n over the nodes of error class e

Both merges are the same segmented op, ``XorMerge``, whose forward and
hand-written backward run as HIP kernels (``gp_xor_merge_forward`` /
``gp_xor_merge_backward``).  The splits are elementwise torch ops.

Import order: this module imports torch before loading the native library,
so both share torch's HIP runtime.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .dem import DetectorErrorModel
from .error_analysis import CorrelationLevel, circuit_text


def _stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _check_index_tensors(values: torch.Tensor, index: torch.Tensor, seg_start: torch.Tensor,
                         require_cuda: bool = True) -> None:
    """Host-side checks that need no device round trip: the kernels read
    values as float64 and index and seg_start as contiguous uint32 arrays,
    all on one device."""
    if not isinstance(values, torch.Tensor) or values.dtype != torch.float64:
        raise TypeError("values must be a float64 tensor")
    if require_cuda and not values.is_cuda:
        raise TypeError("values must be a CUDA tensor")
    if values.dim() != 1:
        raise ValueError("values must be one-dimensional")
    for name, t in (("index", index), ("seg_start", seg_start)):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.int32:
            raise TypeError(f"{name} must be an int32 tensor (read as uint32), got "
                            f"{getattr(t, 'dtype', type(t).__name__)}")
        if t.device != values.device:
            raise ValueError(f"{name} is on {t.device}, values on {values.device}")
        if t.dim() != 1 or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous one-dimensional tensor")
    if seg_start.numel() < 1:
        raise ValueError("seg_start needs at least one entry")


def _check_segments(values: torch.Tensor, index: torch.Tensor, seg_start: torch.Tensor) -> bool:
    """Checks the segment structure on the tensors' device (one round trip):
    seg_start[0] == 0, non-decreasing, seg_start[-1] <= index.numel(), every
    used index (positions below seg_start[-1]) within values.  Returns
    whether a used index repeats."""
    s = seg_start.to(torch.int64)
    used = index.to(torch.int64)
    pos = torch.arange(index.numel(), device=index.device)
    in_use = pos < s[-1]
    bad_idx = ((used < 0) | (used >= values.numel())) & in_use
    # positions past the segments get distinct keys above every index
    key = torch.sort(torch.where(in_use, used, values.numel() + pos)).values
    repeated = (key[1:] == key[:-1]).any() if key.numel() > 1 else torch.zeros((), dtype=torch.bool,
                                                                               device=key.device)
    decreasing = (s[1:] < s[:-1]).any() if s.numel() > 1 else torch.zeros((), dtype=torch.bool,
                                                                          device=s.device)
    flags = torch.stack([s[0] != 0, decreasing, s[-1] > index.numel(), bad_idx.any(),
                         repeated]).cpu().tolist()
    if flags[0]:
        raise ValueError("seg_start[0] must be 0")
    if flags[1]:
        raise ValueError("seg_start must be non-decreasing")
    if flags[2]:
        raise ValueError(f"seg_start[-1] exceeds index.numel() = {index.numel()}")
    if flags[3]:
        raise ValueError(f"index holds entries outside values (numel {values.numel()})")
    return bool(flags[4])


class XorMerge(torch.autograd.Function):
    """out[s] = XOR-merge of values[index[k]], k in [seg_start[s], seg_start[s+1]).

    The kernels need each value in at most one position of ``index``;
    ``xor_merge`` checks that (and the segment structure) and routes repeated
    indices through a gathered copy."""

    @staticmethod
    def forward(ctx, values: torch.Tensor, index: torch.Tensor, seg_start: torch.Tensor) -> torch.Tensor:
        _check_index_tensors(values, index, seg_start)
        values = values.contiguous()
        nseg = seg_start.numel() - 1
        out = torch.empty(nseg, dtype=torch.float64, device=values.device)
        lib = _native.load()
        _native.check(lib.gp_xor_merge_forward(values.data_ptr(), index.data_ptr(), seg_start.data_ptr(),
                                               nseg, out.data_ptr(), _stream_ptr(values.device)))
        ctx.save_for_backward(values, index, seg_start)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        values, index, seg_start = ctx.saved_tensors
        grad = torch.zeros_like(values)
        g = grad_out.contiguous().to(torch.float64)
        lib = _native.load()
        _native.check(lib.gp_xor_merge_backward(values.data_ptr(), index.data_ptr(), seg_start.data_ptr(),
                                                seg_start.numel() - 1, g.data_ptr(), grad.data_ptr(),
                                                _stream_ptr(values.device)))
        return grad, None, None


def xor_merge(values: torch.Tensor, index: torch.Tensor, seg_start: torch.Tensor,
              validate: bool = True) -> torch.Tensor:
    """Segmented XOR-merge with a hand-written backward (see XorMerge).

    index and seg_start are int32 tensors (their bits are read as uint32) on
    values' device.  With ``validate`` (default) the segment structure is
    checked on the device first (one synchronization) and repeated indices
    are handled by merging a gathered copy of the values, whose gradient is
    then summed back per value; ``validate=False`` is for index arrays known
    to be valid partial permutations (a DEM plan's own)."""
    _check_index_tensors(values, index, seg_start)
    if validate and _check_segments(values, index, seg_start):
        n = index.numel()
        gathered = values[index.to(torch.int64)]
        ident = torch.arange(n, dtype=torch.int32, device=values.device)
        return XorMerge.apply(gathered, ident, seg_start)
    return XorMerge.apply(values, index, seg_start)


class NodeRates(torch.autograd.Function):
    """q[n] = XOR-merge of the split rates of node n's components, from the
    location rates theta, in one kernel (gp_node_rates_forward); backward in
    two (per-node products, then one ordered sum per location)."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, plan: "DifferentiableDem") -> torch.Tensor:
        if theta.dtype != torch.float64 or not theta.is_cuda:
            raise TypeError("theta must be a float64 CUDA tensor")
        theta = theta.contiguous()
        q = torch.empty(plan.num_nodes, dtype=torch.float64, device=theta.device)
        lib = _native.load()
        _native.check(lib.gp_node_rates_forward(theta.data_ptr(), plan.comp_loc32.data_ptr(),
                                                plan.comp_kind8.data_ptr(), plan.node_start.data_ptr(),
                                                plan.node_comp.data_ptr(), plan.num_nodes, q.data_ptr(),
                                                _stream_ptr(theta.device)))
        ctx.save_for_backward(theta)
        ctx.plan = plan
        return q

    @staticmethod
    def backward(ctx, grad_q: torch.Tensor):
        (theta,) = ctx.saved_tensors
        plan = ctx.plan
        g = grad_q.contiguous().to(torch.float64)
        grad_comp = torch.empty(max(plan.num_components, 1), dtype=torch.float64, device=theta.device)
        grad_theta = torch.empty_like(theta)
        lib = _native.load()
        _native.check(lib.gp_node_rates_backward(theta.data_ptr(), plan.comp_loc32.data_ptr(),
                                                 plan.comp_kind8.data_ptr(), plan.node_start.data_ptr(),
                                                 plan.node_comp.data_ptr(), plan.num_nodes, g.data_ptr(),
                                                 grad_comp.data_ptr(), plan.loc_start.data_ptr(),
                                                 plan.loc_comp.data_ptr(), plan.num_locations,
                                                 grad_theta.data_ptr(), _stream_ptr(theta.device)))
        return grad_theta, None


class DifferentiableDem:
    """Structure of a circuit's DEM, compiled once on the GPU; probabilities
    (and their gradients) for any per-location rates theta.

    theta has one entry per fault location: each DEPOLARIZE1 / X_ERROR / M
    target and each DEPOLARIZE2 pair, in circuit order."""

    def __init__(self, circuit, correlation_level: CorrelationLevel = CorrelationLevel.L2,
                 device: int | torch.device = 0):
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        dev_index = self.device.index or 0
        lib = _native.load()
        h = ctypes.c_void_p()
        _native.check(lib.gp_grad_plan_create(circuit_text(circuit).encode(), int(correlation_level),
                                               dev_index, ctypes.byref(h)))
        try:
            sizes = np.zeros(4, np.uint32)
            _native.check(lib.gp_grad_plan_sizes(h, sizes.ctypes.data))
            L, K, R, C = (int(x) for x in sizes)
            self.num_locations, self.num_components, self.num_nodes, self.num_errors = L, K, R, C
            comp_loc = np.empty(max(K, 1), np.uint32)
            comp_kind = np.empty(max(K, 1), np.uint8)
            node_start = np.empty(R + 1, np.uint32)
            node_comp = np.empty(max(K, 1), np.uint32)
            class_start = np.empty(C + 1, np.uint32)
            class_node = np.empty(max(R, 1), np.uint32)
            _native.check(lib.gp_grad_plan_copy_arrays(h, comp_loc.ctypes.data, comp_kind.ctypes.data,
                                                       node_start.ctypes.data, node_comp.ctypes.data,
                                                       class_start.ctypes.data, class_node.ctypes.data))
            offsets = np.empty(C + 1, np.uint32)
            # targets: size from the final offset
            _native.check(lib.gp_grad_plan_dem_structure(h, offsets.ctypes.data, None))
            targets = np.empty(max(int(offsets[-1]), 1), np.uint32)
            _native.check(lib.gp_grad_plan_dem_structure(h, offsets.ctypes.data, targets.ctypes.data))
        finally:
            lib.gp_grad_plan_destroy(h)
        dev = self.device
        as_i32 = lambda a: torch.from_numpy(a.astype(np.int64)).to(dev)  # noqa: E731
        self.comp_loc = as_i32(comp_loc[:K])
        kind = torch.from_numpy(comp_kind[:K].astype(np.int64)).to(dev)
        self.is_dep1 = kind == 1
        self.is_dep2 = kind == 2
        # uint32 index arrays for the kernels (int32 storage, same bits)
        self.node_comp = torch.from_numpy(node_comp[:K].view(np.int32)).to(dev)
        self.node_start = torch.from_numpy(node_start.view(np.int32)).to(dev)
        self.class_node = torch.from_numpy(class_node[:R].view(np.int32)).to(dev)
        self.class_start = torch.from_numpy(class_start.view(np.int32)).to(dev)
        # the fused kernels' inputs: components' locations and kinds, and
        # each location's components (ascending), a CSR over locations
        self.comp_loc32 = torch.from_numpy(comp_loc[:max(K, 1)].view(np.int32)).to(dev)
        self.comp_kind8 = torch.from_numpy(comp_kind[:max(K, 1)].copy()).to(dev)
        order = np.argsort(comp_loc[:K], kind="stable").astype(np.uint32)
        loc_start = np.zeros(L + 1, np.uint32)
        np.cumsum(np.bincount(comp_loc[:K], minlength=L)[:L], out=loc_start[1:])
        self.loc_start = torch.from_numpy(loc_start.view(np.int32)).to(dev)
        self.loc_comp = torch.from_numpy(order[:max(K, 1)].view(np.int32) if K else np.zeros(1, np.int32)).to(dev)
        self.offsets = offsets
        self.targets = targets[: int(offsets[-1])]
        self.num_detectors = None
        self.num_observables = None

    def component_rates(self, theta: torch.Tensor) -> torch.Tensor:
        """The components' split rates as plain tensor operations (the
        reference formula the fused kernels are tested against)."""
        t = theta[self.comp_loc]
        s1 = 0.5 - 0.5 * torch.sqrt(torch.clamp(1 - 4 * t / 3, min=1e-300))
        s2 = 0.5 - 0.5 * torch.pow(torch.clamp(1 - 16 * t / 15, min=1e-300), 0.125)
        return torch.where(self.is_dep1, s1, torch.where(self.is_dep2, s2, t))

    def node_probabilities(self, theta: torch.Tensor) -> torch.Tensor:
        """q_n(theta) for every error node, fused on the GPU (NodeRates)."""
        return NodeRates.apply(theta, self)

    def probabilities(self, theta: torch.Tensor) -> torch.Tensor:
        """DEM error probabilities p_e(theta), in the DEM's output order."""
        if theta.shape != (self.num_locations,):
            raise ValueError(f"theta must have shape ({self.num_locations},)")
        q = self.node_probabilities(theta.to(self.device, torch.float64))
        return xor_merge(q, self.class_node, self.class_start, validate=False)

    def dem(self, theta: torch.Tensor, num_detectors: int, num_observables: int) -> DetectorErrorModel:
        p = self.probabilities(theta).detach().cpu().numpy()
        return DetectorErrorModel(p, self.offsets.copy(), self.targets.copy(), num_detectors, num_observables)


def logit_loss(p: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """L = sum_e w_e * logit(p_e)."""
    return torch.sum(w * (torch.log(p) - torch.log1p(-p)))
