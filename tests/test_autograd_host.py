"""Argument checks of the differentiable merge, on the CPU: the kernels read
index and seg_start as contiguous uint32 arrays on values' device, so
anything else is rejected before a pointer reaches the library, and a
repeated index is detected (the merge then runs over a gathered copy)."""

from __future__ import annotations

import pytest

torch = pytest.importorskip("torch")

from greenpeas_amd import autograd as ag  # noqa: E402


def i32(x):
    return torch.tensor(x, dtype=torch.int32)


V = torch.linspace(0.01, 0.4, 10, dtype=torch.float64)
IDX = i32([3, 1, 2, 0])
SEG = i32([0, 2, 4])


def test_valid_partial_permutation_passes():
    ag._check_index_tensors(V, IDX, SEG, require_cuda=False)
    assert ag._check_segments(V, IDX, SEG) is False


def test_repeats_are_detected_only_inside_the_segments():
    assert ag._check_segments(V, i32([3, 1, 3, 0]), SEG) is True
    # positions past seg_start[-1] are never read: repeats there do not count
    assert ag._check_segments(V, i32([3, 1, 2, 0, 3, 3]), SEG) is False


@pytest.mark.parametrize("seg, msg", [([1, 2, 4], "must be 0"), ([0, 3, 2], "non-decreasing"),
                                      ([0, 2, 5], "exceeds index.numel")])
def test_bad_segments_are_rejected(seg, msg):
    with pytest.raises(ValueError, match=msg):
        ag._check_segments(V, IDX, i32(seg))


def test_index_outside_values_is_rejected():
    with pytest.raises(ValueError, match="outside values"):
        ag._check_segments(V, i32([3, 1, 10, 0]), SEG)
    with pytest.raises(ValueError, match="outside values"):
        ag._check_segments(V, i32([3, -1, 2, 0]), SEG)


def test_wrong_dtypes_layouts_and_devices_are_rejected():
    with pytest.raises(TypeError, match="int32"):
        ag._check_index_tensors(V, IDX.long(), SEG, require_cuda=False)  # torch's default int64
    with pytest.raises(TypeError, match="int32"):
        ag._check_index_tensors(V, IDX, [0, 2, 4], require_cuda=False)
    with pytest.raises(TypeError, match="float64"):
        ag._check_index_tensors(V.float(), IDX, SEG, require_cuda=False)
    with pytest.raises(ValueError, match="contiguous"):
        ag._check_index_tensors(V, i32([3, 9, 1, 9, 2, 9, 0, 9])[::2], SEG, require_cuda=False)
    with pytest.raises(ValueError, match="at least one"):
        ag._check_index_tensors(V, IDX, i32([]), require_cuda=False)
    # CPU values never reach the kernels
    with pytest.raises(TypeError, match="CUDA"):
        ag.xor_merge(V, IDX, SEG)
    meta = torch.empty(4, dtype=torch.int32, device="meta")
    with pytest.raises(ValueError, match="is on meta"):
        ag._check_index_tensors(V, meta, SEG, require_cuda=False)
