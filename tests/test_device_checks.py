"""The Python wrappers reject bad device arrays with ValueError before any library call
(the C ABI takes raw pointers and cannot check dtypes, shapes, lengths or index ranges).
Runs without a GPU: every check before the device check is reached with host tensors."""
from types import SimpleNamespace

import pytest
import torch

from rayspace3d_amd import _lib
from rayspace3d_amd import device as dv

CTX = SimpleNamespace(device=0, lib=None, handle=None)   # never reached: checks come first
BVH = SimpleNamespace(handle=None)


def _pts(n=10, dtype=torch.float32):
    return torch.zeros((n, 3), dtype=dtype)


@pytest.mark.parametrize("points,msg", [
    (_pts(dtype=torch.float64), "float32"),
    (torch.zeros((10, 4)), "shape"),
    (torch.zeros((3, 10)).t(), "contiguous"),
    (torch.zeros((10, 6))[:, ::2], "contiguous"),
    (_pts(), "cuda:0"),
    ("not a tensor", "torch tensor"),
])
def test_trace_points_rejects_bad_points(points, msg):
    with pytest.raises(ValueError, match=msg):
        dv.trace_points(CTX, BVH, points)


def test_trace_points_rejects_index_map_problems():
    p = _pts(10)
    # torch.randperm's default int64 would be read as pairs of uint32
    with pytest.raises(ValueError, match="4-byte"):
        dv.check_index_map(torch.arange(10), 10, 0)
    with pytest.raises(ValueError, match=r"\[0, 10\)"):
        dv.check_index_map(torch.tensor([0, 3, 10], dtype=torch.int32), 10, 0)
    with pytest.raises(ValueError, match=r"\[0, 10\)"):
        dv.check_index_map(torch.tensor([-1, 3], dtype=torch.int32), 10, 0)
    assert p.shape == (10, 3)


@pytest.mark.parametrize("kind,out,msg", [
    (_lib.OUT_OBJ_I32, torch.zeros(10, dtype=torch.int64), "int32"),
    (_lib.OUT_OBJ_I32, torch.zeros(9, dtype=torch.int32), "9 slots, 10 needed"),
    (_lib.OUT_FLAG_U8, torch.zeros(10, dtype=torch.int32), "uint8"),
    (_lib.OUT_RAY_RESULT, torch.zeros(10, dtype=torch.int32), "shape"),
    (_lib.OUT_RAY_RESULT, torch.zeros((10, 2), dtype=torch.int32), "cuda:0"),
    (7, torch.zeros(10, dtype=torch.int32), "out_kind"),
])
def test_out_checks(kind, out, msg):
    with pytest.raises(ValueError, match=msg):
        dv.check_out(out, kind, 10, 0)


def test_filter_refine_and_stats_reject_bad_inputs():
    fr = dv.FilterRefinePipeline(CTX)
    with pytest.raises(ValueError, match="float32"):
        fr.execute(BVH, _pts(dtype=torch.float16))
    with pytest.raises(ValueError, match="slots"):
        dv.check_out(torch.zeros(3, dtype=torch.int32), _lib.OUT_OBJ_I32, 10, 0)
    with pytest.raises(ValueError, match="shape"):
        dv.trace_stats(CTX, BVH, torch.zeros(30))
    with pytest.raises(ValueError, match="shape"):
        dv.axis_grid_stats(CTX, BVH, torch.zeros(30))


def test_mesh_array_checks():
    V = torch.zeros((4, 3))
    I = torch.tensor([[0, 1, 2], [1, 2, 3]], dtype=torch.int32)
    O = torch.zeros(2, dtype=torch.int32)
    with pytest.raises(ValueError, match=r"\[0, 4\)"):
        dv.check_mesh_arrays(V, torch.tensor([[0, 1, 4]], dtype=torch.int32), None, 0)
    with pytest.raises(ValueError, match="4-byte"):
        dv.check_mesh_arrays(V, I.long(), O, 0)
    with pytest.raises(ValueError, match="one per triangle"):
        dv.check_mesh_arrays(V, I, torch.zeros(3, dtype=torch.int32), 0)
    with pytest.raises(ValueError, match="cuda:0"):
        dv.check_mesh_arrays(V, I, O, 0)


def test_other_wrappers():
    with pytest.raises(ValueError, match="int64"):
        dv.sort_unique_u64(CTX, torch.zeros(5, dtype=torch.int32))
    with pytest.raises(ValueError, match="float32"):
        dv.tri_intersect(CTX, torch.zeros((2, 3, 3), dtype=torch.float64), torch.zeros((2, 3, 3)))
    with pytest.raises(ValueError, match="points"):
        dv._batch_array([{"kind": "contain", "points": _pts(dtype=torch.int32),
                          "out": torch.zeros(10, dtype=torch.int32), "bvh": BVH}], 0)
