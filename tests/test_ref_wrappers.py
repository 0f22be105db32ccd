"""The reference-named C entry points never report a failure as a count: like the
reference's CUDA_CHECK (query/src/optix/OptixHelpers.h:15-29) they print the error and end
the process with EXIT_FAILURE; the dedup halves its input when the sort's scratch would not
fit (query/src/cuda/mesh_query_deduplication.cu:56-90). Each case runs in a child process."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import ctypes as C, sys
sys.path.insert(0, {root!r})
{pre}
from rayspace3d_amd import _lib
L = _lib.load()
{body}
print("returned", flush=True)
"""


def _run(body, env=None, timeout=120, pre=""):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, "-c", _CHILD.format(root=ROOT, body=body, pre=pre)], env=e,
                          capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("name,call", [
    ("count_hits_gpu", "L.count_hits_gpu(None, 5)"),
    ("compact_hits_gpu", "n = C.c_int(-1); L.compact_hits_gpu(None, None, 5, C.byref(n))"),
    ("exclusive_scan_gpu", "L.exclusive_scan_gpu(None, None, 7)"),
    ("merge_and_deduplicate_pairs_gpu", "L.merge_and_deduplicate_pairs_gpu(None, 3, None, 4, None)"),
    ("compact_hash_table_pairs", "L.compact_hash_table_pairs(None, 16, None, 8)"),
])
def test_wrappers_fail_fast_instead_of_returning_zero(name, call):
    # without a usable device (here) or with missing buffers (anywhere) the call must not
    # return: the process exits with status 1 and names the function and the error
    r = _run(call)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "returned" not in r.stdout
    assert f"librs3d: {name} failed:" in r.stderr


@pytest.mark.parametrize("call,expect", [
    ("print(L.count_hits_gpu(None, 0))", "0"),
    ("print(L.exclusive_scan_gpu(None, None, 0))", "0"),
    ("print(L.merge_and_deduplicate_pairs_gpu(None, 0, None, 0, None))", "0"),
    ("print(L.compact_hash_table_pairs(None, 0, None, 8))", "0"),
])
def test_empty_inputs_return_zero_without_a_device(call, expect):
    r = _run(call)
    assert r.returncode == 0 and r.stdout.split()[0] == expect, (r.stdout, r.stderr)


_DEDUP = r"""
import numpy as np
rng = np.random.Generator(np.random.PCG64(12))
n1, n2 = 23_000_000, 17_000_000
p = rng.integers(0, 3000, size=(n1 + n2, 2)).astype(np.int32)
d1 = torch.from_numpy(p[:n1]).cuda(); d2 = torch.from_numpy(p[n1:]).cuda()
out = torch.zeros((n1 + n2, 2), dtype=torch.int32, device="cuda")
torch.cuda.synchronize()
u = L.merge_and_deduplicate_pairs_gpu(C.c_void_p(d1.data_ptr()), n1, C.c_void_p(d2.data_ptr()), n2,
                                      C.c_void_p(out.data_ptr()))
np.save({path!r}, out[:u].cpu().numpy())
"""


@pytest.mark.gpu
def test_dedup_halving_path_gives_the_same_pairs(tmp_path):
    """40 M raw pairs: with 700 MB granted (RS3D_DEDUP_FREE_BYTES) one sort of all keys does
    not fit, so the dedup sorts halves, merges their unique keys and sorts again; the pair
    list equals numpy's unique set in the reference's (b, a) order and the unhalved call."""
    rng = np.random.Generator(np.random.PCG64(12))
    p = rng.integers(0, 3000, size=(40_000_000, 2)).astype(np.int32)
    key = (p[:, 1].astype(np.uint64) << np.uint64(32)) | p[:, 0].astype(np.uint64)
    uk = np.unique(key)
    exp = np.stack([(uk & np.uint64(0xffffffff)).astype(np.int32), (uk >> np.uint64(32)).astype(np.int32)], 1)
    for tag, env in (("halved", {"RS3D_DEDUP_FREE_BYTES": str(700 << 20)}), ("whole", {})):
        path = str(tmp_path / f"{tag}.npy")
        print(f"dedup child, {tag}", flush=True)
        r = _run(_DEDUP.format(path=path), env=env, timeout=300, pre="import torch")
        assert r.returncode == 0, (tag, r.stdout, r.stderr[-2000:])
        got = np.load(path)
        assert np.array_equal(got, exp), tag
