"""d3d_gjk_epa_batch_noalloc (the fused GJK+EPA pipeline with its workspace in
the library's own device array, no allocation) against d3d_gjk_epa_batch_ws
(the same pipeline with a caller-owned workspace), through the C ABI.

Both entry points run the same kernels on the same inputs, so every output
must be identical to the last bit: distances, closest points, GJK status and
iteration counts, EPA rows and EPA status.  Two calls in a row share the
library's array; the second is smaller and uses another scene, so anything
the first call left in the array that the second failed to clear would show
up as a difference.
"""
import ctypes

import numpy as np
import pytest

from distance3d_amd import _native, kernels, workloads

pytestmark = pytest.mark.gpu


def _outputs(n, device):
    import torch
    f64 = dict(dtype=torch.float64, device=device)
    i32 = dict(dtype=torch.int32, device=device)
    # filled with a pattern so unwritten rows would differ between the runs
    return {"dist": torch.full((n,), -7.0, **f64), "cp_a": torch.full((n, 3), -7.0, **f64),
            "cp_b": torch.full((n, 3), -7.0, **f64), "status": torch.full((n,), -7, **i32),
            "iterations": torch.full((n,), -7, **i32), "epa_out": torch.full((n, 13), -7.0, **f64),
            "epa_status": torch.full((n,), -7, **i32)}


def _call(noalloc, dev, a, b, lanes=None):
    import torch
    lib = _native.lib()
    P = kernels.ptr
    n = a.numel()
    o = _outputs(n, a.device)
    prm = _native.gjk_params(lanes_per_pair=lanes)
    args = [dev.c, P(a), P(b), n, ctypes.byref(prm), 64, 128, 1e-8, P(o["dist"]), P(o["cp_a"]),
            P(o["cp_b"]), P(o["status"]), P(o["iterations"]), P(o["epa_out"]),
            P(o["epa_status"])]
    if noalloc:
        rc = lib.d3d_gjk_epa_batch_noalloc(*args, kernels.stream_ptr())
        name = "d3d_gjk_epa_batch_noalloc"
    else:
        ws = torch.empty(int(lib.d3d_gjk_epa_workspace_bytes(n)), dtype=torch.uint8,
                         device=a.device)
        rc = lib.d3d_gjk_epa_batch_ws(*args, P(ws), ws.numel(), kernels.stream_ptr())
        name = "d3d_gjk_epa_batch_ws"
    _native.check(rc, name)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in o.items()}


def _same(x, y):
    assert x.keys() == y.keys()
    for k in x:
        assert np.array_equal(x[k], y[k], equal_nan=x[k].dtype.kind == "f"), k


def test_noalloc_matches_workspace_entry_point_two_calls(gpu):
    import torch
    print("noalloc: call 1, mixed_pairs seed 9, 16384 pairs", flush=True)
    w1 = workloads.mixed_pairs(seed=9, n_pairs=_native.NOALLOC_MAX_PAIRS, max_vertices=1024)
    d1 = w1.scene.to_device("cuda")
    a1 = torch.from_numpy(w1.pair_a).cuda()
    b1 = torch.from_numpy(w1.pair_b).cuda()
    first = _call(True, d1, a1, b1)
    want1 = _call(False, d1, a1, b1)
    _same(first, want1)
    assert np.all(first["status"] >= 0)
    hits1 = int(np.sum(first["epa_status"] >= 0))
    assert hits1 > 0

    print("noalloc: call 2, epa_pairs seed 12, 3001 pairs", flush=True)
    w2 = workloads.epa_pairs(seed=12, n_pairs=3001)
    d2 = w2.scene.to_device("cuda")
    a2 = torch.from_numpy(w2.pair_a).cuda()
    b2 = torch.from_numpy(w2.pair_b).cuda()
    second = _call(True, d2, a2, b2)
    want2 = _call(False, d2, a2, b2)
    _same(second, want2)
    # about 90% of this workload's pairs intersect, so EPA ran on most of them
    assert np.mean(second["epa_status"] >= 0) > 0.8

    print("noalloc: call 3, the first input again after the smaller one", flush=True)
    _same(_call(True, d1, a1, b1), want1)


def test_noalloc_rejects_more_than_its_maximum(gpu):
    import torch
    lib = _native.lib()
    n = _native.NOALLOC_MAX_PAIRS + 1
    w = workloads.hull_pairs(n_pairs=n)
    dev = w.scene.to_device("cuda")
    a = torch.from_numpy(w.pair_a).cuda()
    b = torch.from_numpy(w.pair_b).cuda()
    o = _outputs(n, a.device)
    prm = _native.gjk_params()
    P = kernels.ptr
    rc = lib.d3d_gjk_epa_batch_noalloc(dev.c, P(a), P(b), n, ctypes.byref(prm), 64, 128, 1e-8,
                                       P(o["dist"]), P(o["cp_a"]), P(o["cp_b"]), P(o["status"]),
                                       None, P(o["epa_out"]), P(o["epa_status"]),
                                       kernels.stream_ptr())
    assert rc == 1
    assert b"D3D_NOALLOC_MAX_PAIRS" in lib.d3d_last_error()
    torch.cuda.synchronize()
    assert torch.all(o["dist"] == -7.0)
