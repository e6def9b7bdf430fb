"""ct_locate_points_ordered_noalloc (the allocation-free point location) through the C ABI.

Two calls in a row on one tree -- the first at the maximum size, the second smaller, so it
runs on the scratch contents the first call left -- each compared bit for bit with
ct_locate_points_ordered (the same walk with a caller-supplied workspace), and the second
also with the oracle.  Sizes above the maximum are rejected before any launch.
"""

import ctypes
import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _ordered(tree, pts, tol):
    import torch

    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    n = int(pts.shape[0])
    ws = D.empty(int(_native.lib().ct_locate_workspace_bytes(n)), torch.uint8)
    out = torch.full((n,), -7, dtype=torch.int64, device=pts.device)
    _native.call("ct_locate_points_ordered", tree.tree_ref, D.ptr(pts), n, tol, D.ptr(ws), ws.numel(), D.ptr(out),
                 D.stream())
    return out


def _noalloc(tree, pts, tol):
    import torch

    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    n = int(pts.shape[0])
    out = torch.full((n,), -7, dtype=torch.int64, device=pts.device)
    _native.call("ct_locate_points_ordered_noalloc", tree.tree_ref, D.ptr(pts), n, tol, D.ptr(out), D.stream())
    return out


def test_noalloc_matches_workspace_entry_point():
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import _native
    from oracle import refcpu as R

    v, f = S.jittered_grid(301, 1000.0, 0.25, 1.0 / 3.0, 3)
    tree = M.CellTree2d(v, f, -1)
    tol = tree._default_tolerance()
    n_max = _native.CT_NOALLOC_MAX_POINTS
    for n in (n_max, 150_001):
        print(f"[noalloc] mixed mesh of {len(f)} cells, {n} points", file=sys.stderr, flush=True)
        pts = torch.from_numpy(S.points(n, 1000.0, 0.95, 50.0, 7 + n % 5)).cuda()
        got = _noalloc(tree, pts, tol)
        want = _ordered(tree, pts, tol)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        assert int((got == -7).sum()) == 0
    ref = R.OracleCellTree2d(v, f, -1).locate_points(pts.cpu().numpy())
    assert np.array_equal(got.cpu().numpy(), ref)


def test_noalloc_rejects_oversized_input():
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    v, f, _ = S.triangle_grid_20k()
    tree = M.CellTree2d(v, f, -1)
    n = _native.CT_NOALLOC_MAX_POINTS + 1
    pts = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    rc = _native.lib().ct_locate_points_ordered_noalloc(tree.tree_ref, D.ptr(pts), n, 1e-12, D.ptr(out),
                                                        D.stream())
    assert rc == _native.CT_ERR_ARG
    assert "allocation-free maximum" in _native.last_error()
    # a zero-size call is a no-op
    assert _native.lib().ct_locate_points_ordered_noalloc(tree.tree_ref, ctypes.c_void_p(0), 0, 1e-12,
                                                          ctypes.c_void_p(0), D.stream()) == 0
