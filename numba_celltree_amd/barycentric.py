"""Differentiable barycentric weights (no reference counterpart for the adjoints; the
forwards are barycentric_triangle.py:28-64 and barycentric_wachspress.py:38-107, as
celltree.py:306-351 dispatches them: Wachspress when the faces have more than 3 columns).

Each point's containing face is located once (not differentiable) and held fixed, and so is
the forward's branch (Wachspress: interior weights, or the interpolation along an edge the
point is within tolerance of); the weights are then differentiable with respect to the
points and the mesh vertices, through the hand-written adjoints
ct_barycentric_triangle_backward and ct_barycentric_wachspress_backward.
"""

from typing import Optional

import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native
from numba_celltree_amd.regrid import _as_f64, _require


class BarycentricFunction(torch.autograd.Function):
    """Weights of points (n, 2) in faces face_index (n,) of a mesh (vertices (n_vertices, 2),
    faces (n_faces, k) int32), all on the current GPU; points and vertices float64.
    ``wachspress``: Wachspress weights (n, k) for k >= 3 (k <= 16 for the adjoint), else
    triangle weights (n, 3) of the first three vertices.  ``barycentric_weights`` converts and
    checks its inputs."""

    @staticmethod
    def forward(ctx, points, vertices, faces, face_index, tolerance: float, wachspress: bool = False):
        _require("points", points, (None, 2))
        _require("vertices", vertices, (None, 2))
        if faces.dtype != torch.int32 or faces.ndim != 2 or int(faces.shape[1]) < 3 or not faces.is_cuda:
            raise ValueError("faces must be a (n_faces, k >= 3) int32 device tensor")
        if face_index.dtype != torch.int64 or face_index.shape != points.shape[:1] or not face_index.is_cuda:
            raise ValueError("face_index must be an int64 device tensor with one entry per point")
        points = points.contiguous()
        vertices = vertices.contiguous()
        n, k = int(points.shape[0]), int(faces.shape[1])
        w = D.empty((n, k if wachspress else 3), torch.float64)
        fn = "ct_barycentric_wachspress_weights" if wachspress else "ct_barycentric_triangle_weights"
        _native.call(fn, D.ptr(points), D.ptr(face_index), n, D.ptr(faces), k, D.ptr(vertices), float(tolerance),
                     D.ptr(w), D.stream())
        ctx.save_for_backward(points, vertices, faces, face_index)
        ctx.tolerance = float(tolerance)
        ctx.wachspress = bool(wachspress)
        return w

    @staticmethod
    def backward(ctx, gw):
        points, vertices, faces, face_index = ctx.saved_tensors
        n = int(points.shape[0])
        gp = D.empty((n, 2), torch.float64) if ctx.needs_input_grad[0] else None
        gv = torch.zeros_like(vertices) if ctx.needs_input_grad[1] else None
        gw = gw.to(torch.float64).contiguous()
        if ctx.wachspress:
            _native.call("ct_barycentric_wachspress_backward", D.ptr(points), D.ptr(face_index), n, D.ptr(faces),
                         int(faces.shape[1]), D.ptr(vertices), ctx.tolerance, D.ptr(gw), D.ptr(gp), D.ptr(gv),
                         D.stream())
        else:
            _native.call("ct_barycentric_triangle_backward", D.ptr(points), D.ptr(face_index), n, D.ptr(faces),
                         int(faces.shape[1]), D.ptr(vertices), D.ptr(gw), D.ptr(gp), D.ptr(gv), D.stream())
        return gp, gv, None, None, None, None


def barycentric_weights(tree, points: torch.Tensor, vertices: Optional[torch.Tensor] = None,
                        tolerance: Optional[float] = None):
    """(face_index, weights) like ``CellTree2d.compute_barycentric_weights``, with weights
    differentiable in ``points`` and ``vertices`` (faces and branches held fixed): triangle
    weights on a 3-column tree, Wachspress weights on a wider one (up to 16 columns)."""
    k = int(tree.d_elements.shape[1])
    if k > 16:
        raise NotImplementedError("the Wachspress adjoint covers faces of at most 16 vertices")
    if tolerance is None:
        tolerance = tree._default_tolerance()
    if vertices is None:
        vertices = tree.d_vertices
    # Host arrays, other float dtypes and other devices are converted (differentiably).
    points, vertices = (_as_f64(x) for x in (points, vertices))
    _require("points", points, (None, 2))
    _require("vertices", vertices, (int(tree.d_vertices.shape[0]), 2))
    face = tree._locate_points_dev(points.detach().contiguous(), float(tolerance))
    w = BarycentricFunction.apply(points, vertices, tree.d_elements, face, float(tolerance), k > 3)
    return face, w
