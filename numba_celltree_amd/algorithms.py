"""Batch geometry kernels with the signatures of numba_celltree.algorithms
(numba_celltree/algorithms/__init__.py:1-23) and numba_celltree.geometry_utils, on the GPU.

Host inputs give host outputs; device tensors give device tensors.
"""

import numpy as np
import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native


def _idx(x) -> torch.Tensor:
    return D.to_dev(x, torch.int64)


def _f64(x) -> torch.Tensor:
    return D.to_dev(x, torch.float64)


def polygons_intersect(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    """SAT intersection test per pair (separating_axis.py:58-75)."""
    io = D.Io(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b)
    fa, fb = D.faces_to_dev(faces_a), D.faces_to_dev(faces_b)
    ia, ib = _idx(indices_a), _idx(indices_b)
    n = int(ia.numel())
    out = D.empty(n, torch.uint8)
    _native.call(
        "ct_polygons_intersect", D.ptr(_f64(vertices_a)), D.ptr(_f64(vertices_b)), D.ptr(fa),
        int(fa.shape[1]), D.ptr(fb), int(fb.shape[1]), D.ptr(ia), D.ptr(ib), n, D.ptr(out), D.stream(),
    )
    return io.out(out.bool())


def area_of_intersection(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    """Sutherland-Hodgman overlap area per pair (sutherland_hodgman.py:151-168)."""
    io = D.Io(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b)
    fa, fb = D.faces_to_dev(faces_a), D.faces_to_dev(faces_b)
    ia, ib = _idx(indices_a), _idx(indices_b)
    n = int(ia.numel())
    out = D.empty(n, torch.float64)
    _native.call(
        "ct_area_of_intersection", D.ptr(_f64(vertices_a)), D.ptr(_f64(vertices_b)), D.ptr(fa),
        int(fa.shape[1]), D.ptr(fb), int(fb.shape[1]), D.ptr(ia), D.ptr(ib), n, D.ptr(out), D.stream(),
    )
    return io.out(out)


def box_area_of_intersection(bbox_coords, vertices, faces, indices_bbox, indices_face):
    """Overlap area of axis-aligned boxes with faces (sutherland_hodgman.py:171-187)."""
    io = D.Io(bbox_coords, vertices, faces, indices_bbox, indices_face)
    f = D.faces_to_dev(faces)
    ib, iface = _idx(indices_bbox), _idx(indices_face)
    n = int(ib.numel())
    out = D.empty(n, torch.float64)
    _native.call(
        "ct_box_area_of_intersection", D.ptr(_f64(bbox_coords)), D.ptr(_f64(vertices)), D.ptr(f),
        int(f.shape[1]), D.ptr(ib), D.ptr(iface), n, D.ptr(out), D.stream(),
    )
    return io.out(out)


def _bary(fn, points, face_indices, faces, vertices, tolerance, ncols):
    io = D.Io(points, face_indices, faces, vertices)
    pts, fi, f = _f64(points), _idx(face_indices), D.faces_to_dev(faces)
    n = int(pts.shape[0])
    out = D.empty((n, ncols(int(f.shape[1]))), torch.float64)
    _native.call(fn, D.ptr(pts), D.ptr(fi), n, D.ptr(f), int(f.shape[1]), D.ptr(_f64(vertices)),
                 float(tolerance), D.ptr(out), D.stream())
    return io.out(out)


def barycentric_triangle_weights(points, face_indices, faces, vertices, tolerance):
    """(n, 3) triangle weights, zeros where face_indices == -1 (barycentric_triangle.py:46-64)."""
    return _bary("ct_barycentric_triangle_weights", points, face_indices, faces, vertices, tolerance,
                 lambda nmax: 3)


def barycentric_wachspress_weights(points, face_indices, faces, vertices, tolerance):
    """(n, n_max_vert) Wachspress weights (barycentric_wachspress.py:88-107)."""
    return _bary("ct_barycentric_wachspress_weights", points, face_indices, faces, vertices, tolerance,
                 lambda nmax: nmax)


def counter_clockwise(vertices, faces) -> None:
    """Reorders clockwise face rows in place (geometry_utils.py:541-561)."""
    d_faces = D.faces_to_dev(faces)
    n, nmax = int(d_faces.shape[0]), int(d_faces.shape[1])
    _native.call("ct_counter_clockwise", D.ptr(_f64(vertices)), D.ptr(d_faces), n, nmax, D.stream())
    if isinstance(faces, torch.Tensor):
        faces.copy_(d_faces.to(faces.dtype))
    else:
        faces[...] = d_faces.cpu().numpy().astype(faces.dtype)


def build_face_bboxes(faces, vertices):
    """(n, 4) face bounding boxes (geometry_utils.py:443-456)."""
    io = D.Io(faces, vertices)
    f = D.faces_to_dev(faces)
    n = int(f.shape[0])
    out = D.empty((n, 4), torch.float64)
    _native.call("ct_build_face_bboxes", D.ptr(f), n, int(f.shape[1]), D.ptr(_f64(vertices)), D.ptr(out), D.stream())
    return io.out(out)


def build_edge_bboxes(edges, vertices, tolerance):
    """(n, 4) edge bounding boxes padded by tolerance (geometry_utils.py:473-487)."""
    io = D.Io(edges, vertices)
    e = D.faces_to_dev(edges)
    n = int(e.shape[0])
    out = D.empty((n, 4), torch.float64)
    _native.call("ct_build_edge_bboxes", D.ptr(e), n, D.ptr(_f64(vertices)), float(tolerance), D.ptr(out), D.stream())
    return io.out(out)


def sort_intersections_by_edge(ii, jj, xy, edge_coords):
    """Stable sort by (edge, t) (geometry_utils.py:564-574); ``ii`` must be grouped by edge
    in ascending order, as every query of this package returns it."""
    io = D.Io(ii, jj, xy, edge_coords)
    d_ii, d_jj, d_xy = _idx(ii), _idx(jj).clone(), _f64(xy).clone()
    d_e = _f64(edge_coords)
    n = int(d_e.shape[0])
    if d_ii.numel() > 1 and bool((d_ii[1:] < d_ii[:-1]).any()):
        raise ValueError("ii must be sorted ascending (grouped by edge)")
    counts = torch.bincount(d_ii, minlength=n).to(torch.int64) if d_ii.numel() else torch.zeros(n, dtype=torch.int64, device=d_e.device)
    offsets, _ = D.exclusive_scan(counts)
    _native.call("ct_sort_intersections_by_edge", D.ptr(d_e), n, D.ptr(offsets), D.ptr(d_jj), D.ptr(d_xy), D.stream())
    return io.out(d_ii), io.out(d_jj), io.out(d_xy)
