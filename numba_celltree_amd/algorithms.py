"""Batch geometry kernels with the signatures of numba_celltree.algorithms
(numba_celltree/algorithms/__init__.py:1-23) and numba_celltree.geometry_utils, on the GPU.

Host inputs give host outputs; device tensors give device tensors.  Every device tensor
handed to the library is bound to a local name until the call has been enqueued: a
temporary freed earlier could be handed by torch's caching allocator to the next
temporary of the same call.
"""

import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native


def _f64(x) -> torch.Tensor:
    return D.to_dev(x, torch.float64)


def _pair_inputs(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    va, vb = _f64(vertices_a), _f64(vertices_b)
    fa, fb = D.faces_to_dev(faces_a, va.shape[0]), D.faces_to_dev(faces_b, vb.shape[0])
    ia, ib = D.index_to_dev(indices_a, fa.shape[0]), D.index_to_dev(indices_b, fb.shape[0])
    if ia.numel() != ib.numel():
        raise ValueError("indices_a and indices_b must have the same length")
    return va, vb, fa, fb, ia, ib


def polygons_intersect(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    """SAT intersection test per pair (separating_axis.py:58-75)."""
    io = D.Io(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b)
    va, vb, fa, fb, ia, ib = _pair_inputs(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b)
    n = int(ia.numel())
    out = D.empty(n, torch.uint8)
    _native.call(
        "ct_polygons_intersect", D.ptr(va), D.ptr(vb), D.ptr(fa), int(fa.shape[1]), D.ptr(fb),
        int(fb.shape[1]), D.ptr(ia), D.ptr(ib), n, D.ptr(out), D.stream(),
    )
    return io.out(out.bool())


def area_of_intersection(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    """Sutherland-Hodgman overlap area per pair (sutherland_hodgman.py:151-168)."""
    io = D.Io(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b)
    va, vb, fa, fb, ia, ib = _pair_inputs(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b)
    n = int(ia.numel())
    out = D.empty(n, torch.float64)
    _native.call(
        "ct_area_of_intersection", D.ptr(va), D.ptr(vb), D.ptr(fa), int(fa.shape[1]), D.ptr(fb),
        int(fb.shape[1]), D.ptr(ia), D.ptr(ib), n, D.ptr(out), D.stream(),
    )
    return io.out(out)


def box_area_of_intersection(bbox_coords, vertices, faces, indices_bbox, indices_face):
    """Overlap area of axis-aligned boxes with faces (sutherland_hodgman.py:171-187)."""
    io = D.Io(bbox_coords, vertices, faces, indices_bbox, indices_face)
    bx, v = _f64(bbox_coords), _f64(vertices)
    f = D.faces_to_dev(faces, v.shape[0])
    ib, iface = D.index_to_dev(indices_bbox, bx.shape[0]), D.index_to_dev(indices_face, f.shape[0])
    if ib.numel() != iface.numel():
        raise ValueError("indices_bbox and indices_face must have the same length")
    n = int(ib.numel())
    out = D.empty(n, torch.float64)
    _native.call(
        "ct_box_area_of_intersection", D.ptr(bx), D.ptr(v), D.ptr(f), int(f.shape[1]), D.ptr(ib),
        D.ptr(iface), n, D.ptr(out), D.stream(),
    )
    return io.out(out)


def _bary(fn, points, face_indices, faces, vertices, tolerance, wachspress: bool):
    io = D.Io(points, face_indices, faces, vertices)
    pts, v = _f64(points), _f64(vertices)
    f = D.faces_to_dev(faces, v.shape[0])
    fi = D.index_to_dev(face_indices, f.shape[0], allow_missing=True)
    n = int(pts.shape[0])
    if fi.numel() != n:
        raise ValueError("face_indices must have one entry per point")
    nmax = int(f.shape[1])
    out = D.empty((n, nmax if wachspress else 3), torch.float64)
    _native.call(fn, D.ptr(pts), D.ptr(fi), n, D.ptr(f), nmax, D.ptr(v), float(tolerance), D.ptr(out), D.stream())
    return io.out(out)


def barycentric_triangle_weights(points, face_indices, faces, vertices, tolerance):
    """(n, 3) triangle weights, zeros where face_indices == -1 (barycentric_triangle.py:46-64)."""
    return _bary("ct_barycentric_triangle_weights", points, face_indices, faces, vertices, tolerance, False)


def barycentric_wachspress_weights(points, face_indices, faces, vertices, tolerance):
    """(n, n_max_vert) Wachspress weights (barycentric_wachspress.py:88-107)."""
    return _bary("ct_barycentric_wachspress_weights", points, face_indices, faces, vertices, tolerance, True)


def counter_clockwise(vertices, faces) -> None:
    """Reorders clockwise face rows in place (geometry_utils.py:541-561)."""
    v = _f64(vertices)
    d_faces = D.faces_to_dev(faces, v.shape[0])
    n, nmax = int(d_faces.shape[0]), int(d_faces.shape[1])
    _native.call("ct_counter_clockwise", D.ptr(v), D.ptr(d_faces), n, nmax, D.stream())
    if isinstance(faces, torch.Tensor):
        faces.copy_(d_faces.to(faces.dtype))
    else:
        faces[...] = d_faces.cpu().numpy().astype(faces.dtype)


def build_face_bboxes(faces, vertices):
    """(n, 4) face bounding boxes (geometry_utils.py:443-456)."""
    io = D.Io(faces, vertices)
    v = _f64(vertices)
    f = D.faces_to_dev(faces, v.shape[0])
    n = int(f.shape[0])
    out = D.empty((n, 4), torch.float64)
    _native.call("ct_build_face_bboxes", D.ptr(f), n, int(f.shape[1]), D.ptr(v), D.ptr(out), D.stream())
    return io.out(out)


def build_edge_bboxes(edges, vertices, tolerance):
    """(n, 4) edge bounding boxes padded by tolerance (geometry_utils.py:473-487)."""
    io = D.Io(edges, vertices)
    v = _f64(vertices)
    e = D.faces_to_dev(edges, v.shape[0], allow_fill=False)
    n = int(e.shape[0])
    out = D.empty((n, 4), torch.float64)
    _native.call("ct_build_edge_bboxes", D.ptr(e), n, D.ptr(v), float(tolerance), D.ptr(out), D.stream())
    return io.out(out)


def sort_intersections_by_edge(ii, jj, xy, edge_coords):
    """Stable sort by (edge, t) (geometry_utils.py:564-574); ``ii`` must be grouped by edge
    in ascending order, as every query of this package returns it."""
    io = D.Io(ii, jj, xy, edge_coords)
    d_e = _f64(edge_coords)
    n = int(d_e.shape[0])
    d_ii, d_jj, d_xy = D.index_to_dev(ii, n), D.to_dev(jj, torch.int64), _f64(xy)
    if d_jj.numel() != d_ii.numel() or d_xy.numel() != 4 * d_ii.numel():
        raise ValueError("ii, jj and xy must describe the same rows")
    if d_ii.numel() > 1 and bool((d_ii[1:] < d_ii[:-1]).any()):
        raise ValueError("ii must be sorted ascending (grouped by edge)")
    counts = torch.bincount(d_ii, minlength=n).to(torch.int64)
    offsets, _ = D.exclusive_scan(counts)
    jj_out = torch.empty_like(d_jj)
    xy_out = torch.empty_like(d_xy)
    _native.call(
        "ct_sort_intersections_by_edge", D.ptr(d_e), D.ptr(offsets), D.ptr(d_ii), D.ptr(d_jj), D.ptr(d_xy),
        int(d_ii.numel()), D.ptr(jj_out), D.ptr(xy_out), D.stream(),
    )
    return io.out(d_ii), io.out(jj_out), io.out(xy_out)


def _bbox_stats(bb_coords):
    bb = _f64(bb_coords)
    if bb.ndim != 2 or int(bb.shape[1]) != 4:
        raise ValueError(f"bb_coords must have shape (n, 4), received {tuple(bb.shape)}")
    n = int(bb.shape[0])
    dist = D.empty((n, 3), torch.float64)
    stats = D.empty(5, torch.float64)
    _native.call("ct_bbox_stats", D.ptr(bb), n, D.ptr(dist), D.ptr(stats), D.stream())
    return dist, stats


def bbox_tree(bb_coords):
    """(xmin, xmax, ymin, ymax) over all boxes (celltree_base.py:20-25)."""
    io = D.Io(bb_coords)
    _, stats = _bbox_stats(bb_coords)
    return io.out(stats[:4].clone())


def bbox_distances(bb_coords):
    """(n, 3) dx, dy and diagonal length of every box (celltree_base.py:28-48)."""
    io = D.Io(bb_coords)
    dist, _ = _bbox_stats(bb_coords)
    return io.out(dist)


def default_tolerance(bb_diagonal) -> float:
    """max(MIN_TOLERANCE, TOLERANCE_FACTOR * max(bb_diagonal)) (celltree_base.py:51-52)."""
    from numba_celltree_amd.celltree_base import default_tolerance as _tol

    if isinstance(bb_diagonal, torch.Tensor):
        bb_diagonal = bb_diagonal.detach().cpu().numpy()
    return _tol(bb_diagonal)
