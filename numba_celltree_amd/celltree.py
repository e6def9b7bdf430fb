"""CellTree2d on MI355X (drop-in for numba_celltree/celltree.py:37-351).

Construction and every query run as HIP kernels (libcelltree_hip.so): counter-clockwise
reordering, face bounding boxes, the level-synchronous tree build, traversals, the FP64
clippers and the barycentric weights.  Host inputs give host (numpy) outputs; device
tensors give device tensors.
"""

from typing import Optional, Tuple

import numpy as np
import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native
from numba_celltree_amd.cast import cast_bboxes, cast_edges, cast_faces, cast_vertices
from numba_celltree_amd.celltree_base import CellTree2dBase


class CellTree2d(CellTree2dBase):
    """
    Construct a cell tree from 2D vertices and a faces indexing array.

    Parameters
    ----------
    vertices: array of floats with shape ``(n_point, 2)`` (numpy or device tensor)
    faces: array of integers with shape ``(n_face, n_max_vert)``
    fill_value: int
        Fill value marking empty nodes in ``faces``.
    n_buckets: int, optional, default: 4 (2 .. 64)
    cells_per_leaf: int, optional, default: 2
    """

    def __init__(self, vertices, faces, fill_value: int, n_buckets: int = 4, cells_per_leaf: int = 2):
        if n_buckets < 2:
            raise ValueError("n_buckets must be >= 2")
        if cells_per_leaf < 1:
            raise ValueError("cells_per_leaf must be >= 1")
        vertices = cast_vertices(vertices, copy=True)
        faces = cast_faces(faces, fill_value)
        d_vertices = D.to_dev(vertices, torch.float64)
        d_faces = D.faces_to_dev(faces, d_vertices.shape[0])
        n, nmax = int(d_faces.shape[0]), int(d_faces.shape[1])
        if n == 0:
            raise ValueError("faces is empty")
        if nmax < 3:
            raise ValueError("faces must have at least 3 columns")
        st = D.stream()
        _native.call("ct_counter_clockwise", D.ptr(d_vertices), D.ptr(d_faces), n, nmax, st)
        bb = D.empty((n, 4), torch.float64)
        _native.call("ct_build_face_bboxes", D.ptr(d_faces), n, nmax, D.ptr(d_vertices), D.ptr(bb), st)
        self._build_tree(d_vertices, d_faces, bb, n_buckets, cells_per_leaf)

    @property
    def faces(self) -> np.ndarray:
        """The counter-clockwise reordered faces (int64, fill -1)."""
        return self._elements_host()

    # ----------------------------------------------------------------- queries
    # Point batches at least this large are processed in Morton order (ct_query_order, then
    # ct_locate_points_in_order): neighbouring lanes then share their root-to-leaf paths.
    ORDERED_MIN = 1 << 18

    def locate_points(self, points, tolerance: Optional[float] = None):
        """Index of the face containing each point, -1 outside (celltree.py:99-128)."""
        if tolerance is None:
            tolerance = self._default_tolerance()
        io = D.Io(points)
        pts = D.to_dev(cast_vertices(points), torch.float64)
        out = self._locate_points_dev(pts, float(tolerance))
        return io.out(out)

    def _locate_points_dev(self, pts: torch.Tensor, tolerance: float, order=None, sorted_pts=None) -> torch.Tensor:
        """``order`` / ``sorted_pts``: a processing order from ``D.query_order`` and, optionally,
        the points in that order (then read contiguously)."""
        n = int(pts.shape[0])
        out = D.empty(n, torch.int64)
        if order is None and self.ORDERED_MIN is not None and n >= self.ORDERED_MIN:
            # The order alone: writing the points in sorted order costs more in the sort
            # (0.3 -> 0.8 ms at 1e7 points) than the contiguous read saves in the traversal.
            order = D.query_order(0, pts, self._bbox)
        if order is not None:
            _native.call("ct_locate_points_in_order", self.tree_ref, D.ptr(pts), n, D.ptr(order), D.ptr(sorted_pts),
                         tolerance, D.ptr(out), D.stream())
        else:
            _native.call("ct_locate_points", self.tree_ref, D.ptr(pts), n, tolerance, D.ptr(out), D.stream())
        return out

    def _locate_boxes_dev(self, d_boxes: torch.Tensor):
        n = int(d_boxes.shape[0])
        st = D.stream()
        order, srt = self._order(1, d_boxes)
        split = self._split(1, d_boxes, order)
        if split is not None:
            subs, part_offsets, m = split
            counts = D.empty(max(m, 1), torch.int64)
            hits, slots = self._hit_record(m, self.BOX_HIT_SLOTS, self.BOX_HIT_BUDGET)
            _native.call("ct_locate_boxes_count_sub", self.tree_ref, D.ptr(d_boxes), D.ptr(subs), m, D.ptr(counts),
                         D.ptr(hits), slots, st)
            offsets, total = self._sub_offsets(order, n, part_offsets, counts)
            ii = D.empty(total, torch.int64)
            jj = D.empty(total, torch.int64)
            _native.call("ct_locate_boxes_write_sub", self.tree_ref, D.ptr(d_boxes), D.ptr(subs), m, D.ptr(offsets),
                         D.ptr(counts), D.ptr(hits), slots, D.ptr(ii), D.ptr(jj), st)
            return ii, jj
        counts = D.empty(n, torch.int64)
        hits, slots = self._hit_record(n, self.BOX_HIT_SLOTS, self.BOX_HIT_BUDGET)
        _native.call("ct_locate_boxes_count_rec", self.tree_ref, D.ptr(d_boxes), n, D.ptr(order), D.ptr(srt),
                     D.ptr(counts), D.ptr(hits), slots, st)
        offsets, total = D.exclusive_scan(counts)
        ii = D.empty(total, torch.int64)
        jj = D.empty(total, torch.int64)
        _native.call(
            "ct_locate_boxes_write_rec", self.tree_ref, D.ptr(d_boxes), n, D.ptr(order), D.ptr(srt),
            D.ptr(offsets), D.ptr(hits), slots, D.ptr(ii), D.ptr(jj), st,
        )
        return ii, jj

    def locate_boxes(self, bbox_coords) -> Tuple:
        """All (box, face) pairs whose bounding boxes strictly overlap (celltree.py:130-148)."""
        io = D.Io(bbox_coords)
        d_boxes = D.to_dev(cast_bboxes(bbox_coords), torch.float64)
        ii, jj = self._locate_boxes_dev(d_boxes)
        return io.out(ii), io.out(jj)

    def _compact(self, ii, jj, area, keep):
        offsets, total = D.exclusive_scan(keep)
        io_ = D.empty(total, torch.int64)
        jo = D.empty(total, torch.int64)
        ao = D.empty(total, torch.float64) if area is not None else None
        _native.call(
            "ct_compact_pairs", D.ptr(ii), D.ptr(jj), D.ptr(area), D.ptr(keep), D.ptr(offsets),
            int(ii.numel()), D.ptr(io_), D.ptr(jo), D.ptr(ao), D.stream(),
        )
        return io_, jo, ao

    def intersect_boxes(self, bbox_coords) -> Tuple:
        """Box-face pairs with overlap area > 0 (celltree.py:150-184)."""
        io = D.Io(bbox_coords)
        d_boxes = D.to_dev(cast_bboxes(bbox_coords), torch.float64)
        ii, jj = self._locate_boxes_dev(d_boxes)
        n = int(ii.numel())
        area = D.empty(n, torch.float64)
        keep = D.empty(n, torch.int64)
        _native.call(
            "ct_box_pairs", D.ptr(d_boxes), D.ptr(self.d_vertices), D.ptr(self.d_elements),
            int(self.d_elements.shape[1]), D.ptr(ii), D.ptr(jj), n, D.ptr(area), D.ptr(keep), D.stream(),
        )
        i, j, a = self._compact(ii, jj, area, keep)
        return io.out(i), io.out(j), io.out(a)

    def _prepare_faces(self, d_vertices: torch.Tensor, d_faces: torch.Tensor) -> torch.Tensor:
        n, nmax = int(d_faces.shape[0]), int(d_faces.shape[1])
        st = D.stream()
        _native.call("ct_counter_clockwise", D.ptr(d_vertices), D.ptr(d_faces), n, nmax, st)
        bb = D.empty((n, 4), torch.float64)
        _native.call("ct_build_face_bboxes", D.ptr(d_faces), n, nmax, D.ptr(d_vertices), D.ptr(bb), st)
        return bb

    def _face_pairs(self, d_vertices, d_faces, with_area: bool):
        """locate_boxes on the faces' bboxes, then the fused SAT (+ clip area) pair stage."""
        bb = self._prepare_faces(d_vertices, d_faces)
        ii, jj = self._locate_boxes_dev(bb)
        n = int(ii.numel())
        area = D.empty(n, torch.float64) if with_area else None
        keep = D.empty(n, torch.int64)
        _native.call(
            "ct_face_pairs", D.ptr(d_vertices), D.ptr(d_faces), int(d_faces.shape[1]),
            D.ptr(self.d_vertices), D.ptr(self.d_elements), int(self.d_elements.shape[1]),
            D.ptr(ii), D.ptr(jj), n, int(with_area), D.ptr(area), D.ptr(keep), D.stream(),
        )
        return self._compact(ii, jj, area, keep)

    def locate_faces(self, vertices, faces) -> Tuple:
        """Face pairs that intersect or touch (SAT), celltree.py:186-226.

        As in the reference, a host ``faces`` array is reordered counter-clockwise in place.
        """
        io = D.Io(vertices, faces)
        d_vertices = D.to_dev(cast_vertices(vertices), torch.float64)
        d_faces = D.faces_to_dev(faces, d_vertices.shape[0])
        i, j, _ = self._face_pairs(d_vertices, d_faces, with_area=False)
        if isinstance(faces, np.ndarray) and faces.ndim == 2:
            faces[...] = d_faces.cpu().numpy().astype(faces.dtype)
        elif isinstance(faces, torch.Tensor) and faces.is_cuda and faces.dtype == torch.int32:
            faces.copy_(d_faces)
        return io.out(i), io.out(j)

    def intersect_faces(self, vertices, faces, fill_value: int) -> Tuple:
        """Face pairs with overlap area > 0, and the area (celltree.py:228-269)."""
        io = D.Io(vertices, faces)
        d_vertices = D.to_dev(cast_vertices(vertices), torch.float64)
        d_faces = D.faces_to_dev(cast_faces(faces, fill_value), d_vertices.shape[0])
        i, j, a = self._face_pairs(d_vertices, d_faces, with_area=True)
        return io.out(i), io.out(j), io.out(a)

    def intersect_edges(self, edge_coords) -> Tuple:
        """Clipped sub-segment of every segment inside every face, ordered along each
        segment (celltree.py:271-304)."""
        io = D.Io(edge_coords)
        d_edges = D.to_dev(cast_edges(edge_coords), torch.float64)
        ii, jj, xy = self._locate_edges(d_edges, 1)
        return io.out(ii), io.out(jj), io.out(xy)

    def compute_barycentric_weights(self, points, tolerance: Optional[float] = None) -> Tuple:
        """Face index and barycentric weights of each point (celltree.py:306-351):
        Wachspress weights when faces have more than 3 columns, else triangle weights."""
        if tolerance is None:
            tolerance = self._default_tolerance()
        io = D.Io(points)
        pts = D.to_dev(cast_vertices(points), torch.float64)
        n = int(pts.shape[0])
        nmax = int(self.d_elements.shape[1])
        face = D.empty(n, torch.int64)
        w = D.empty((n, nmax), torch.float64)
        _native.call(
            "ct_compute_barycentric_weights", self.tree_ref, D.ptr(pts), n, float(tolerance),
            D.ptr(face), D.ptr(w), D.stream(),
        )
        return io.out(face), io.out(w)
