"""EdgeCellTree2d on MI355X (drop-in for numba_celltree/edge_celltree.py:25-146)."""

from typing import Optional, Tuple

import numpy as np
import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native
from numba_celltree_amd.cast import cast_edges, cast_vertices
from numba_celltree_amd.celltree_base import CellTree2dBase
from numba_celltree_amd.constants import MIN_TOLERANCE, TOLERANCE_FACTOR


class EdgeCellTree2d(CellTree2dBase):
    """
    Construct a cell tree from 2D vertices and an edges indexing array ``(n_edge, 2)``.
    Edge bounding boxes are padded by max(1e-15, 1e-12 * extent of the vertices) so that
    axis-aligned edges get a width (edge_celltree.py:60-64).
    """

    def __init__(self, vertices, edges, n_buckets: int = 4, cells_per_leaf: int = 2):
        if n_buckets < 2:
            raise ValueError("n_buckets must be >= 2")
        if cells_per_leaf < 1:
            raise ValueError("cells_per_leaf must be >= 1")
        vertices = cast_vertices(vertices, copy=True)
        d_vertices = D.to_dev(vertices, torch.float64)
        d_edges = D.faces_to_dev(edges, d_vertices.shape[0], allow_fill=False)
        if d_edges.ndim != 2 or d_edges.shape[1] != 2:
            raise ValueError("edges must have shape (n_edge, 2)")
        st = D.stream()
        tol = self._edge_tolerance(d_vertices)
        n = int(d_edges.shape[0])
        bb = D.empty((n, 4), torch.float64)
        _native.call("ct_build_edge_bboxes", D.ptr(d_edges), n, D.ptr(d_vertices), tol, D.ptr(bb), st)
        self._build_tree(d_vertices, d_edges, bb, n_buckets, cells_per_leaf)

    KIND = "edges"

    @staticmethod
    def _edge_tolerance(d_vertices: torch.Tensor) -> float:
        """max(1e-15, 1e-12 * the larger extent of the vertices) (edge_celltree.py:60-64)."""
        ext = D.empty(4, torch.float64)
        _native.call("ct_vertex_extent", D.ptr(d_vertices), int(d_vertices.shape[0]), D.ptr(ext), D.stream())
        xmin, xmax, ymin, ymax = ext.cpu().numpy().tolist()
        return max(MIN_TOLERANCE, TOLERANCE_FACTOR * max(xmax - xmin, ymax - ymin))

    @property
    def edges(self) -> np.ndarray:
        return self._elements_host()

    def locate_points(self, points, tolerance: Optional[float] = None):
        """Index of the edge each point lies on (within tolerance), else -1 (edge_celltree.py:85-112)."""
        return self._locate("ct_locate_points_on_edge", points, tolerance)

    def intersect_edges(self, edge_coords) -> Tuple:
        """Intersections of segments with the tree's edges, ordered along each segment
        (edge_celltree.py:114-146)."""
        io = D.Io(edge_coords)
        d_edges = D.to_dev(cast_edges(edge_coords), torch.float64)
        ii, jj, xy = self._locate_edges(d_edges, 0)
        return io.out(ii), io.out(jj), io.out(xy[:, 0].contiguous())
