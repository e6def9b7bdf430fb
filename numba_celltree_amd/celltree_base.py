"""Shared base of CellTree2d / EdgeCellTree2d (numba_celltree/celltree_base.py:20-114).

A tree lives on the GPU: vertices, elements, bounding boxes, nodes and bb_indices are
device tensors, built by the HIP library.  The reference's public attributes (``nodes``,
``bb_indices``, ``bb_coords``, ``bbox``, ``bb_distances``, ``vertices``,
``celltree_data``) are host numpy copies made on first access and cached, so an edit of
``tree.nodes`` (as the reference's tests do) is seen by ``node_bounds`` and
``validate_node_bounds``.  Queries always use the device tree as it was built.
"""

import ctypes
from typing import Optional

import numpy as np
import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native
from numba_celltree_amd.constants import MIN_TOLERANCE, TOLERANCE_FACTOR, CellTreeData


# Side of the start-node grid of a tree (ct_start_grid): the power of two at or above
# sqrt(n_nodes), at most START_GRID_MAX; trees of fewer than 64 nodes get none.  An int
# overrides the rule (0: no grid); results are the same either way.
START_GRID_N = None
# Every tree also keeps its cell boxes in bb_indices order (ct_pack_leaf_bboxes, 32 bytes per
# cell) for box and segment queries; False builds trees without them (same results).
LEAF_BBOXES = True
START_GRID_MAX = 1024


def start_grid_size(n_nodes: int) -> int:
    if START_GRID_N is not None:
        return int(START_GRID_N)
    if n_nodes < 64:
        return 0
    g = 1 << int(np.ceil(np.log2(np.sqrt(n_nodes))))
    return min(g, START_GRID_MAX)


def default_tolerance(bb_diagonal) -> float:  # celltree_base.py:51-52
    return max(MIN_TOLERANCE, TOLERANCE_FACTOR * max(bb_diagonal))


class CellTree2dBase:
    # Box and segment batches at least this large are processed in the order of
    # ct_query_order (similar work per wavefront); results do not depend on it.
    ORDER_MIN = 4096
    # Whether ordered box / segment passes also read their queries from a sorted copy.
    SORTED_QUERIES = False

    def _order(self, kind: int, queries: torch.Tensor):
        """(order, sorted queries or None) for a box / segment batch, or (None, None)."""
        if int(queries.shape[0]) < self.ORDER_MIN:
            return None, None
        if self.SORTED_QUERIES:
            return D.query_order(kind, queries, self._bbox, with_sorted=True)
        return D.query_order(kind, queries, self._bbox), None

    # ------------------------------------------------------------------ build
    def _build_tree(self, vertices: torch.Tensor, elements: torch.Tensor, bb_coords: torch.Tensor,
                    n_buckets: int, cells_per_leaf: int) -> None:
        n = int(elements.shape[0])
        if n == 0:
            raise ValueError("a cell tree needs at least one element")
        L = _native.lib()
        st = D.stream()
        ws = D.empty(int(L.ct_build_workspace_bytes(n, n_buckets)), torch.uint8)
        cap = int(L.ct_max_nodes(n))
        nodes_raw = D.empty(32 * cap, torch.uint8)
        bb_indices = D.empty(n, torch.int32)
        n_nodes = ctypes.c_int64(0)
        depth = ctypes.c_int32(0)
        _native.call(
            "ct_build", D.ptr(bb_coords), n, n_buckets, cells_per_leaf, D.ptr(ws), ws.numel(),
            D.ptr(nodes_raw), D.ptr(bb_indices), ctypes.byref(n_nodes), ctypes.byref(depth), st,
        )
        del ws
        self._adopt(vertices, elements, bb_coords, nodes_raw[: 32 * n_nodes.value], bb_indices,
                    int(n_nodes.value), int(depth.value), n_buckets, cells_per_leaf)

    def _adopt(self, vertices: torch.Tensor, elements: torch.Tensor, bb_coords: torch.Tensor,
               nodes_raw: torch.Tensor, bb_indices: torch.Tensor, n_nodes: int, depth: int,
               n_buckets: int, cells_per_leaf: int) -> None:
        """Takes a built (or imported) tree's device arrays and derives the rest: the bbox
        statistics, the ct_tree descriptor and the device-only layouts."""
        n = int(elements.shape[0])
        st = D.stream()
        distances = D.empty((n, 3), torch.float64)
        stats = D.empty(5, torch.float64)
        _native.call("ct_bbox_stats", D.ptr(bb_coords), n, D.ptr(distances), D.ptr(stats), st)
        host = stats.cpu().numpy()

        self.n_buckets = n_buckets
        self.cells_per_leaf = cells_per_leaf
        self.d_vertices = vertices
        self.d_elements = elements
        self.d_bb_coords = bb_coords
        self.d_nodes = nodes_raw
        self.d_bb_indices = bb_indices
        self.d_bb_distances = distances
        self.n_nodes = int(n_nodes)
        self.depth = int(depth)
        self._bbox = host[:4].copy()
        self._max_diagonal = float(host[4])
        t = _native.CtTree()
        t.vertices = vertices.data_ptr()
        t.elements = elements.data_ptr()
        t.bb_coords = bb_coords.data_ptr()
        t.nodes = self.d_nodes.data_ptr()
        t.bb_indices = bb_indices.data_ptr()
        t.n_vertices = int(vertices.shape[0])
        t.n_elements = n
        t.n_nodes = self.n_nodes
        t.n_max_vert = int(elements.shape[1])
        t.depth = self.depth
        for k in range(4):
            t.bbox[k] = float(self._bbox[k])
        t.leaf_coords = None
        self._ctree = t
        self._host = {}
        nmax = int(elements.shape[1])
        self.d_leaf_coords = None
        if 3 <= nmax <= 8:
            # Leaf-ordered vertex copy: a leaf's cells become one contiguous read.
            self.d_leaf_coords = D.empty((n, nmax, 2), torch.float64)
            _native.call("ct_pack_leaf_coords", self.tree_ref, D.ptr(self.d_leaf_coords), st)
            t.leaf_coords = self.d_leaf_coords.data_ptr()
        t.leaf_bb = None
        self.d_leaf_bb = None
        if LEAF_BBOXES and n > 0:
            # Leaf-ordered cell boxes: box and segment queries read a leaf's boxes directly.
            self.d_leaf_bb = D.empty((n, 4), torch.float64)
            _native.call("ct_pack_leaf_bboxes", self.tree_ref, D.ptr(self.d_leaf_bb), st)
            t.leaf_bb = self.d_leaf_bb.data_ptr()
        t.parents = None
        self.d_parents = None
        if self.depth > _native.CT_MAX_DEPTH:
            # Deeper than the largest traversal stack: queries climb parent links instead.
            self.d_parents = D.empty(self.n_nodes, torch.int32)
            _native.call("ct_node_parents", self.tree_ref, D.ptr(self.d_parents), st)
            t.parents = self.d_parents.data_ptr()
        t.start_grid = None
        t.start_grid_n = 0
        self.d_start_grid = None
        g = start_grid_size(self.n_nodes)
        if g:
            # Per grid cell, the node where every point walk through it starts.
            self.d_start_grid = D.empty((g, g), torch.int32)
            _native.call("ct_start_grid", self.tree_ref, g, D.ptr(self.d_start_grid), st)
            t.start_grid = self.d_start_grid.data_ptr()
            t.start_grid_n = g
        t.node_boxes = None
        self.d_node_boxes = None
        if nmax >= 3:
            # Tight subtree boxes: segment-face queries skip subtrees a segment cannot reach.
            self.d_node_boxes = D.empty((self.n_nodes + 1, 4), torch.float64)
            scratch = D.empty(self.n_nodes, torch.int32)
            _native.call("ct_node_boxes", self.tree_ref, D.ptr(scratch), D.ptr(self.d_node_boxes), st)
            del scratch
            t.node_boxes = self.d_node_boxes.data_ptr()

    # ------------------------------------------------------------ export / import
    KIND = "faces"

    def to_arrays(self) -> dict:
        """The tree as host arrays in the reference's layouts (numba_celltree_amd.treeio):
        vertices, elements (CCW faces or edges, int64), nodes (NodeDType), bb_indices (intp),
        bb_coords, n_buckets, cells_per_leaf and kind.  Copies from the device."""
        return {
            "vertices": self.d_vertices.cpu().numpy(),
            "elements": self._elements_host().copy(),
            "nodes": D.nodes_to_host(self.d_nodes, self.n_nodes),
            "bb_indices": self.d_bb_indices.cpu().numpy().astype(np.intp),
            "bb_coords": self.d_bb_coords.cpu().numpy(),
            "n_buckets": int(self.n_buckets),
            "cells_per_leaf": int(self.cells_per_leaf),
            "kind": self.KIND,
        }

    def _compact_arrays(self) -> dict:
        """The tree in the compact file form: int32 elements and bb_indices, nodes as device
        records, no bb_coords (numba_celltree_amd.treeio)."""
        from numba_celltree_amd import treeio

        return {
            "vertices": self.d_vertices.cpu().numpy(),
            "elements": self.d_elements.cpu().numpy(),
            "nodes": self.d_nodes.cpu().numpy().view(treeio.NODE_RECORD),
            "bb_indices": self.d_bb_indices.cpu().numpy(),
            "n_buckets": int(self.n_buckets),
            "cells_per_leaf": int(self.cells_per_leaf),
            "kind": self.KIND,
        }

    @classmethod
    def from_arrays(cls, vertices, elements, nodes, bb_indices, bb_coords=None, n_buckets: int = 4,
                    cells_per_leaf: int = 2, kind: Optional[str] = None, **_ignored):
        """A tree from exported arrays (``to_arrays``, or a reference tree's own attributes:
        vertices, faces / edges, nodes, bb_indices, bb_coords), without rebuilding it.  The
        arrays are validated on the host; given bb_coords must be the elements' bounding
        boxes; the device layouts are derived again."""
        from numba_celltree_amd import treeio

        arrays = treeio.validate({
            "vertices": vertices, "elements": elements, "nodes": nodes, "bb_indices": bb_indices,
            "bb_coords": bb_coords, "n_buckets": n_buckets, "cells_per_leaf": cells_per_leaf,
            "kind": cls.KIND if kind is None else kind,
        })
        return cls._from_validated(arrays)

    @classmethod
    def _from_validated(cls, arrays: dict):
        if arrays["kind"] != cls.KIND:
            raise ValueError(f"{cls.__name__} holds a {cls.KIND} tree, the arrays are a {arrays['kind']} tree")
        tree = cls.__new__(cls)
        tree._import(arrays)
        return tree

    def _import(self, a: dict) -> None:
        """Uploads a validated tree (treeio.validate's compact form) and derives the rest."""
        d_vertices = D.to_dev(a["vertices"], torch.float64)
        faces = a["kind"] == "faces"
        d_elements = D.faces_to_dev(a["elements"], d_vertices.shape[0], allow_fill=faces)
        n, nmax = int(d_elements.shape[0]), int(d_elements.shape[1])
        st = D.stream()
        d_bb = D.empty((n, 4), torch.float64)
        if faces:
            _native.call("ct_build_face_bboxes", D.ptr(d_elements), n, nmax, D.ptr(d_vertices), D.ptr(d_bb), st)
        else:
            _native.call("ct_build_edge_bboxes", D.ptr(d_elements), n, D.ptr(d_vertices),
                         self._edge_tolerance(d_vertices), D.ptr(d_bb), st)
        if a.get("bb_coords") is not None:
            # given boxes must be the elements' own: queries prune with them
            given = D.to_dev(a["bb_coords"], torch.float64)
            same = (given == d_bb) | (torch.isnan(given) & torch.isnan(d_bb))
            if not bool(same.all()):
                raise ValueError("bb_coords are not the bounding boxes of the elements")
        d_nodes = D.to_dev(a["nodes"].view(np.uint8), torch.uint8)
        d_bb_indices = D.to_dev(a["bb_indices"], torch.int32)
        self._adopt(d_vertices, d_elements, d_bb, d_nodes, d_bb_indices, len(a["nodes"]), a["depth"],
                    a["n_buckets"], a["cells_per_leaf"])

    def save(self, file) -> None:
        """Writes the tree to ``file`` (a path or binary file object) in the compact file form
        of numba_celltree_amd.treeio."""
        from numba_celltree_amd import treeio

        treeio.save(file, self._compact_arrays())

    @classmethod
    def load(cls, file):
        """Reads a tree written by ``save`` (validated on the host; not rebuilt)."""
        from numba_celltree_amd import treeio

        return cls._from_validated(treeio.load(file))

    def __getstate__(self):
        return self._compact_arrays()

    def __setstate__(self, state):
        from numba_celltree_amd import treeio

        self._import(treeio.validate(state))

    @property
    def tree_ref(self):
        return ctypes.byref(self._ctree)

    def _default_tolerance(self) -> float:
        return max(MIN_TOLERANCE, TOLERANCE_FACTOR * self._max_diagonal)

    # ------------------------------------------------------- host attributes
    def _cached(self, key, make):
        if key not in self._host:
            self._host[key] = make()
        return self._host[key]

    @property
    def nodes(self) -> np.ndarray:
        return self._cached("nodes", lambda: D.nodes_to_host(self.d_nodes, self.n_nodes))

    @property
    def bb_indices(self) -> np.ndarray:
        return self._cached("bb_indices", lambda: self.d_bb_indices.cpu().numpy().astype(np.intp))

    @property
    def bb_coords(self) -> np.ndarray:
        return self._cached("bb_coords", lambda: self.d_bb_coords.cpu().numpy())

    @property
    def bb_distances(self) -> np.ndarray:
        return self._cached("bb_distances", lambda: self.d_bb_distances.cpu().numpy())

    @property
    def bbox(self) -> np.ndarray:
        return self._bbox

    @property
    def vertices(self) -> np.ndarray:
        return self._cached("vertices", lambda: self.d_vertices.cpu().numpy())

    def _elements_host(self) -> np.ndarray:
        return self._cached("elements", lambda: self.d_elements.cpu().numpy().astype(np.intp))

    @property
    def celltree_data(self) -> CellTreeData:
        return CellTreeData(
            self._elements_host(), self.vertices, self.nodes, self.bb_indices, self.bb_coords,
            self.bbox, self.cells_per_leaf,
        )

    # ------------------------------------------------------------ validation
    def _nodes_for_check(self) -> torch.Tensor:
        if "nodes" in self._host:  # the host copy may have been edited
            return D.nodes_to_device(self._host["nodes"])
        return self.d_nodes

    @property
    def node_bounds(self) -> np.ndarray:
        """Bounds (xmin, xmax, ymin, ymax) of every node (query.py:568-622)."""
        return self._node_bounds(self._nodes_for_check()).cpu().numpy()

    def _node_bounds(self, d_nodes: torch.Tensor) -> torch.Tensor:
        n = d_nodes.numel() // 32
        out = D.empty((n, 4), torch.float64)
        parent = D.empty(n, torch.int32)
        bbox = (ctypes.c_double * 4)(*[float(x) for x in self._bbox])
        _native.call("ct_collect_node_bounds", D.ptr(d_nodes), n, bbox, D.ptr(parent), D.ptr(out), D.stream())
        return out

    def validate_node_bounds(self) -> np.ndarray:
        """Whether every node's children / cells lie in its bounds (query.py:625-664)."""
        d_nodes = self._nodes_for_check()
        nb = self._node_bounds(d_nodes)
        n = d_nodes.numel() // 32
        t = _native.CtTree.from_buffer_copy(self._ctree)
        t.nodes = d_nodes.data_ptr()
        t.n_nodes = n
        out = D.empty(n, torch.uint8)
        _native.call("ct_validate_node_bounds", ctypes.byref(t), D.ptr(nb), D.ptr(out), D.stream())
        return out.cpu().numpy().astype(bool)

    def to_dict_of_lists(self):  # celltree_base.py:83-114
        d = {}
        for parent_index, left in enumerate(self.nodes["child"]):
            d[parent_index] = [] if left == -1 else [int(left), int(left) + 1]
        return d

    # ---------------------------------------------------------------- points
    def _locate(self, fn: str, points, tolerance: Optional[float]):
        from numba_celltree_amd.cast import cast_vertices

        if tolerance is None:
            tolerance = self._default_tolerance()
        io = D.Io(points)
        pts = D.to_dev(cast_vertices(points), torch.float64)
        n = int(pts.shape[0])
        out = D.empty(n, torch.int64)
        _native.call(fn, self.tree_ref, D.ptr(pts), n, float(tolerance), D.ptr(out), D.stream())
        return io.out(out)

    # Hit cells the count pass records per segment (slot-major, so only slots in use are
    # touched); segments with at most that many hits are written from the record instead of
    # a second tree walk.  The long segments carry most of the rows, so the record pays off
    # only when it covers them: up to EDGE_HIT_SLOTS slots within EDGE_HIT_BUDGET int32
    # entries (1 GiB); below 4 slots the write pass walks the tree again.
    EDGE_HIT_SLOTS = 256
    EDGE_HIT_BUDGET = 1 << 28

    # With the cells, the count pass of split segment queries keeps the clipped segments of the
    # first EDGE_XY_SLOTS hits (32 bytes each, within EDGE_XY_BUDGET bytes), so that the write
    # pass copies the rows of a subquery whose hits all fit instead of clipping its cells again.
    EDGE_XY_SLOTS = 32
    EDGE_XY_BUDGET = 1 << 31

    # Box queries: a box's hits are its output rows, and wide boxes have hundreds (up to ~900
    # for boxes of 20 cell widths on a million-cell mesh), so the record is larger: up to BOX_HIT_SLOTS slots within BOX_HIT_BUDGET int32
    # entries (4 GiB).  All four are class defaults; setting them on a tree instance changes
    # that tree only (results never depend on them).
    BOX_HIT_SLOTS = 1024
    BOX_HIT_BUDGET = 1 << 30

    @staticmethod
    def _hit_record(n: int, max_slots: int, budget: int):
        """(hits, slots): a record of up to max_slots slots per query within budget int32
        entries.  When the device cannot hold it, fewer slots are tried, down to no record
        (slots 0: the write pass walks the tree again), so memory pressure costs speed, not
        the call."""
        slots = min(int(max_slots), int(budget) // max(n, 1))
        while slots >= 4:
            try:
                return D.empty(slots * max(n, 1), torch.int32), slots
            except torch.cuda.OutOfMemoryError:
                torch.cuda.empty_cache()
                slots //= 2
        return None, 0

    @staticmethod
    def _xy_record(n: int, slots: int, max_slots: int, budget: int):
        """(hit_xy, xy_slots): the clipped-segment record beside a cell record of `slots` slots,
        up to max_slots slots within budget bytes; (None, 0) below 4 slots or when the device
        cannot hold it."""
        xs = min(int(max_slots), int(slots), int(budget) // (32 * max(n, 1)))
        while xs >= 4:
            try:
                return D.empty((xs * max(n, 1), 4), torch.float64), xs
            except torch.cuda.OutOfMemoryError:
                torch.cuda.empty_cache()
                xs //= 2
        return None, 0

    # Work splitting of the largest box and segment queries (ct_split_queries): a query expected
    # to meet more than SPLIT_MIN_CELLS cells is cut into at most SPLIT_MAX_PARTS subqueries by
    # subtree, expanding subtrees of more than SPLIT_LEAF_CELLS (boxes) or EDGE_SPLIT_LEAF_CELLS
    # (segments) cells; results are identical.  Segments take larger leaves: fewer subqueries
    # keep their hit records small.
    SPLIT_MIN_CELLS = 16
    SPLIT_LEAF_CELLS = 32
    EDGE_SPLIT_LEAF_CELLS = 64
    SPLIT_MAX_PARTS = 16

    def _split(self, kind: int, queries: torch.Tensor, order):
        """(subqueries (m, 2) int32, part offsets (n + 1), m) for kind 1 boxes, 2 segment-face,
        3 segment-edge, or None when the batch is not split.  Batches too small to be ordered
        (order None: queries in their own order) are split too: with few lanes, one long walk
        would set the call's time."""
        n = int(queries.shape[0])
        if n == 0 or self.depth > _native.CT_MAX_DEPTH or int(self.SPLIT_MAX_PARTS) <= 1:
            return None
        st = D.stream()
        leaf = self.SPLIT_LEAF_CELLS if kind == 1 else self.EDGE_SPLIT_LEAF_CELLS
        args = (self.tree_ref, kind, D.ptr(queries), n, D.ptr(order), int(self.SPLIT_MIN_CELLS), int(leaf),
                int(self.SPLIT_MAX_PARTS))
        parts = D.empty(n, torch.int64)
        keep = D.empty(n * int(self.SPLIT_MAX_PARTS), torch.int32)  # frontiers, for the write call
        _native.call("ct_split_queries", *args, D.ptr(parts), None, None, D.ptr(keep), st)
        part_offsets, m = D.exclusive_scan(parts)
        subs = D.empty((max(m, 1), 2), torch.int32)
        _native.call("ct_split_queries", *args, None, D.ptr(part_offsets), D.ptr(subs), D.ptr(keep), st)
        return subs, part_offsets, m

    def _sub_offsets(self, order, n: int, part_offsets, counts):
        """Turns per-subquery counts into prefixes within each query; returns the query offsets
        (n + 1) and the row total."""
        per_query = D.empty(n, torch.int64)
        _native.call("ct_subquery_offsets", D.ptr(order), n, D.ptr(part_offsets), D.ptr(counts), D.ptr(per_query),
                     D.stream())
        return D.exclusive_scan(per_query)

    def _sort_edges(self, edge_coords, offsets, ii, jj, xy, total):
        jj_sorted = torch.empty_like(jj)
        xy_sorted = torch.empty_like(xy)
        _native.call(
            "ct_sort_intersections_by_edge", D.ptr(edge_coords), D.ptr(offsets), D.ptr(ii), D.ptr(jj),
            D.ptr(xy), total, D.ptr(jj_sorted), D.ptr(xy_sorted), D.stream(),
        )
        return ii, jj_sorted, xy_sorted

    def _locate_edges(self, edge_coords, kind: int):
        """count (recording hits) -> scan -> write, then the per-segment sort by t (device
        tensors)."""
        n = int(edge_coords.shape[0])
        st = D.stream()
        order, srt = self._order(2, edge_coords)
        split = self._split(2 if kind == 1 else 3, edge_coords, order)
        if split is not None:
            subs, part_offsets, m = split
            counts = D.empty(max(m, 1), torch.int64)
            hits, slots = self._hit_record(m, self.EDGE_HIT_SLOTS, self.EDGE_HIT_BUDGET)
            hit_xy, xy_slots = (None, 0) if hits is None else self._xy_record(m, slots, self.EDGE_XY_SLOTS,
                                                                              self.EDGE_XY_BUDGET)
            _native.call("ct_locate_edges_count_sub", self.tree_ref, D.ptr(edge_coords), D.ptr(subs), m, kind,
                         D.ptr(counts), D.ptr(hits), slots, D.ptr(hit_xy), xy_slots, st)
            # staging rows in processing order, then the per-segment sort into the output rows
            sub_offsets, total = D.exclusive_scan(counts[:m])
            per_query = D.empty(n, torch.int64)
            query_start = D.empty(n, torch.int64)
            _native.call("ct_subquery_spans", D.ptr(order), n, D.ptr(part_offsets), D.ptr(sub_offsets),
                         D.ptr(per_query), D.ptr(query_start), st)
            offsets, _ = D.exclusive_scan(per_query, total=False)
            rows_query = D.empty(total, torch.int32)
            rows_cell = D.empty(total, torch.int32)
            rows_xy = D.empty((total, 4), torch.float64)
            rows_t = D.empty(total, torch.float64)
            _native.call("ct_locate_edges_write_sub", self.tree_ref, D.ptr(edge_coords), D.ptr(subs), m, kind,
                         D.ptr(sub_offsets), D.ptr(hits), slots, D.ptr(hit_xy), xy_slots, D.ptr(rows_query),
                         D.ptr(rows_cell), D.ptr(rows_xy), D.ptr(rows_t), st)
            ii = D.empty(total, torch.int64)
            jj = D.empty(total, torch.int64)
            xy = D.empty((total, 2, 2), torch.float64)
            _native.call("ct_sort_staged_rows", D.ptr(offsets), D.ptr(query_start), D.ptr(rows_query),
                         D.ptr(rows_cell), D.ptr(rows_xy), D.ptr(rows_t), total, D.ptr(ii), D.ptr(jj), D.ptr(xy), st)
            return ii, jj, xy
        counts = D.empty(n, torch.int64)
        hits, slots = self._hit_record(n, self.EDGE_HIT_SLOTS, self.EDGE_HIT_BUDGET)
        _native.call(
            "ct_locate_edges_count_rec", self.tree_ref, D.ptr(edge_coords), n, kind, D.ptr(order), D.ptr(srt),
            D.ptr(counts), D.ptr(hits), slots, st
        )
        offsets, total = D.exclusive_scan(counts)
        ii = D.empty(total, torch.int64)
        jj = D.empty(total, torch.int64)
        xy = D.empty((total, 2, 2), torch.float64)
        _native.call(
            "ct_locate_edges_write_rec", self.tree_ref, D.ptr(edge_coords), n, kind, D.ptr(order), D.ptr(srt),
            D.ptr(offsets), D.ptr(hits), slots, D.ptr(ii), D.ptr(jj), D.ptr(xy), st,
        )
        return self._sort_edges(edge_coords, offsets, ii, jj, xy, total)
