"""CPU oracle for the cell-tree hot path: ctypes wrapper around ``refcpu.cpp``.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module; the product package never
does.  The classes below mirror numba_celltree v0.4.2's public API so that tests
read like the reference's own:

* ``OracleCellTree2d``      -- numba_celltree/celltree.py:37-351
* ``OracleEdgeCellTree2d``  -- numba_celltree/edge_celltree.py:25-146

Input casting restates numba_celltree/cast.py:14-54, the tree record restates
constants.py:82-106 (``NodeDType`` is the reference's packed 41-byte layout) and
``sort_intersections_by_edge`` restates geometry_utils.py:564-574 with the same
``np.lexsort`` call.
"""

import ctypes
import os
from typing import NamedTuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librefcpu.so")

IntDType = np.intp
FloatDType = np.float64
FILL_VALUE = -1
MAX_N_VERTEX = 32
MIN_TOLERANCE = 1e-15
TOLERANCE_FACTOR = 1e-12

# constants.py:92-106 -- packed record, 41 bytes.
NodeDType = np.dtype(
    [
        ("child", IntDType),
        ("Lmax", FloatDType),
        ("Rmin", FloatDType),
        ("ptr", IntDType),
        ("size", IntDType),
        ("dim", bool),
    ]
)
assert NodeDType.itemsize == 41


class CellTreeData(NamedTuple):  # constants.py:82-89
    elements: np.ndarray
    vertices: np.ndarray
    nodes: np.ndarray
    bb_indices: np.ndarray
    bb_coords: np.ndarray
    bbox: np.ndarray
    cells_per_leaf: int


class _Tree(ctypes.Structure):
    _fields_ = [
        ("elements", ctypes.c_void_p),
        ("n_elements", ctypes.c_int64),
        ("n_max_vert", ctypes.c_int64),
        ("vertices", ctypes.c_void_p),
        ("nodes", ctypes.c_void_p),
        ("n_nodes", ctypes.c_int64),
        ("bb_indices", ctypes.c_void_p),
        ("bb_coords", ctypes.c_void_p),
        ("bbox", ctypes.c_double * 4),
    ]


class _Pairs(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("ii", ctypes.POINTER(ctypes.c_int64)),
        ("jj", ctypes.POINTER(ctypes.c_int64)),
        ("xy", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None

# CPU features each ISA level of build_oracle.py needs (the x86-64 psABI levels).
ISA_FEATURES = {
    "x86-64-v4": ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"),
    "x86-64-v3": ("avx2", "bmi1", "bmi2", "fma", "movbe", "abm", "f16c"),
}


def host_cpu():
    """(model name, flags) of this host's CPU from /proc/cpuinfo ("unknown", set() elsewhere)."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return model, flags


def native_isa_level() -> str:
    """The highest ISA level of build_oracle.py this host's CPU executes."""
    _, flags = host_cpu()
    for level in ("x86-64-v4", "x86-64-v3"):
        if set(ISA_FEATURES[level]) <= flags:
            return level
    return "x86-64-v2"


def use_isa_level(level: str) -> str:
    """Load the oracle compiled for ``level`` from now on (same results at every level);
    returns the library path."""
    global _lib, _LIB_PATH
    from oracle import build_oracle

    path = build_oracle.library_path(level)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (python oracle/build_oracle.py)")
    if path != _LIB_PATH:
        _LIB_PATH = path
        _lib = None
    return path


def lib():
    """Load the oracle library (``librefcpu.so`` unless use_isa_level chose another),
    building it first when it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            from oracle import build_oracle

            build_oracle.build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, f64 = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        sigs = {
            "ref_num_threads": (i32, []),
            "ref_pessimistic_n_nodes": (i64, [i64]),
            "ref_initialize": (i64, [P, i64, i32, i32, P, P]),
            "ref_counter_clockwise": (None, [P, P, i64, i32]),
            "ref_build_face_bboxes": (None, [P, i64, i32, P, P]),
            "ref_build_edge_bboxes": (None, [P, i64, P, f64, P]),
            "ref_locate_points": (None, [P, P, i64, f64, P]),
            "ref_locate_points_counted": (None, [P, P, i64, f64, P, P]),
            "ref_locate_points_on_edge": (None, [P, P, i64, f64, P]),
            "ref_locate_boxes": (None, [P, P, i64, P]),
            "ref_locate_edges": (None, [P, P, i64, i32, P]),
            "ref_pairs_free": (None, [P]),
            "ref_polygons_intersect": (None, [P, P, P, i32, P, i32, P, P, i64, P]),
            "ref_area_of_intersection": (None, [P, P, P, i32, P, i32, P, P, i64, P]),
            "ref_face_pair_flops": (None, [P, P, P, i32, P, i32, P, P, i64, P, P, P, P]),
            "ref_pair_area_gradient": (None, [P, P, P, i32, P, i32, P, P, i64, P, P, P, P]),
            "ref_box_area_of_intersection": (None, [P, P, P, i32, P, P, i64, P]),
            "ref_barycentric_triangle_weights": (None, [P, P, i64, P, i32, P, f64, P]),
            "ref_barycentric_wachspress_weights": (None, [P, P, i64, P, i32, P, f64, P]),
            "ref_collect_node_bounds": (None, [P, P]),
            "ref_validate_node_bounds": (None, [P, P, P]),
            "ref_polygon_area": (f64, [P, i32]),
            "ref_polygon_polygon_clip_area": (f64, [P, i32, P, i32]),
            "ref_sh_intersection": (i32, [P, P, P, P, P]),
            "ref_separating_axes": (i32, [P, i32, P, i32]),
            "ref_point_in_polygon_or_on_edge": (i32, [P, P, i32, f64]),
            "ref_point_on_edge": (i32, [P, P, f64]),
            "ref_cohen_sutherland": (i32, [P, P, P, P]),
            "ref_cyrus_beck": (i32, [P, P, P, i32, f64, P]),
            "ref_lines_intersect": (i32, [P, P, P, P, P]),
            "ref_bary_triangle": (None, [P, P, P]),
            "ref_wachspress": (None, [P, i32, P, P, i32, f64]),
            "ref_interp_edge_case": (None, [P, P, P, P, i32, i32, i32]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


# --------------------------------------------------------------------------- cast.py
def cast_vertices(vertices, copy: bool = False) -> np.ndarray:  # cast.py:14-21
    if isinstance(vertices, np.ndarray):
        vertices = vertices.astype(FloatDType, copy=copy)
    else:
        vertices = np.ascontiguousarray(vertices, dtype=FloatDType)
    if vertices.ndim != 2 or vertices.shape[1] != 2:
        raise ValueError("vertices must have shape (n_points, 2)")
    return np.ascontiguousarray(vertices)


def cast_faces(faces, fill_value: int) -> np.ndarray:  # cast.py:24-40
    if isinstance(faces, np.ndarray):
        faces = faces.astype(IntDType, copy=True)
    else:
        faces = np.ascontiguousarray(faces, dtype=IntDType)
    if faces.ndim != 2:
        raise ValueError("faces must have shape (n_face, n_max_vert)")
    if faces.shape[1] > MAX_N_VERTEX:
        raise ValueError(f"faces contains up to {faces.shape[1]} vertices for a single face.")
    if fill_value != FILL_VALUE:
        faces[faces == fill_value] = FILL_VALUE
    return np.ascontiguousarray(faces)


def cast_bboxes(bbox_coords) -> np.ndarray:  # cast.py:43-47
    bbox_coords = np.ascontiguousarray(bbox_coords, dtype=FloatDType)
    if bbox_coords.ndim != 2 or bbox_coords.shape[1] != 4:
        raise ValueError("bbox_coords must have shape (n_box, 4)")
    return bbox_coords


def cast_edges(edges) -> np.ndarray:  # cast.py:50-54
    edges = np.ascontiguousarray(edges, dtype=FloatDType)
    if edges.ndim != 3 or edges.shape[1] != 2 or edges.shape[2] != 2:
        raise ValueError("edges must have shape (n_edge, 2, 2)")
    return edges


# ------------------------------------------------------------- geometry / algorithms
def counter_clockwise(vertices: np.ndarray, faces: np.ndarray) -> None:
    """In place, geometry_utils.py:541-561. ``faces`` must be a C-contiguous int64 array."""
    assert faces.dtype == np.int64 and faces.flags.c_contiguous
    lib().ref_counter_clockwise(_p(_f64(vertices)), _p(faces), len(faces), faces.shape[1])


def build_face_bboxes(faces: np.ndarray, vertices: np.ndarray) -> np.ndarray:
    faces = _i64(faces)
    out = np.empty((len(faces), 4), dtype=np.float64)
    lib().ref_build_face_bboxes(_p(faces), len(faces), faces.shape[1], _p(_f64(vertices)), _p(out))
    return out


def build_edge_bboxes(edges: np.ndarray, vertices: np.ndarray, tolerance: float) -> np.ndarray:
    edges = _i64(edges)
    out = np.empty((len(edges), 4), dtype=np.float64)
    lib().ref_build_edge_bboxes(_p(edges), len(edges), _p(_f64(vertices)), tolerance, _p(out))
    return out


def initialize(elements, bb_coords, n_buckets: int = 4, cells_per_leaf: int = 2):
    """creation.py:384-413 -> (nodes[NodeDType], bb_indices[intp])."""
    bb_coords = _f64(bb_coords)
    n = len(elements)
    cap = lib().ref_pessimistic_n_nodes(n)
    nodes = np.empty(cap, dtype=NodeDType)
    bb_indices = np.empty(n, dtype=np.int64)
    n_nodes = lib().ref_initialize(_p(bb_coords), n, n_buckets, cells_per_leaf, _p(nodes), _p(bb_indices))
    if n_nodes < 0:
        raise IndexError("degenerate bucketing: a centroid falls outside every bucket")
    return nodes[:n_nodes].copy(), bb_indices


def bbox_tree(bb_coords: np.ndarray) -> np.ndarray:  # celltree_base.py:20-25
    return np.array(
        [bb_coords[:, 0].min(), bb_coords[:, 1].max(), bb_coords[:, 2].min(), bb_coords[:, 3].max()],
        dtype=np.float64,
    )


def bbox_distances(bb_coords: np.ndarray) -> np.ndarray:  # celltree_base.py:28-48
    d = np.empty((bb_coords.shape[0], 3), dtype=np.float64)
    d[:, 0] = bb_coords[:, 1] - bb_coords[:, 0]
    d[:, 1] = bb_coords[:, 3] - bb_coords[:, 2]
    d[:, 2] = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
    return d


def default_tolerance(bb_diagonal: np.ndarray) -> float:  # celltree_base.py:51-52
    return max(MIN_TOLERANCE, TOLERANCE_FACTOR * max(bb_diagonal))


def sort_intersections_by_edge(ii, jj, xy, edge_coords):  # geometry_utils.py:564-574
    a = edge_coords[ii, 0]
    b = edge_coords[ii, 1]
    c = xy[:, 0]
    t = np.sum((c - a) * (b - a), axis=1)
    order = np.lexsort((t, ii))
    return ii[order], jj[order], xy[order]


def polygons_intersect(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    fa, fb = _i64(faces_a), _i64(faces_b)
    ia, ib = _i64(indices_a), _i64(indices_b)
    out = np.empty(len(ia), dtype=np.uint8)
    lib().ref_polygons_intersect(
        _p(_f64(vertices_a)), _p(_f64(vertices_b)), _p(fa), fa.shape[1], _p(fb), fb.shape[1],
        _p(ia), _p(ib), len(ia), _p(out),
    )
    return out.astype(bool)


def area_of_intersection(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    fa, fb = _i64(faces_a), _i64(faces_b)
    ia, ib = _i64(indices_a), _i64(indices_b)
    out = np.empty(len(ia), dtype=np.float64)
    lib().ref_area_of_intersection(
        _p(_f64(vertices_a)), _p(_f64(vertices_b)), _p(fa), fa.shape[1], _p(fb), fb.shape[1],
        _p(ia), _p(ib), len(ia), _p(out),
    )
    return out


def face_pair_flops(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b):
    """Per candidate pair of intersect_faces' pair stage: (sat_ops, passed, clip_ops, area).

    sat_ops counts the FP64 adds, subs, muls and divs polygons_intersect executes
    (separating_axis.py:58-75, both directions, short-circuit), clip_ops those of
    area_of_intersection (sutherland_hodgman.py:84-168 plus polygon_area) on pairs that
    passed (0 otherwise); area is that area (0 where not computed)."""
    fa, fb = _i64(faces_a), _i64(faces_b)
    ia, ib = _i64(indices_a), _i64(indices_b)
    n = len(ia)
    sat = np.empty(n, dtype=np.int64)
    passed = np.empty(n, dtype=np.uint8)
    clip = np.empty(n, dtype=np.int64)
    area = np.empty(n, dtype=np.float64)
    lib().ref_face_pair_flops(
        _p(_f64(vertices_a)), _p(_f64(vertices_b)), _p(fa), fa.shape[1], _p(fb), fb.shape[1],
        _p(ia), _p(ib), n, _p(sat), _p(passed), _p(clip), _p(area),
    )
    return sat, passed.astype(bool), clip, area


def pair_area_gradient(vertices_a, vertices_b, faces_a, faces_b, indices_a, indices_b, g_area):
    """Per pair (subject faces_a[indices_a], clipper faces_b[indices_b], convex, CCW): g_area
    times the derivative of area_of_intersection with respect to the subject's and the
    clipper's vertices, (ga (n, na_max, 2), gb (n, nb_max, 2)) by vertex slot, and the FP64
    operations of each (ref_pair_area_gradient; no reference counterpart)."""
    fa, fb = _i64(faces_a), _i64(faces_b)
    ia, ib = _i64(indices_a), _i64(indices_b)
    n = len(ia)
    g = np.ascontiguousarray(np.broadcast_to(np.asarray(g_area, dtype=np.float64), (n,)))
    ga = np.zeros((n, fa.shape[1], 2))
    gb = np.zeros((n, fb.shape[1], 2))
    ops = np.empty(n, dtype=np.int64)
    lib().ref_pair_area_gradient(
        _p(_f64(vertices_a)), _p(_f64(vertices_b)), _p(fa), fa.shape[1], _p(fb), fb.shape[1],
        _p(ia), _p(ib), n, _p(g), _p(ga), _p(gb), _p(ops),
    )
    return ga, gb, ops


def box_area_of_intersection(bbox_coords, vertices, faces, indices_bbox, indices_face):
    f = _i64(faces)
    ib, iface = _i64(indices_bbox), _i64(indices_face)
    out = np.empty(len(ib), dtype=np.float64)
    lib().ref_box_area_of_intersection(
        _p(_f64(bbox_coords)), _p(_f64(vertices)), _p(f), f.shape[1], _p(ib), _p(iface), len(ib), _p(out)
    )
    return out


def barycentric_triangle_weights(points, face_indices, faces, vertices, tolerance):
    pts, fi, f = _f64(points), _i64(face_indices), _i64(faces)
    out = np.empty((len(pts), 3), dtype=np.float64)
    lib().ref_barycentric_triangle_weights(
        _p(pts), _p(fi), len(pts), _p(f), f.shape[1], _p(_f64(vertices)), tolerance, _p(out)
    )
    return out


def barycentric_wachspress_weights(points, face_indices, faces, vertices, tolerance):
    pts, fi, f = _f64(points), _i64(face_indices), _i64(faces)
    out = np.empty((len(pts), f.shape[1]), dtype=np.float64)
    lib().ref_barycentric_wachspress_weights(
        _p(pts), _p(fi), len(pts), _p(f), f.shape[1], _p(_f64(vertices)), tolerance, _p(out)
    )
    return out


# Single primitives (algorithm unit tests).
def polygon_area(poly) -> float:
    poly = _f64(poly)
    return lib().ref_polygon_area(_p(poly), len(poly))


def polygon_polygon_clip_area(subject, clipper) -> float:
    s, c = _f64(subject), _f64(clipper)
    return lib().ref_polygon_polygon_clip_area(_p(s), len(s), _p(c), len(c))


def sh_intersection(a, V, r, N):
    out = np.empty(2)
    ok = lib().ref_sh_intersection(_p(_f64(a)), _p(_f64(V)), _p(_f64(r)), _p(_f64(N)), _p(out))
    return bool(ok), out


def separating_axes(a, b) -> bool:
    a, b = _f64(a), _f64(b)
    return bool(lib().ref_separating_axes(_p(a), len(a), _p(b), len(b)))


def point_in_polygon_or_on_edge(p, poly, tolerance) -> bool:
    poly = _f64(poly)
    return bool(lib().ref_point_in_polygon_or_on_edge(_p(_f64(p)), _p(poly), len(poly), tolerance))


def point_on_edge(p, edge, tolerance) -> bool:
    return bool(lib().ref_point_on_edge(_p(_f64(p)), _p(_f64(edge)), tolerance))


def cohen_sutherland_line_box_clip(a, b, box):
    cd = np.empty(4)
    ok = lib().ref_cohen_sutherland(_p(_f64(a)), _p(_f64(b)), _p(_f64(box)), _p(cd))
    return bool(ok), cd[:2].copy(), cd[2:].copy()


def cyrus_beck_line_polygon_clip(a, b, poly, tolerance):
    poly = _f64(poly)
    cd = np.empty(4)
    ok = lib().ref_cyrus_beck(_p(_f64(a)), _p(_f64(b)), _p(poly), len(poly), tolerance, _p(cd))
    return bool(ok), cd[:2].copy(), cd[2:].copy()


def lines_intersect(a, b, p, q):
    xy = np.empty(2)
    ok = lib().ref_lines_intersect(_p(_f64(a)), _p(_f64(b)), _p(_f64(p)), _p(_f64(q)), _p(xy))
    return bool(ok), xy[0], xy[1]


def triangle_weights(triangle, p):
    w = np.empty(3)
    lib().ref_bary_triangle(_p(_f64(triangle)), _p(_f64(p)), _p(w))
    return w


def wachspress_weights(polygon, p, n_weights, tolerance):
    poly = _f64(polygon)
    w = np.zeros(n_weights)
    lib().ref_wachspress(_p(poly), len(poly), _p(_f64(p)), _p(w), n_weights, tolerance)
    return w


def interp_edge_case(a, U, p, n_weights, i, j):
    w = np.zeros(n_weights)
    lib().ref_interp_edge_case(_p(_f64(a)), _p(_f64(U)), _p(_f64(p)), _p(w), n_weights, i, j)
    return w


# --------------------------------------------------------------------------- trees
class _OracleBase:
    def _ctree(self) -> _Tree:
        t = _Tree()
        t.elements = self.celltree_data.elements.ctypes.data
        t.n_elements = len(self.celltree_data.elements)
        t.n_max_vert = self.celltree_data.elements.shape[1]
        t.vertices = self.vertices.ctypes.data
        t.nodes = self.nodes.ctypes.data
        t.n_nodes = len(self.nodes)
        t.bb_indices = self.bb_indices.ctypes.data
        t.bb_coords = self.bb_coords.ctypes.data
        for k in range(4):
            t.bbox[k] = float(self.bbox[k])
        self._keep = t
        return t

    @property
    def node_bounds(self) -> np.ndarray:  # celltree_base.py:62-65, query.py:568-622
        out = np.empty((len(self.nodes), 4), dtype=np.float64)
        lib().ref_collect_node_bounds(ctypes.byref(self._ctree()), _p(out))
        return out

    def validate_node_bounds(self) -> np.ndarray:  # celltree_base.py:67-81, query.py:625-664
        nb = self.node_bounds
        out = np.empty(len(self.nodes), dtype=np.uint8)
        lib().ref_validate_node_bounds(ctypes.byref(self._ctree()), _p(nb), _p(out))
        return out.astype(bool)

    def to_dict_of_lists(self):  # celltree_base.py:83-114
        d = {}
        for parent_index, node in enumerate(self.nodes):
            left = node["child"]
            d[parent_index] = [] if left == -1 else [left, left + 1]
        return d

    @staticmethod
    def _pairs(p: _Pairs, with_xy: bool):
        n = p.n
        ii = np.ctypeslib.as_array(p.ii, shape=(max(n, 1),))[:n].copy()
        jj = np.ctypeslib.as_array(p.jj, shape=(max(n, 1),))[:n].copy()
        xy = None
        if with_xy:
            xy = np.ctypeslib.as_array(p.xy, shape=(max(n, 1) * 4,))[: 4 * n].reshape(n, 2, 2).copy()
        lib().ref_pairs_free(ctypes.byref(p))
        return ii, jj, xy

    def _locate_points(self, points, tolerance, fn):
        if tolerance is None:
            tolerance = default_tolerance(self.bb_distances[:, 2])
        points = cast_vertices(points)
        out = np.empty(len(points), dtype=np.int64)
        fn(ctypes.byref(self._ctree()), _p(points), len(points), tolerance, _p(out))
        return out

    def _locate_edges(self, edge_coords, edge_face: int):
        p = _Pairs()
        lib().ref_locate_edges(ctypes.byref(self._ctree()), _p(edge_coords), len(edge_coords), edge_face, ctypes.byref(p))
        return self._pairs(p, True)


class OracleCellTree2d(_OracleBase):
    """numba_celltree/celltree.py:37-351 on the C++ oracle."""

    def __init__(self, vertices, faces, fill_value, n_buckets: int = 4, cells_per_leaf: int = 2):
        if n_buckets < 2:
            raise ValueError("n_buckets must be >= 2")
        if cells_per_leaf < 1:
            raise ValueError("cells_per_leaf must be >= 1")
        vertices = cast_vertices(vertices, copy=True)
        faces = cast_faces(faces, fill_value)
        counter_clockwise(vertices, faces)
        bb_coords = build_face_bboxes(faces, vertices)
        nodes, bb_indices = initialize(faces, bb_coords, n_buckets, cells_per_leaf)
        self.vertices = vertices
        self.faces = faces
        self.n_buckets = n_buckets
        self.cells_per_leaf = cells_per_leaf
        self.nodes = nodes
        self.bb_indices = bb_indices
        self.bb_coords = bb_coords
        self.bbox = bbox_tree(bb_coords)
        self.bb_distances = bbox_distances(bb_coords)
        self.celltree_data = CellTreeData(
            self.faces, self.vertices, self.nodes, self.bb_indices, self.bb_coords, self.bbox,
            self.cells_per_leaf,
        )

    def locate_points(self, points, tolerance=None):
        return self._locate_points(points, tolerance, lib().ref_locate_points)

    def locate_boxes(self, bbox_coords):
        bbox_coords = cast_bboxes(bbox_coords)
        p = _Pairs()
        lib().ref_locate_boxes(ctypes.byref(self._ctree()), _p(bbox_coords), len(bbox_coords), ctypes.byref(p))
        ii, jj, _ = self._pairs(p, False)
        return ii, jj

    def intersect_boxes(self, bbox_coords):
        bbox_coords = cast_bboxes(bbox_coords)
        i, j = self.locate_boxes(bbox_coords)
        area = box_area_of_intersection(bbox_coords, self.vertices, self.faces, i, j)
        actual = area > 0
        return i[actual], j[actual], area[actual]

    def locate_faces(self, vertices, faces):
        vertices = _f64(vertices)
        counter_clockwise(vertices, faces)
        bbox_coords = build_face_bboxes(faces, vertices)
        si, sj = self.locate_boxes(bbox_coords)
        intersects = polygons_intersect(vertices, self.vertices, faces, self.faces, si, sj)
        return si[intersects], sj[intersects]

    def intersect_faces(self, vertices, faces, fill_value):
        vertices = cast_vertices(vertices)
        faces = cast_faces(faces, fill_value)
        i, j = self.locate_faces(vertices, faces)
        area = area_of_intersection(vertices, self.vertices, faces, self.faces, i, j)
        actual = area > 0
        return i[actual], j[actual], area[actual]

    def intersect_edges(self, edge_coords):
        edge_coords = cast_edges(edge_coords)
        ii, jj, xy = self._locate_edges(edge_coords, 1)
        return sort_intersections_by_edge(ii, jj, xy, edge_coords)

    def compute_barycentric_weights(self, points, tolerance=None):
        if tolerance is None:
            tolerance = default_tolerance(self.bb_distances[:, 2])
        face_indices = self.locate_points(points, tolerance)
        f = barycentric_wachspress_weights if self.faces.shape[1] > 3 else barycentric_triangle_weights
        return face_indices, f(cast_vertices(points), face_indices, self.faces, self.vertices, tolerance)


class OracleEdgeCellTree2d(_OracleBase):
    """numba_celltree/edge_celltree.py:25-146 on the C++ oracle."""

    def __init__(self, vertices, edges, n_buckets: int = 4, cells_per_leaf: int = 2):
        if n_buckets < 2:
            raise ValueError("n_buckets must be >= 2")
        if cells_per_leaf < 1:
            raise ValueError("cells_per_leaf must be >= 1")
        vertices = cast_vertices(vertices, copy=True)
        x, y = vertices.T
        dx = x.max() - x.min()
        dy = y.max() - y.min()
        global_tolerance = max(MIN_TOLERANCE, TOLERANCE_FACTOR * max(dx, dy))
        edges = _i64(edges)
        bb_coords = build_edge_bboxes(edges, vertices, global_tolerance)
        nodes, bb_indices = initialize(edges, bb_coords, n_buckets, cells_per_leaf)
        self.vertices = vertices
        self.edges = edges
        self.n_buckets = n_buckets
        self.cells_per_leaf = cells_per_leaf
        self.nodes = nodes
        self.bb_indices = bb_indices
        self.bb_coords = bb_coords
        self.bbox = bbox_tree(bb_coords)
        self.bb_distances = bbox_distances(bb_coords)
        self.celltree_data = CellTreeData(
            self.edges, self.vertices, self.nodes, self.bb_indices, self.bb_coords, self.bbox,
            self.cells_per_leaf,
        )

    def locate_points(self, points, tolerance=None):
        return self._locate_points(points, tolerance, lib().ref_locate_points_on_edge)

    def intersect_edges(self, edge_coords):
        edge_coords = cast_edges(edge_coords)
        ii, jj, xy = self._locate_edges(edge_coords, 0)
        ii, jj, xy = sort_intersections_by_edge(ii, jj, xy, edge_coords)
        return ii, jj, np.ascontiguousarray(xy[:, 0])
