"""ctypes binding of libcelltree_hip.so (the C ABI declared in include/celltree_hip.h).

The library is the only compute path of this package: there is no CPU fallback.  When the
shared object is missing, or no ROCm GPU is visible, every entry point raises.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcelltree_hip.so")

CT_OK = 0
CT_ERR_ARG = -1
CT_ERR_HIP = -2
CT_ERR_DEGENERATE = -3
CT_ERR_DEPTH = -4
CT_ERR_WORKSPACE = -5
CT_MAX_N_VERTEX = 32
CT_MAX_DEPTH = 254
CT_NOALLOC_MAX_POINTS = 1 << 21


class CtNode(ctypes.Structure):
    """Device node record (32 bytes), ``ct_node`` in the C header."""

    _fields_ = [
        ("Lmax", ctypes.c_double),
        ("Rmin", ctypes.c_double),
        ("child", ctypes.c_int32),
        ("ptr", ctypes.c_int32),
        ("size", ctypes.c_int32),
        ("dim", ctypes.c_int32),
    ]


class CtTree(ctypes.Structure):
    """``ct_tree``: device pointers of a built tree plus its sizes and bbox."""

    _fields_ = [
        ("vertices", ctypes.c_void_p),
        ("elements", ctypes.c_void_p),
        ("bb_coords", ctypes.c_void_p),
        ("nodes", ctypes.c_void_p),
        ("bb_indices", ctypes.c_void_p),
        ("n_vertices", ctypes.c_int64),
        ("n_elements", ctypes.c_int64),
        ("n_nodes", ctypes.c_int64),
        ("n_max_vert", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("bbox", ctypes.c_double * 4),
        ("leaf_coords", ctypes.c_void_p),
        ("node_boxes", ctypes.c_void_p),
        ("start_grid", ctypes.c_void_p),
        ("start_grid_n", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("leaf_bb", ctypes.c_void_p),
        ("parents", ctypes.c_void_p),
    ]


class CtBatchTile(ctypes.Structure):
    """``ct_batch_tile`` (16 bytes)."""

    _fields_ = [("begin", ctypes.c_int64), ("count", ctypes.c_int32), ("tree", ctypes.c_int32)]


class CtBatch(ctypes.Structure):
    """``ct_batch``: host description of a batch of trees (device arrays + sizes)."""

    _fields_ = [
        ("trees", ctypes.c_void_p),
        ("tolerances", ctypes.c_void_p),
        ("query_offsets", ctypes.c_void_p),
        ("tiles", ctypes.c_void_p),
        ("queue_starts", ctypes.c_void_p),
        ("counters", ctypes.c_void_p),
        ("n_tiles", ctypes.c_int64),
        ("n_trees", ctypes.c_int32),
        ("n_queues", ctypes.c_int32),
        ("max_n_vert", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("stack", ctypes.c_int32),
        ("n_workgroups", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class CtRegridTopology(ctypes.Structure):
    """``ct_regrid_topology``: the fixed pair structure of a regrid (device arrays)."""

    _fields_ = [
        ("tgt_faces", ctypes.c_void_p),
        ("src_faces", ctypes.c_void_p),
        ("tt", ctypes.c_void_p),
        ("ss", ctypes.c_void_p),
        ("tgt_offsets", ctypes.c_void_p),
        ("src_offsets", ctypes.c_void_p),
        ("perm_by_src", ctypes.c_void_p),
        ("tgt_vertex_offsets", ctypes.c_void_p),
        ("tgt_vertex_corners", ctypes.c_void_p),
        ("src_vertex_offsets", ctypes.c_void_p),
        ("src_vertex_corners", ctypes.c_void_p),
        ("n_pairs", ctypes.c_int64),
        ("n_tgt", ctypes.c_int64),
        ("n_src", ctypes.c_int64),
        ("n_tgt_vertices", ctypes.c_int64),
        ("n_src_vertices", ctypes.c_int64),
        ("n_max_t", ctypes.c_int32),
        ("n_max_s", ctypes.c_int32),
        ("rest_targets", ctypes.c_void_p),
        ("n_rest_targets", ctypes.c_int64),
    ]


assert ctypes.sizeof(CtNode) == 32
assert ctypes.sizeof(CtTree) == 152
assert ctypes.sizeof(CtBatchTile) == 16

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F64 = ctypes.c_double
_SZ = ctypes.c_size_t
_TP = ctypes.POINTER(CtTree)
_BP = ctypes.POINTER(CtBatch)
_RP = ctypes.POINTER(CtRegridTopology)

# name -> (restype, argtypes); mirrors include/celltree_hip.h one to one.
SIGNATURES = {
    "ct_last_error": (ctypes.c_char_p, []),
    "ct_version": (ctypes.c_int, []),
    "ct_counter_clockwise": (ctypes.c_int, [_P, _P, _I64, _I32, _P]),
    "ct_build_face_bboxes": (ctypes.c_int, [_P, _I64, _I32, _P, _P, _P]),
    "ct_build_edge_bboxes": (ctypes.c_int, [_P, _I64, _P, _F64, _P, _P]),
    "ct_bbox_stats": (ctypes.c_int, [_P, _I64, _P, _P, _P]),
    "ct_vertex_extent": (ctypes.c_int, [_P, _I64, _P, _P]),
    "ct_max_nodes": (_I64, [_I64]),
    "ct_build_workspace_bytes": (_SZ, [_I64, _I32]),
    "ct_build": (
        ctypes.c_int,
        [_P, _I64, _I32, _I32, _P, _SZ, _P, _P, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _P],
    ),
    "ct_collect_node_bounds": (ctypes.c_int, [_P, _I64, ctypes.POINTER(_F64), _P, _P, _P]),
    "ct_validate_node_bounds": (ctypes.c_int, [_TP, _P, _P, _P]),
    "ct_locate_points": (ctypes.c_int, [_TP, _P, _I64, _F64, _P, _P]),
    "ct_pack_leaf_coords": (ctypes.c_int, [_TP, _P, _P]),
    "ct_node_boxes": (ctypes.c_int, [_TP, _P, _P, _P]),
    "ct_start_grid": (ctypes.c_int, [_TP, _I32, _P, _P]),
    "ct_pack_leaf_bboxes": (ctypes.c_int, [_TP, _P, _P]),
    "ct_locate_workspace_bytes": (_SZ, [_I64]),
    "ct_locate_points_ordered": (ctypes.c_int, [_TP, _P, _I64, _F64, _P, _SZ, _P, _P]),
    "ct_locate_points_in_order": (ctypes.c_int, [_TP, _P, _I64, _P, _P, _F64, _P, _P]),
    "ct_check_tree": (ctypes.c_int, [_P, _I64, _P, _P, _I64, _I32, _I64, _I32, _P]),
    "ct_node_parents": (ctypes.c_int, [_TP, _P, _P]),
    "ct_barycentric_wachspress_backward": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _P, _F64, _P, _P, _P, _P]),
    "ct_split_queries": (ctypes.c_int, [_TP, _I32, _P, _I64, _P, _I32, _I32, _I32, _P, _P, _P, _P, _P]),
    "ct_subquery_offsets": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P]),
    "ct_locate_boxes_count_sub": (ctypes.c_int, [_TP, _P, _P, _I64, _P, _P, _I32, _P]),
    "ct_locate_boxes_write_sub": (ctypes.c_int, [_TP, _P, _P, _I64, _P, _P, _P, _I32, _P, _P, _P]),
    "ct_locate_edges_count_sub": (ctypes.c_int, [_TP, _P, _P, _I64, _I32, _P, _P, _I32, _P, _I32, _P]),
    "ct_locate_edges_write_sub": (ctypes.c_int, [_TP, _P, _P, _I64, _I32, _P, _P, _I32, _P, _I32, _P, _P, _P, _P,
                                                 _P]),
    "ct_subquery_spans": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P]),
    "ct_sort_staged_rows": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "ct_locate_points_ordered_noalloc": (ctypes.c_int, [_TP, _P, _I64, _F64, _P, _P]),
    "ct_locate_points_on_edge": (ctypes.c_int, [_TP, _P, _I64, _F64, _P, _P]),
    "ct_order_workspace_bytes": (_SZ, [_I64]),
    "ct_query_order": (ctypes.c_int, [_I32, _P, _I64, ctypes.POINTER(_F64), _P, _SZ, _P, _P, _P]),
    "ct_locate_boxes_count": (ctypes.c_int, [_TP, _P, _I64, _P, _P, _P, _P]),
    "ct_locate_boxes_write": (ctypes.c_int, [_TP, _P, _I64, _P, _P, _P, _P, _P, _P]),
    "ct_locate_edges_count": (ctypes.c_int, [_TP, _P, _I64, _I32, _P, _P, _P, _P]),
    "ct_locate_edges_write": (ctypes.c_int, [_TP, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P]),
    "ct_locate_boxes_count_rec": (ctypes.c_int, [_TP, _P, _I64, _P, _P, _P, _P, _I32, _P]),
    "ct_locate_boxes_write_rec": (ctypes.c_int, [_TP, _P, _I64, _P, _P, _P, _P, _I32, _P, _P, _P]),
    "ct_locate_edges_count_rec": (ctypes.c_int, [_TP, _P, _I64, _I32, _P, _P, _P, _P, _I32, _P]),
    "ct_locate_edges_write_rec": (ctypes.c_int, [_TP, _P, _I64, _I32, _P, _P, _P, _P, _I32, _P, _P, _P, _P]),
    "ct_sort_intersections_by_edge": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _P, _P, _P]),
    "ct_scan_workspace_bytes": (_SZ, [_I64]),
    "ct_exclusive_scan": (ctypes.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "ct_polygons_intersect": (ctypes.c_int, [_P, _P, _P, _I32, _P, _I32, _P, _P, _I64, _P, _P]),
    "ct_area_of_intersection": (ctypes.c_int, [_P, _P, _P, _I32, _P, _I32, _P, _P, _I64, _P, _P]),
    "ct_box_area_of_intersection": (ctypes.c_int, [_P, _P, _P, _I32, _P, _P, _I64, _P, _P]),
    "ct_face_pairs": (ctypes.c_int, [_P, _P, _I32, _P, _P, _I32, _P, _P, _I64, _I32, _P, _P, _P]),
    "ct_box_pairs": (ctypes.c_int, [_P, _P, _P, _I32, _P, _P, _I64, _P, _P, _P]),
    "ct_compact_pairs": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "ct_barycentric_triangle_weights": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _P, _F64, _P, _P]),
    "ct_barycentric_wachspress_weights": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _P, _F64, _P, _P]),
    "ct_compute_barycentric_weights": (ctypes.c_int, [_TP, _P, _I64, _F64, _P, _P, _P]),
    "ct_regrid_pair_areas": (ctypes.c_int, [_P, _P, _I32, _P, _P, _I32, _P, _P, _I64, _P, _P]),
    "ct_polygon_areas": (ctypes.c_int, [_P, _P, _I32, _I64, _P, _P]),
    "ct_regrid_apply": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _I64, _P, _P]),
    "ct_regrid_apply_transpose": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _I64, _P, _P]),
    "ct_barycentric_triangle_backward": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _P, _P, _P, _P, _P]),
    "ct_regrid_rest_targets": (ctypes.c_int, [_RP, _P, _P, _P]),
    "ct_regrid_backward_workspace_bytes": (_SZ, [_RP]),
    "ct_regrid_vertex_backward": (ctypes.c_int, [_RP, _P, _P, _P, _P, _P, _P, _I32, _P, _SZ, _P, _P, _P]),
    "ct_batch_locate_points": (ctypes.c_int, [_BP, _P, _P, _P]),
    "ct_batch_edges_count": (ctypes.c_int, [_BP, _P, _P, _P]),
    "ct_batch_edges_write": (ctypes.c_int, [_BP, _P, _P, _I64, _P, _P, _P, _P]),
    "ct_batch_edges_count_rec": (ctypes.c_int, [_BP, _P, _I64, _P, _P, _I32, _P]),
    "ct_batch_edges_write_rec": (ctypes.c_int, [_BP, _P, _I64, _P, _P, _I32, _I64, _P, _P, _P, _P]),
    "ct_batch_sort_intersections": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _I64, _I32, _P, _P, _P]),
}

_lib = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    """Load the shared library (no GPU work happens at load time)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is not built; build it with "
                        "`python numba_celltree_amd/csrc/build.py`"
                    )
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def last_error() -> str:
    msg = lib().ct_last_error()
    return msg.decode() if msg else ""


def check(rc: int, name: str) -> None:
    """Raise the Python exception matching a CT_ERR_* status."""
    if rc == CT_OK:
        return
    msg = f"{name}: {last_error()}"
    if rc == CT_ERR_ARG:
        raise ValueError(msg)
    if rc == CT_ERR_DEGENERATE:
        raise IndexError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
