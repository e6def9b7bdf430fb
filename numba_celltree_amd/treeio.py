"""Tree export / import: a built CellTree2d or EdgeCellTree2d as plain host arrays or a file.

The reference keeps a tree as plain numpy attributes (celltree.py:80-97,
edge_celltree.py:65-83), which pickle with the object; it has no file format of its own.
Here a tree lives on the GPU, so export copies its arrays to the host and import uploads
them and derives the device-only layouts again, without rebuilding the tree.

Two host forms:

* the reference's attributes (``CellTree2d.to_arrays`` / ``from_arrays``): ``vertices``
  (n_vertices, 2) float64, ``elements`` (n, n_max_vert) int64 (CCW faces, -1 padded; or an
  EdgeCellTree2d's (n, 2) edges), ``nodes`` NodeDType (constants.py:92-106, DFS numbering),
  ``bb_indices`` intp, ``bb_coords`` (n, 4) float64, ``n_buckets``, ``cells_per_leaf``.  A tree
  the reference built imports from these attributes directly;
* the compact file form (``save`` / ``load``): the same tree with int32 elements and
  bb_indices, and nodes as the 32-byte device record ``ct_node`` (NODE_RECORD), so that an
  import is a straight upload; bb_coords are not stored (one kernel recomputes them).

The file is little-endian: 8 magic bytes ``CELLTREE``, the header length (uint64), a JSON
header (format, kind, n_buckets, cells_per_leaf, and per array its name, dtype, shape and
byte offset), then the arrays, each 64-byte aligned.  Reading it parses only JSON and raw
numbers (np.frombuffer of the listed dtypes); nothing in a file is executed.

Everything imported is checked on the host before it reaches a kernel (the traversal
kernels do not bounds-check node links): node links form one tree with children after their
parent, cell ranges lie in [0, n), bb_indices is a permutation, vertex indices are in range
and faces have at least 3 vertices.
"""

import json
from typing import Dict

import numpy as np

from numba_celltree_amd.constants import NodeDType

FORMAT = "numba_celltree_amd.tree/1"
KINDS = ("faces", "edges")
MAGIC = b"CELLTREE"

# ct_node of include/celltree_hip.h: the device node record (32 bytes).
NODE_RECORD = np.dtype([("Lmax", "<f8"), ("Rmin", "<f8"), ("child", "<i4"), ("ptr", "<i4"), ("size", "<i4"),
                        ("dim", "<i4")])
_FILE_DTYPES = {"<f8": np.dtype("<f8"), "<i4": np.dtype("<i4"), "ct_node": NODE_RECORD}
_FILE_ARRAYS = ("vertices", "elements", "nodes", "bb_indices")


def to_node_records(nodes: np.ndarray) -> np.ndarray:
    """NodeDType (or NODE_RECORD) array -> NODE_RECORD array."""
    if nodes.dtype == NODE_RECORD:
        return nodes
    out = np.empty(len(nodes), dtype=NODE_RECORD)
    for name in NODE_RECORD.names:
        out[name] = nodes[name]
    return out


def to_node_dtype(records: np.ndarray) -> np.ndarray:
    """NODE_RECORD array -> the reference's NodeDType array."""
    out = np.empty(len(records), dtype=NodeDType)
    for name in NodeDType.names:
        out[name] = records[name] if name != "dim" else records["dim"] != 0
    return out


def check_tree(nodes: np.ndarray, bb_indices: np.ndarray, elements: np.ndarray, n_vertices: int,
               faces: bool) -> int:
    """The tree's depth, after the library's host-side check (ct_check_tree) that the arrays
    describe a tree the kernels can walk: child links forwards and inside the node array,
    one parent per non-root node, cell ranges in [0, n), bb_indices a permutation, element
    indices in range.  ``nodes`` NODE_RECORD, ``bb_indices`` and ``elements`` int32, all
    C-contiguous host arrays.  Raises ValueError."""
    import ctypes

    from numba_celltree_amd import _native

    depth = ctypes.c_int32(-1)
    rc = _native.lib().ct_check_tree(
        nodes.ctypes.data_as(ctypes.c_void_p), len(nodes), bb_indices.ctypes.data_as(ctypes.c_void_p),
        elements.ctypes.data_as(ctypes.c_void_p), elements.shape[0], elements.shape[1], int(n_vertices),
        int(faces), ctypes.byref(depth))
    if rc != _native.CT_OK:
        raise ValueError(_native.last_error())
    return int(depth.value)


def _to_int32(a: np.ndarray, name: str) -> np.ndarray:
    if a.dtype != np.int32:
        lo, hi = (int(a.min()), int(a.max())) if a.size else (0, 0)
        if lo < -(2**31) or hi >= 2**31:
            raise ValueError(f"{name} out of int32 range")
    return np.ascontiguousarray(a, dtype=np.int32)


def _int_array(x, name: str) -> np.ndarray:
    a = np.asarray(x)
    if a.dtype.kind not in "iu":
        raise ValueError(f"{name} must be an integer array")
    return a


def validate(arrays: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """Checks an exported tree (either host form) and returns it in the compact form:
    vertices float64, elements int32, nodes NODE_RECORD, bb_indices int32, bb_coords (float64
    or None), n_buckets, cells_per_leaf, kind and depth.  Raises ValueError."""
    missing = {"vertices", "elements", "nodes", "bb_indices", "n_buckets", "cells_per_leaf", "kind"} - set(arrays)
    if missing:
        raise ValueError(f"tree arrays lack {sorted(missing)}")
    fmt = str(np.asarray(arrays.get("format", FORMAT)))
    if fmt != FORMAT:
        raise ValueError(f"unsupported tree format {fmt!r} (expected {FORMAT!r})")
    kind = str(np.asarray(arrays["kind"]))
    if kind not in KINDS:
        raise ValueError(f"kind must be one of {KINDS}, got {kind!r}")
    vertices = np.ascontiguousarray(arrays["vertices"], dtype=np.float64)
    elements = _int_array(arrays["elements"], "elements")
    bb_indices = _int_array(arrays["bb_indices"], "bb_indices")
    n_buckets = int(np.asarray(arrays["n_buckets"]))
    cells_per_leaf = int(np.asarray(arrays["cells_per_leaf"]))
    if vertices.ndim != 2 or vertices.shape[1] != 2:
        raise ValueError("vertices must have shape (n_vertices, 2)")
    if elements.ndim != 2 or elements.shape[0] == 0:
        raise ValueError("elements must have shape (n, n_max_vert) with n >= 1")
    n, width = elements.shape
    if n >= 2**31 or len(vertices) >= 2**31:
        raise ValueError("trees are limited to 2**31 cells and vertices")
    if kind == "edges" and width != 2:
        raise ValueError("an edge tree's elements must have shape (n, 2)")
    if kind == "faces" and not 3 <= width <= 32:
        raise ValueError("a face tree's elements must have 3 to 32 columns")
    if bb_indices.shape != (n,):
        raise ValueError("bb_indices must have shape (n,)")
    bb_coords = arrays.get("bb_coords")
    if bb_coords is not None:
        bb_coords = np.ascontiguousarray(bb_coords, dtype=np.float64)
        if bb_coords.shape != (n, 4):
            raise ValueError("bb_coords must have shape (n, 4)")
    if n_buckets < 2 or cells_per_leaf < 1:
        raise ValueError("n_buckets must be >= 2 and cells_per_leaf >= 1")
    nodes_in = np.asarray(arrays["nodes"])
    names = nodes_in.dtype.names or ()
    if not {"child", "Lmax", "Rmin", "ptr", "size", "dim"} <= set(names) or nodes_in.ndim != 1 or len(nodes_in) == 0:
        raise ValueError("nodes must be a non-empty 1-d array with NodeDType's fields")
    if len(nodes_in) >= 2**31:
        raise ValueError("trees are limited to 2**31 nodes")
    nodes = np.ascontiguousarray(to_node_records(nodes_in))
    if nodes_in.dtype != NODE_RECORD:  # NodeDType: its 64-bit fields must fit the records
        for name in ("child", "ptr", "size"):
            if (nodes_in[name] != nodes[name]).any():
                raise ValueError(f"invalid tree: node {name} out of int32 range")
    elements32 = _to_int32(elements, "elements")
    bb32 = _to_int32(bb_indices, "bb_indices")
    depth = check_tree(nodes, bb32, elements32, len(vertices), kind == "faces")
    return {
        "vertices": vertices, "elements": elements32, "nodes": nodes, "bb_indices": bb32, "bb_coords": bb_coords,
        "n_buckets": n_buckets, "cells_per_leaf": cells_per_leaf, "kind": kind, "depth": depth,
    }


def save(file, arrays: Dict[str, np.ndarray]) -> None:
    """Writes a tree in the compact file form (see the module docstring) to a path or a
    binary file object.  ``arrays``: vertices, elements, nodes, bb_indices, n_buckets,
    cells_per_leaf, kind (bb_coords, if present, is not stored)."""
    conv = {
        "vertices": np.ascontiguousarray(arrays["vertices"], dtype="<f8"),
        "elements": np.ascontiguousarray(arrays["elements"], dtype="<i4"),
        "nodes": np.ascontiguousarray(to_node_records(np.asarray(arrays["nodes"]))),
        "bb_indices": np.ascontiguousarray(arrays["bb_indices"], dtype="<i4"),
    }
    header = {"format": FORMAT, "kind": str(arrays["kind"]), "n_buckets": int(arrays["n_buckets"]),
              "cells_per_leaf": int(arrays["cells_per_leaf"]), "arrays": []}
    offset = 0
    for name in _FILE_ARRAYS:
        a = conv[name]
        header["arrays"].append({"name": name, "dtype": "ct_node" if name == "nodes" else a.dtype.str,
                                 "shape": list(a.shape), "offset": offset})
        offset += (a.nbytes + 63) // 64 * 64
    hb = json.dumps(header).encode()
    head = MAGIC + np.uint64(len(hb)).astype("<u8").tobytes() + hb
    pad = (-len(head)) % 64
    own = isinstance(file, (str, bytes)) or hasattr(file, "__fspath__")
    fh = open(file, "wb") if own else file
    try:
        fh.write(head + b"\0" * pad)
        for name in _FILE_ARRAYS:
            a = conv[name]
            fh.write(a.tobytes())
            fh.write(b"\0" * ((-a.nbytes) % 64))
    finally:
        if own:
            fh.close()


def load(file) -> Dict[str, np.ndarray]:
    """Reads a tree written by ``save`` (a path, a binary file object or bytes) and validates
    it.  The arrays returned are read-only views of the file's bytes."""
    # The arrays are views of these bytes (read-only; nothing writes to them).
    if isinstance(file, (bytes, bytearray, memoryview)):
        buf = file
    elif isinstance(file, str) or hasattr(file, "__fspath__"):
        with open(file, "rb") as fh:
            buf = fh.read()
    else:
        buf = file.read()
    if len(buf) < 16 or bytes(buf[:8]) != MAGIC:
        raise ValueError("not a cell tree file (bad magic)")
    hlen = int(np.frombuffer(buf, dtype="<u8", count=1, offset=8)[0])
    if hlen > len(buf) - 16:
        raise ValueError("truncated cell tree file")
    try:
        header = json.loads(bytes(buf[16:16 + hlen]).decode())
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise ValueError(f"corrupt cell tree header: {e}") from None
    if not isinstance(header, dict) or header.get("format") != FORMAT:
        raise ValueError(f"unsupported tree format (expected {FORMAT!r})")
    base = 16 + hlen + (-(16 + hlen)) % 64
    arrays = {"format": FORMAT, "kind": header.get("kind"), "n_buckets": header.get("n_buckets"),
              "cells_per_leaf": header.get("cells_per_leaf")}
    for entry in header.get("arrays", []):
        name, dt = entry.get("name"), _FILE_DTYPES.get(entry.get("dtype"))
        shape = entry.get("shape")
        if name not in _FILE_ARRAYS or dt is None or not isinstance(shape, list) \
                or not all(isinstance(d, int) and d >= 0 for d in shape):
            raise ValueError(f"corrupt cell tree header entry {entry!r}")
        count = int(np.prod(shape, dtype=np.int64))
        off = base + int(entry.get("offset", -1))
        if off < base or off + count * dt.itemsize > len(buf):
            raise ValueError(f"cell tree file too short for array {name!r}")
        arrays[name] = np.frombuffer(buf, dtype=dt, count=count, offset=off).reshape(shape)
    if any(arrays.get(k) is None for k in ("kind", "n_buckets", "cells_per_leaf")):
        raise ValueError("corrupt cell tree header")
    return validate(arrays)
