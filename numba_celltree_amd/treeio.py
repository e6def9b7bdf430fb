"""Tree export / import: a built CellTree2d or EdgeCellTree2d as plain host arrays.

The reference keeps a tree as plain numpy attributes (celltree.py:80-97,
edge_celltree.py:65-83), which pickle with the object; it has no file format of its own.
Here a tree lives on the GPU, so export copies its arrays to the host in the reference's own
layouts, and import uploads them and derives the device-only layouts again, without
rebuilding the tree:

    ``vertices``   (n_vertices, 2) float64
    ``elements``   (n, n_max_vert) int64, faces in counter-clockwise order, -1 padded
                   (an EdgeCellTree2d: its (n, 2) edges)
    ``nodes``      (n_nodes,) NodeDType (constants.py:92-106), the reference's DFS numbering
    ``bb_indices`` (n,) intp
    ``bb_coords``  (n, 4) float64
    ``n_buckets``, ``cells_per_leaf``, ``kind`` ("faces" / "edges"), ``format``

These are exactly the attributes a reference tree holds, so a tree the reference built
imports too (``CellTree2d.from_arrays(vertices=t.vertices, elements=t.faces,
nodes=t.nodes, bb_indices=t.bb_indices, bb_coords=t.bb_coords, ...)``).

Everything imported is checked on the host before it reaches a kernel (the traversal
kernels do not bounds-check node links): node links, cell ranges, the bb_indices
permutation, vertex indices, and that bb_coords are the bounding boxes of the elements.
"""

from typing import Dict

import numpy as np

from numba_celltree_amd.constants import NodeDType

FORMAT = "numba_celltree_amd.tree/1"
KINDS = ("faces", "edges")


def tree_depth(nodes: np.ndarray) -> int:
    """Longest root-to-leaf path (in edges) of a node array, checking on the way that the
    child links form one tree over all nodes: every non-root node is reached exactly once,
    and children come after their parent (DFS numbering), so there is no cycle."""
    n = len(nodes)
    child = nodes["child"].astype(np.int64)
    internal = child != -1
    idx = np.arange(n, dtype=np.int64)
    if (child[internal] <= idx[internal]).any() or (child[internal] + 1 >= n).any():
        raise ValueError("invalid tree: a child link points backwards or past the last node")
    if (child < -1).any():
        raise ValueError("invalid tree: negative child index")
    seen = np.zeros(n, dtype=np.int64)
    seen[0] = 1
    frontier = np.array([0], dtype=np.int64)
    depth = 0
    while True:
        c = child[frontier]
        c = c[c != -1]
        if c.size == 0:
            break
        frontier = np.concatenate([c, c + 1])
        np.add.at(seen, frontier, 1)
        depth += 1
        if depth > n:
            raise ValueError("invalid tree: deeper than its node count")
    if (seen != 1).any():
        raise ValueError("invalid tree: the child links do not reach every node exactly once")
    return depth


def validate(arrays: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """Checks an exported tree and returns it with canonical dtypes (raises ValueError)."""
    missing = {"vertices", "elements", "nodes", "bb_indices", "bb_coords", "n_buckets",
               "cells_per_leaf", "kind"} - set(arrays)
    if missing:
        raise ValueError(f"tree arrays lack {sorted(missing)}")
    fmt = str(np.asarray(arrays.get("format", FORMAT)))
    if fmt != FORMAT:
        raise ValueError(f"unsupported tree format {fmt!r} (expected {FORMAT!r})")
    kind = str(np.asarray(arrays["kind"]))
    if kind not in KINDS:
        raise ValueError(f"kind must be one of {KINDS}, got {kind!r}")
    vertices = np.ascontiguousarray(arrays["vertices"], dtype=np.float64)
    elements = np.ascontiguousarray(arrays["elements"]).astype(np.int64)
    nodes_in = np.asarray(arrays["nodes"])
    bb_indices = np.ascontiguousarray(arrays["bb_indices"]).astype(np.int64)
    bb_coords = np.ascontiguousarray(arrays["bb_coords"], dtype=np.float64)
    n_buckets = int(np.asarray(arrays["n_buckets"]))
    cells_per_leaf = int(np.asarray(arrays["cells_per_leaf"]))
    if vertices.ndim != 2 or vertices.shape[1] != 2:
        raise ValueError("vertices must have shape (n_vertices, 2)")
    if elements.ndim != 2 or elements.shape[0] == 0:
        raise ValueError("elements must have shape (n, n_max_vert) with n >= 1")
    n, width = elements.shape
    if kind == "edges" and width != 2:
        raise ValueError("an edge tree's elements must have shape (n, 2)")
    if kind == "faces" and not 3 <= width <= 32:
        raise ValueError("a face tree's elements must have 3 to 32 columns")
    lo = -1 if kind == "faces" else 0
    if elements.min() < lo or elements.max() >= len(vertices) or elements[:, : min(width, 3)].min() < 0:
        raise ValueError("element indices out of range of the vertices (or a face with fewer than 3 vertices)")
    if bb_indices.shape != (n,) or bb_coords.shape != (n, 4):
        raise ValueError("bb_indices must have shape (n,) and bb_coords (n, 4)")
    if bb_indices.min() < 0 or bb_indices.max() >= n or not (np.bincount(bb_indices, minlength=n) == 1).all():
        raise ValueError("bb_indices must be a permutation of range(n)")
    if n_buckets < 2 or cells_per_leaf < 1:
        raise ValueError("n_buckets must be >= 2 and cells_per_leaf >= 1")
    names = nodes_in.dtype.names or ()
    if not {"child", "Lmax", "Rmin", "ptr", "size", "dim"} <= set(names) or nodes_in.ndim != 1 or len(nodes_in) == 0:
        raise ValueError("nodes must be a non-empty 1-d array with NodeDType's fields")
    nodes = np.empty(len(nodes_in), dtype=NodeDType)
    for name in NodeDType.names:
        nodes[name] = nodes_in[name]
    ptr = nodes["ptr"].astype(np.int64)
    size = nodes["size"].astype(np.int64)
    if (ptr < 0).any() or (size < 0).any() or (ptr + size > n).any():
        raise ValueError("invalid tree: a node's cell range leaves [0, n)")
    if len(nodes) >= 2**31 or n >= 2**31:
        raise ValueError("trees are limited to 2**31 nodes and cells")
    depth = tree_depth(nodes)
    return {
        "vertices": vertices, "elements": elements, "nodes": nodes, "bb_indices": bb_indices,
        "bb_coords": bb_coords, "n_buckets": n_buckets, "cells_per_leaf": cells_per_leaf, "kind": kind,
        "depth": depth,
    }


def save(file, arrays: Dict[str, np.ndarray]) -> None:
    """Writes an exported tree as an .npz archive (no pickled objects)."""
    out = dict(arrays)
    out["format"] = np.array(FORMAT)
    out["kind"] = np.array(str(out["kind"]))
    np.savez(file, **out)


def load(file) -> Dict[str, np.ndarray]:
    """Reads an .npz tree archive (numpy's loader with allow_pickle=False) and validates it."""
    with np.load(file, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    return validate(arrays)
