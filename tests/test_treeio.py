"""Tree export / import on the host (numba_celltree_amd.treeio): the validation that stands
between a file and the kernels, checked on trees the oracle builds (the reference's own
layouts) and on corrupted copies of them.  The device round trips are in
tests/test_gpu_treeio.py."""

import io

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from numba_celltree_amd import treeio
from oracle import refcpu as R


def _depth_by_recursion(nodes, i=0):
    c = int(nodes["child"][i])
    if c == -1:
        return 0
    return 1 + max(_depth_by_recursion(nodes, c), _depth_by_recursion(nodes, c + 1))


def _arrays(tree, kind="faces"):
    return {
        "vertices": tree.vertices, "elements": tree.faces if kind == "faces" else tree.edges,
        "nodes": tree.nodes, "bb_indices": tree.bb_indices, "bb_coords": tree.bb_coords,
        "n_buckets": tree.n_buckets, "cells_per_leaf": tree.cells_per_leaf, "kind": kind,
    }


@pytest.fixture(scope="module")
def oracle_tree():
    v, f, _ = S.triangle_grid_20k()
    return R.OracleCellTree2d(v, f, -1)


def test_validate_accepts_reference_layout_and_finds_depth(oracle_tree):
    a = treeio.validate(_arrays(oracle_tree))
    assert a["depth"] == _depth_by_recursion(oracle_tree.nodes)
    assert np.array_equal(treeio.to_node_dtype(a["nodes"]), oracle_tree.nodes)
    assert a["nodes"].dtype.itemsize == 32 and a["elements"].dtype == np.int32
    v, f = S.jittered_grid(61, 100.0, 0.25, 1.0 / 3.0, 2)
    mixed = R.OracleCellTree2d(v, f, -1, n_buckets=3, cells_per_leaf=1)
    assert treeio.validate(_arrays(mixed))["depth"] == _depth_by_recursion(mixed.nodes)
    e = R.OracleEdgeCellTree2d(np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [2.0, 1.0]]),
                               np.array([[0, 1], [1, 2], [2, 3]]))
    assert treeio.validate(_arrays(e, "edges"))["depth"] == 1


def test_save_load_round_trip(oracle_tree):
    buf = io.BytesIO()
    treeio.save(buf, _arrays(oracle_tree))
    buf.seek(0)
    a = treeio.load(buf)
    for k in ("vertices", "elements", "bb_indices"):
        assert np.array_equal(a[k], getattr(oracle_tree, "faces" if k == "elements" else k)), k
    assert np.array_equal(treeio.to_node_dtype(a["nodes"]), oracle_tree.nodes)
    assert a["bb_coords"] is None  # not stored: recomputed on import
    assert (a["n_buckets"], a["cells_per_leaf"], a["kind"]) == (4, 2, "faces")
    assert a["depth"] == _depth_by_recursion(oracle_tree.nodes)


def test_file_corruption_is_rejected(oracle_tree):
    """Truncations, a bad magic, a bad header and out-of-range array entries raise
    ValueError; a flipped child link inside the node array fails validation."""
    import json

    buf = io.BytesIO()
    treeio.save(buf, _arrays(oracle_tree))
    raw = buf.getvalue()
    for bad in (raw[:10], b"NOTATREE" + raw[8:], raw[: len(raw) // 2]):
        with pytest.raises(ValueError):
            treeio.load(bad)
    hlen = int(np.frombuffer(raw, dtype="<u8", count=1, offset=8)[0])
    header = json.loads(raw[16:16 + hlen])
    for mutate in (lambda h: h["arrays"][2].__setitem__("dtype", "|O"),
                   lambda h: h["arrays"][0].__setitem__("shape", [10**9, 2]),
                   lambda h: h["arrays"][1].__setitem__("offset", -64),
                   lambda h: h.__setitem__("format", "x")):
        h = json.loads(json.dumps(header))
        mutate(h)
        hb = json.dumps(h).encode()
        hb = hb + b" " * (hlen - len(hb)) if len(hb) <= hlen else None
        if hb is None:
            continue
        with pytest.raises(ValueError):
            treeio.load(raw[:16] + hb + raw[16 + hlen:])
    # a child link pointing backwards inside the node records
    nodes_at = 16 + hlen + (-(16 + hlen)) % 64 + header["arrays"][2]["offset"]
    rec = bytearray(raw)
    internal = int(np.flatnonzero(oracle_tree.nodes["child"] != -1)[1])
    rec[nodes_at + 32 * internal + 16: nodes_at + 32 * internal + 20] = np.int32(0).tobytes()
    with pytest.raises(ValueError, match="invalid tree"):
        treeio.load(bytes(rec))


def _corrupt(base, key, fn):
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in base.items()}
    fn(a[key]) if key else fn(a)
    return a


@pytest.mark.parametrize(
    "name, key, fn",
    [
        ("child backwards", "nodes", lambda n: n["child"].__setitem__(
            int(np.flatnonzero(n["child"] != -1)[1]), 1)),
        ("child past the end", "nodes", lambda n: n["child"].__setitem__(0, len(n) - 1)),
        ("self loop", "nodes", lambda n: n["child"].__setitem__(0, 0)),
        ("shared children", "nodes", lambda n: n["child"].__setitem__(
            int(np.flatnonzero(n["child"] != -1)[1]), n["child"][0])),
        ("cell range past n", "nodes", lambda n: n["size"].__setitem__(len(n) - 1, 10**9)),
        ("negative ptr", "nodes", lambda n: n["ptr"].__setitem__(0, -1)),
        ("bb_indices repeat", "bb_indices", lambda b: b.__setitem__(0, b[1])),
        ("bb_indices range", "bb_indices", lambda b: b.__setitem__(0, len(b))),
        ("vertex index", "elements", lambda e: e.__setitem__((0, 0), 10**7)),
        ("face of 2 vertices", "elements", lambda e: e.__setitem__((0, 2), -1)),
        ("fill below -1", "elements", lambda e: e.__setitem__((0, 0), -2)),
        ("kind", None, lambda a: a.__setitem__("kind", "cells")),
        ("format", None, lambda a: a.__setitem__("format", "something/9")),
        ("n_buckets", None, lambda a: a.__setitem__("n_buckets", 1)),
        ("bb_coords shape", None, lambda a: a.__setitem__("bb_coords", a["bb_coords"][:, :3])),
        ("missing field", None, lambda a: a.pop("nodes")),
    ],
)
def test_validate_rejects_corrupt_trees(oracle_tree, name, key, fn):
    with pytest.raises(ValueError):
        treeio.validate(_corrupt(_arrays(oracle_tree), key, fn))


def test_validate_rejects_edge_tree_with_fill(oracle_tree):
    e = R.OracleEdgeCellTree2d(np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [2.0, 1.0]]),
                               np.array([[0, 1], [1, 2], [2, 3]]))
    a = _arrays(e, "edges")
    a["elements"] = a["elements"].copy()
    a["elements"][1, 1] = -1
    with pytest.raises(ValueError):
        treeio.validate(a)
