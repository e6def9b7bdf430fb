"""Tree export / import on the GPU: save -> load, pickle, and import of the reference's own
tree attributes (built by the oracle) give a tree whose arrays and query results are
identical to the built tree's and to the oracle's."""

import io
import pickle
import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def _meshes():
    v, f, _ = S.triangle_grid_20k()
    yield "triangles", v, f
    v, f = S.jittered_grid(151, 1000.0, 0.25, 1.0 / 3.0, 5)
    yield "mixed", v, f
    from tests.golden import cases as C

    v, f = C.load_voronoi()
    yield "voronoi", v, f


def _same_tree(a, b):
    assert np.array_equal(a.nodes, b.nodes)
    assert np.array_equal(a.bb_indices, b.bb_indices)
    assert np.array_equal(a.bb_coords, b.bb_coords)
    assert np.array_equal(a.faces, b.faces)
    assert np.array_equal(a.vertices, b.vertices)
    assert a.depth == b.depth and a.n_nodes == b.n_nodes
    assert np.array_equal(a.bbox, b.bbox) and a._default_tolerance() == b._default_tolerance()


def _queries(v, seed):
    lo, hi = v.min(0), v.max(0)
    ext = float((hi - lo).max())
    rng = np.random.default_rng(seed)
    pts = rng.uniform(lo - 0.05 * ext, hi + 0.05 * ext, (20_000, 2))
    c = rng.uniform(lo, hi, (2_000, 2))
    w = ext * 10 ** rng.uniform(-3, -1.5, (2_000, 2))
    boxes = np.column_stack([c[:, 0] - w[:, 0], c[:, 0] + w[:, 0], c[:, 1] - w[:, 1], c[:, 1] + w[:, 1]])
    a = rng.uniform(lo, hi, (2_000, 2))
    ang = rng.uniform(0, 2 * np.pi, 2_000)
    ln = ext * 10 ** rng.uniform(-3, -1, 2_000)
    segs = np.stack([a, a + ln[:, None] * np.column_stack([np.cos(ang), np.sin(ang)])], axis=1)
    return pts, boxes, segs


@pytest.mark.parametrize("name, v, f", list(_meshes()), ids=lambda x: x if isinstance(x, str) else "")
def test_save_load_pickle_identical(name, v, f):
    import numba_celltree_amd as M

    print(f"[treeio] {name}: {len(f)} cells", file=sys.stderr, flush=True)
    built = M.CellTree2d(v, f, -1)
    buf = io.BytesIO()
    built.save(buf)
    buf.seek(0)
    loaded = M.CellTree2d.load(buf)
    unpickled = pickle.loads(pickle.dumps(built))
    ora = R.OracleCellTree2d(v, f, -1)
    # the reference's own attributes import without a rebuild
    imported = M.CellTree2d.from_arrays(vertices=ora.vertices, elements=ora.faces, nodes=ora.nodes,
                                        bb_indices=ora.bb_indices, bb_coords=ora.bb_coords,
                                        n_buckets=ora.n_buckets, cells_per_leaf=ora.cells_per_leaf)
    pts, boxes, segs = _queries(v, 3)
    want_p = ora.locate_points(pts)
    want_b = ora.locate_boxes(boxes)
    want_e = ora.intersect_edges(segs)
    tv, tf = S.hex_voronoi(12, 12, float((v.max(0) - v.min(0)).max()), 0.1, 9)
    tv = tv + v.min(0)
    want_f = ora.intersect_faces(tv, tf, -1)
    for label, t in (("loaded", loaded), ("unpickled", unpickled), ("imported", imported)):
        _same_tree(t, built)
        assert np.array_equal(t.locate_points(pts), want_p), label
        for got, want in zip(t.locate_boxes(boxes), want_b):
            assert np.array_equal(got, want), label
        gi, gj, gxy = t.intersect_edges(segs)
        assert np.array_equal(gi, want_e[0]) and np.array_equal(gj, want_e[1]), label
        np.testing.assert_allclose(gxy, want_e[2], rtol=1e-12, atol=1e-12 * np.abs(v).max())
        fi, fj, fa = t.intersect_faces(tv, tf, -1)
        assert np.array_equal(fi, want_f[0]) and np.array_equal(fj, want_f[1]), label
        np.testing.assert_allclose(fa, want_f[2], rtol=1e-12)
        assert t.validate_node_bounds().all()


def test_edge_tree_round_trip():
    import numba_celltree_amd as M
    from tests.golden import cases as C

    built = M.EdgeCellTree2d(C.E_VERTICES, C.E_EDGES)
    buf = io.BytesIO()
    built.save(buf)
    buf.seek(0)
    for t in (M.EdgeCellTree2d.load(buf), pickle.loads(pickle.dumps(built))):
        assert np.array_equal(t.nodes, built.nodes) and np.array_equal(t.edges, built.edges)
        assert np.array_equal(t.bb_coords, built.bb_coords)
        pts = np.array([[0.5, 0.0], [1.5, 0.0], [2.0, 0.5], [0.5, 0.5]])
        np.testing.assert_array_equal(t.locate_points(pts), built.locate_points(pts))
        np.testing.assert_array_equal(t.locate_points(pts), [0, 1, 2, -1])
        segs = np.array([[[1.0, -1.0], [1.0, 1.0]], [[3.0, 1.0], [-1.0, -1.0]]])
        for got, want in zip(t.intersect_edges(segs), built.intersect_edges(segs)):
            assert np.array_equal(got, want)
    buf.seek(0)
    with pytest.raises(ValueError):
        M.CellTree2d.load(buf)  # an edge tree is not a face tree


def test_import_rejects_inconsistent_boxes():
    import numba_celltree_amd as M

    v, f, _ = S.triangle_grid_20k()
    a = M.CellTree2d(v, f, -1).to_arrays()
    a["bb_coords"] = a["bb_coords"].copy()
    a["bb_coords"][17, 1] += 1.0
    with pytest.raises(ValueError, match="bounding boxes"):
        M.CellTree2d.from_arrays(**a)
