"""Deep trees on the GPU: meshes refined geometrically toward a corner, whose cell trees are
about as deep as the number of refinement levels, with overlapping child boxes at every level
(so walks defer far children at many levels).

* 100 levels (depth 101) runs the 128-entry stack class, 200 levels (depth 201) the 256 one:
  both far past the 4 LDS stack slots, deep in the private spill array;
* 300 levels (depth 301) and 900 levels refined by 0.7 (depth ~900) exceed the largest stack
  class (CT_MAX_DEPTH): their queries keep no stack and climb parent links to find deferred
  siblings again.  (The ratio keeps every product of coordinates far above the subnormal
  range, where the CPU oracle would crawl.)

Every query kind is compared with the oracle (whose stacks grow without bound, as the
reference's, utils.py:34-97): the tree itself, points (including mesh vertices, edge
midpoints and centroids, at the default tolerance and at 0), barycentric weights, boxes,
segments crossing many levels, overlap areas, node bounds, and an edge tree of the same mesh.
"""

import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

LEVELS = [(100, 0.5), (200, 0.5), (300, 0.5), (900, 0.7)]


def _queries(v, f, levels, seed, ratio=0.5):
    rng = np.random.default_rng(seed)
    m = 4000
    # log-uniform toward the refined corner: every level gets queries
    scale = ratio ** rng.uniform(0, levels + 1, m)
    pts = np.concatenate([
        scale[:, None] * rng.uniform(0, 1, (m, 2)),
        v,
        0.5 * (v[f[:, 0]] + v[f[:, 1]]),
        v[f].mean(1),
        rng.uniform(-0.1, 1.1, (500, 2)),
    ])
    c = scale[:, None] * rng.uniform(0, 1, (m, 2))
    w = scale[:, None] * rng.uniform(0.01, 1.0, (m, 2))
    boxes = np.column_stack([c[:, 0] - w[:, 0], c[:, 0] + w[:, 0], c[:, 1] - w[:, 1], c[:, 1] + w[:, 1]])
    s2 = ratio ** rng.uniform(0, levels + 1, m)
    a = scale[:, None] * rng.uniform(0, 1, (m, 2))
    b = s2[:, None] * rng.uniform(0, 1, (m, 2))
    segs = np.stack([a, b], axis=1)
    return pts, boxes, segs


@pytest.mark.parametrize("levels, ratio", LEVELS)
def test_deep_tree_matches_oracle(levels, ratio):
    import numba_celltree_amd as M
    from numba_celltree_amd import _native

    v, f = S.graded_corner(levels, ratio)
    gpu = M.CellTree2d(v, f, -1)
    print(f"[deep] {levels} levels by {ratio}, {len(f)} cells, depth {gpu.depth}", file=sys.stderr, flush=True)
    ora = R.OracleCellTree2d(v, f, -1)
    assert gpu.depth >= 0.95 * levels
    assert (gpu.d_parents is not None) == (gpu.depth > _native.CT_MAX_DEPTH)
    assert np.array_equal(gpu.nodes, ora.nodes) and np.array_equal(gpu.bb_indices, ora.bb_indices)
    np.testing.assert_array_equal(gpu.node_bounds, ora.node_bounds)
    assert gpu.validate_node_bounds().all()

    pts, boxes, segs = _queries(v, f, levels, levels, ratio)
    for tol in (None, 0.0):
        assert np.array_equal(gpu.locate_points(pts, tol), ora.locate_points(pts, tol)), tol
    fi, w = gpu.compute_barycentric_weights(pts)
    fo, wo = ora.compute_barycentric_weights(pts)
    assert np.array_equal(fi, fo)
    np.testing.assert_allclose(w, wo, rtol=1e-12, atol=1e-12)
    for got, want in zip(gpu.locate_boxes(boxes), ora.locate_boxes(boxes)):
        assert np.array_equal(got, want)
    gi, gj, gxy = gpu.intersect_edges(segs)
    oi, oj, oxy = ora.intersect_edges(segs)
    assert len(gi) > 10 * len(segs)  # segments cross many levels
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    np.testing.assert_allclose(gxy, oxy, rtol=1e-12, atol=0)
    # overlap areas with squares at several scales
    k = np.arange(0, levels, max(1, levels // 40))
    s = ratio ** k.astype(float)
    tv = np.stack([np.column_stack([0.3 * s, 0.2 * s]), np.column_stack([0.9 * s, 0.2 * s]),
                   np.column_stack([0.9 * s, 0.8 * s]), np.column_stack([0.3 * s, 0.8 * s])], axis=1).reshape(-1, 2)
    tf = np.arange(len(tv)).reshape(-1, 4)
    for got, want in zip(gpu.intersect_faces(tv, tf, -1)[:2], ora.intersect_faces(tv, tf, -1)[:2]):
        assert np.array_equal(got, want)
    np.testing.assert_allclose(gpu.intersect_faces(tv, tf, -1)[2], ora.intersect_faces(tv, tf, -1)[2], rtol=1e-12)


@pytest.mark.parametrize("levels", [200, 400])
def test_deep_edge_tree_matches_oracle(levels):
    import numba_celltree_amd as M

    v, f = S.graded_corner(levels)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    gpu = M.EdgeCellTree2d(v, e)
    ora = R.OracleEdgeCellTree2d(v, e)
    print(f"[deep edges] {levels} levels, {len(e)} edges, depth {gpu.depth}", file=sys.stderr, flush=True)
    assert np.array_equal(gpu.nodes, ora.nodes)
    pts, _, segs = _queries(v, f, levels, 7)
    pts = np.concatenate([pts, 0.5 * (v[e[:, 0]] + v[e[:, 1]])])
    assert np.array_equal(gpu.locate_points(pts), ora.locate_points(pts))
    gi, gj, gxy = gpu.intersect_edges(segs)
    oi, oj, oxy = ora.intersect_edges(segs)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    np.testing.assert_allclose(gxy, oxy, rtol=1e-12, atol=0)


def test_batch_rejects_trees_past_the_stack_classes():
    import numba_celltree_amd as M
    from numba_celltree_amd import batch as B

    deep = M.CellTree2d(*S.graded_corner(300), -1)
    with pytest.raises(ValueError, match="depth"):
        B.TreeBatch([deep])
