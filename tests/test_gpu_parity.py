"""HIP library vs the CPU oracle on seeded synthetic meshes.

Bar: integer and index results bit-exact, in the reference's order (the oracle
reproduces the reference's chunk-concatenated order, so ordered arrays are compared, which
implies equality of the sorted (query, cell) sets); FP64 results (areas, clipped points,
weights) within 1e-12 relative, measured against max(|value|, scale).
"""

import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def log(msg):
    print(f"[gpu parity] {msg}", file=sys.stderr, flush=True)


def assert_close(actual, expected, scale):
    actual = np.asarray(actual)
    expected = np.asarray(expected)
    assert actual.shape == expected.shape
    both_nan = np.isnan(actual) & np.isnan(expected)
    err = np.where(both_nan, 0.0, np.abs(actual - expected))
    bound = RTOL * np.maximum(np.abs(expected), scale)
    assert np.all(err <= bound), f"max err {np.nanmax(err)} vs bound {bound.min()}"


def same_tree(gpu, ora):
    gn, on = gpu.nodes, ora.nodes
    assert gn.shape == on.shape, (gn.shape, on.shape)
    for field in ("child", "ptr", "size", "dim"):
        assert np.array_equal(gn[field], on[field]), field
    for field in ("Lmax", "Rmin"):
        assert np.array_equal(gn[field].view(np.int64), on[field].view(np.int64)), field
    assert np.array_equal(gpu.bb_indices, ora.bb_indices)
    assert np.array_equal(gpu.bb_coords, ora.bb_coords)
    assert np.array_equal(gpu.faces if hasattr(gpu, "faces") else gpu.edges,
                          ora.faces if hasattr(ora, "faces") else ora.edges)


MESHES = {
    "tri20k": lambda: S.triangle_grid_20k()[:2],
    "mixed_45k": lambda: S.mixed_mesh_1m(nv=151, seed=3),
    "tri_voronoi": lambda: S.hex_voronoi(60, 50, 100.0, 0.1, 5),
}


@pytest.fixture(scope="module", params=list(MESHES))
def mesh(request):
    return request.param, MESHES[request.param]()


@pytest.mark.parametrize("n_buckets,cells_per_leaf", [(4, 2), (2, 1), (8, 4), (3, 1)])
def test_tree_structure_identical(mesh, n_buckets, cells_per_leaf):
    import numba_celltree_amd as M

    name, (v, f) = mesh
    log(f"tree {name} nb={n_buckets} cpl={cells_per_leaf}")
    g = M.CellTree2d(v, f, -1, n_buckets=n_buckets, cells_per_leaf=cells_per_leaf)
    o = R.OracleCellTree2d(v, f, -1, n_buckets=n_buckets, cells_per_leaf=cells_per_leaf)
    same_tree(g, o)
    assert np.array_equal(g.node_bounds, o.node_bounds)
    assert g.validate_node_bounds().all()
    assert np.array_equal(g.bb_distances, o.bb_distances)
    assert np.array_equal(g.bbox, o.bbox)


@pytest.fixture(scope="module")
def trees(mesh):
    import numba_celltree_amd as M

    name, (v, f) = mesh
    return name, v, f, M.CellTree2d(v, f, -1), R.OracleCellTree2d(v, f, -1)


def _extent(v):
    return float(max(v[:, 0].max() - v[:, 0].min(), v[:, 1].max() - v[:, 1].min()))


def test_locate_points(trees):
    name, v, f, g, o = trees
    L = _extent(v)
    lo, hi = v.min(0), v.max(0)
    rng = np.random.default_rng(1)
    pts = rng.uniform(lo - 0.05 * L, hi + 0.05 * L, size=(200_000, 2))
    # degenerate queries: every vertex, every edge midpoint, every cell centroid
    fv = f[(f >= 0).all(1)] if f.shape[1] == 3 else f
    a = v[fv[:, 0]]
    b = v[fv[:, 1]]
    special = np.concatenate([v, 0.5 * (a + b), v[f[:, :3]].mean(1)])
    pts = np.concatenate([pts, special])
    log(f"locate_points {name} n={len(pts)}")
    for tol in (None, 1e-9, 1e-3 * L):
        assert np.array_equal(g.locate_points(pts, tol), o.locate_points(pts, tol)), tol


def test_locate_points_every_path(trees):
    """Plain (unpacked leaves), plain (packed leaves) and ordered: identical answers."""
    import torch

    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    name, v, f, g, o = trees
    pts = np.random.default_rng(7).uniform(v.min(0) - 1.0, v.max(0) + 1.0, size=(100_000, 2))
    want = o.locate_points(pts)
    tol = g._default_tolerance()
    d_pts = torch.from_numpy(pts).cuda()
    saved = g._ctree.leaf_coords
    try:
        for packed in (False, True):
            g._ctree.leaf_coords = saved if packed else None
            out = D.empty(len(pts), torch.int64)
            _native.call("ct_locate_points", g.tree_ref, D.ptr(d_pts), len(pts), tol, D.ptr(out), D.stream())
            assert np.array_equal(out.cpu().numpy(), want), f"plain packed={packed}"
            ws = D.empty(int(_native.lib().ct_locate_workspace_bytes(len(pts))), torch.uint8)
            out2 = D.empty(len(pts), torch.int64)
            _native.call(
                "ct_locate_points_ordered", g.tree_ref, D.ptr(d_pts), len(pts), tol, D.ptr(ws), ws.numel(),
                D.ptr(out2), D.stream(),
            )
            assert np.array_equal(out2.cpu().numpy(), want), f"ordered packed={packed}"
            for with_sorted in (False, True):
                order, srt = D.query_order(0, d_pts, g._bbox, with_sorted=True)
                out3 = D.empty(len(pts), torch.int64)
                _native.call(
                    "ct_locate_points_in_order", g.tree_ref, D.ptr(d_pts), len(pts), D.ptr(order),
                    D.ptr(srt if with_sorted else None), tol, D.ptr(out3), D.stream(),
                )
                assert np.array_equal(out3.cpu().numpy(), want), f"in_order sorted={with_sorted} packed={packed}"
    finally:
        g._ctree.leaf_coords = saved


def _grid_points(bbox, n, rng):
    """Points on the start grid's cell borders (the kernel's own arithmetic: lo + k * w with
    w = (hi - lo) / n), on the bbox, outside it, and NaN."""
    xmin, xmax, ymin, ymax = bbox
    wx, wy = (xmax - xmin) / n, (ymax - ymin) / n
    kx = rng.integers(0, n + 1, 4000)
    ky = rng.integers(0, n + 1, 4000)
    bx = xmin + kx * wx
    by = ymin + ky * wy
    free_x = rng.uniform(xmin, xmax, 4000)
    free_y = rng.uniform(ymin, ymax, 4000)
    pts = [np.stack([bx, free_y], 1), np.stack([free_x, by], 1), np.stack([bx, by], 1),
           np.array([[xmin, ymin], [xmax, ymax], [xmin - 1.0, ymin], [np.nan, ymin], [xmin, np.nan],
                     [np.inf, ymin], [-np.inf, -np.inf]])]
    return np.concatenate(pts)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 64, 4096])
def test_locate_points_start_grid(trees, n):
    """Point walks that begin at the start-grid node give the oracle's answers for every grid
    size (none, one cell, fine), including points on grid-cell borders, NaN and inf."""
    import torch

    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    name, v, f, g, o = trees
    rng = np.random.default_rng(11 + n)
    pts = np.concatenate([rng.uniform(v.min(0) - 1.0, v.max(0) + 1.0, size=(50_000, 2)),
                          _grid_points(g._bbox, max(n, 1), rng)])
    want = o.locate_points(pts)
    saved = (g._ctree.start_grid, g._ctree.start_grid_n)
    grid = D.empty((max(n, 1), max(n, 1)), torch.int32)
    try:
        if n:
            _native.call("ct_start_grid", g.tree_ref, n, D.ptr(grid), D.stream())
            g._ctree.start_grid, g._ctree.start_grid_n = grid.data_ptr(), n
        else:
            g._ctree.start_grid, g._ctree.start_grid_n = None, 0
        log(f"start grid {name} n={n}")
        assert np.array_equal(g.locate_points(pts), want)
        d_pts = torch.from_numpy(pts).cuda()
        out = D.empty(len(pts), torch.int64)
        _native.call("ct_locate_points", g.tree_ref, D.ptr(d_pts), len(pts), g._default_tolerance(),
                     D.ptr(out), D.stream())
        assert np.array_equal(out.cpu().numpy(), want)
    finally:
        g._ctree.start_grid, g._ctree.start_grid_n = saved


def test_start_grid_nodes_are_forced():
    """Every start-grid node is on the root path of its cell's points with nothing deferred:
    for sampled points of each cell, the oracle-side forced descent passes through it."""
    from numba_celltree_amd import CellTree2d

    v, f = S.triangle_grid_20k()[:2]
    g = CellTree2d(v, f, -1)
    n = 16
    grid = g.d_start_grid
    if grid is None or grid.shape[0] != n:
        import torch

        from numba_celltree_amd import _device as D
        from numba_celltree_amd import _native

        grid = D.empty((n, n), torch.int32)
        _native.call("ct_start_grid", g.tree_ref, n, D.ptr(grid), D.stream())
    grid = grid.cpu().numpy()
    nodes = g.nodes
    xmin, xmax, ymin, ymax = g._bbox
    rng = np.random.default_rng(5)
    deep = 0
    for gy in range(n):
        for gx in range(n):
            x0, x1 = xmin + gx * (xmax - xmin) / n, xmin + (gx + 1) * (xmax - xmin) / n
            y0, y1 = ymin + gy * (ymax - ymin) / n, ymin + (gy + 1) * (ymax - ymin) / n
            start = int(grid[gy, gx])
            deep += start != 0
            for x, y in rng.uniform([x0, y0], [x1, y1], size=(20, 2)):
                node = 0
                while node != start:
                    nd = nodes[node]
                    assert nd["child"] != -1, (gx, gy)
                    val = y if nd["dim"] else x
                    left, right = val <= nd["Lmax"], val >= nd["Rmin"]
                    assert left != right, (gx, gy, node)  # one child only: nothing deferred
                    node = int(nd["child"]) + (0 if left else 1)
    assert deep > n * n // 2  # the grid does skip levels


def _first_descent_deferred(nodes, x, y):
    """Far children deferred on the near-first descent from the root to the first leaf."""
    node, deferred = 0, 0
    while nodes[node]["child"] != -1:
        nd = nodes[node]
        val = y if nd["dim"] else x
        left, right = val <= nd["Lmax"], val >= nd["Rmin"]
        if not (left or right):
            break
        if left and right:
            deferred += 1
            go_right = (nd["Lmax"] - val) < (val - nd["Rmin"])
        else:
            go_right = right
        node = int(nd["child"]) + (1 if go_right else 0)
    return deferred


def test_locate_points_overlapping_cells():
    """Heavily overlapping cells: many two-way splits per walk, so the walk stack grows past
    its 4 LDS slots into the private spill array and back."""
    from numba_celltree_amd import CellTree2d

    rng = np.random.default_rng(3)
    n = 4000
    c = rng.uniform(0.0, 100.0, size=(n, 2))
    r = rng.uniform(1.0, 30.0, size=n)
    ang = rng.uniform(0.0, 2 * np.pi, size=(n, 1)) + np.array([0.0, 2.1, 4.2])
    v = np.stack([c[:, :1] + r[:, None] * np.cos(ang), c[:, 1:] + r[:, None] * np.sin(ang)], -1).reshape(-1, 2)
    f = np.arange(3 * n).reshape(n, 3)
    g = CellTree2d(v, f, -1)
    o = R.OracleCellTree2d(v, f, -1)
    pts = rng.uniform(-10.0, 110.0, size=(100_000, 2))
    # deferred siblings along the first root-to-leaf descent of a sample of walks: past the
    # 4 LDS slots of the walk stack
    deepest = max(_first_descent_deferred(g.nodes, x, y) for x, y in pts[:300])
    log(f"overlapping cells: depth {g.depth}, deepest first-descent stack {deepest}")
    assert deepest > 4
    assert np.array_equal(g.locate_points(pts), o.locate_points(pts))


def _degenerate_mesh(m, width, seed):
    """Jittered m x m quad grid whose cells are written in degenerate ways: plain quads,
    fill-padded triangle pairs, and rows with repeated vertices (at the start, middle and
    end), in a face array of `width` columns."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(m + 1.0), np.arange(m + 1.0), indexing="xy"), -1).reshape(-1, 2)
    v = g + rng.uniform(-0.2, 0.2, size=g.shape)
    rows = []
    for j in range(m):
        for i in range(m):
            a = j * (m + 1) + i
            b, c, d = a + 1, a + m + 2, a + m + 1
            kind = (i * 7 + j * 3) % 5
            if kind == 0:
                rows += [[a, b, c, d]]
            elif kind == 1:
                rows += [[a, b, c], [a, c, d]]
            elif kind == 2:
                rows += [[a, b, b, c], [a, c, d, d]]
            elif kind == 3:
                rows += [[a, a, b, c], [c, d, a, c]]
            else:
                rows += [[a, b, c, c, d]] if width > 4 else [[a, b, c, d]]
    f = np.full((len(rows), width), -1, dtype=np.int64)
    for r, row in enumerate(rows):
        f[r, : len(row)] = row
    return v, f


@pytest.mark.parametrize("width", [4, 6])
def test_locate_points_degenerate_cells(width):
    """Packed rows padded with a repeated vertex (and rows with repeated vertices of their
    own) answer exactly as the reference's point_in_polygon_or_on_edge loop."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    v, f = _degenerate_mesh(40, width, seed=width)
    g = M.CellTree2d(v, f, -1)
    o = R.OracleCellTree2d(v, f, -1)
    assert g._ctree.leaf_coords is not None
    rng = np.random.default_rng(11)
    fa = np.where(f >= 0, f, f[:, :1])
    mid = 0.5 * (v[fa[:, 0]] + v[fa[:, 1]])
    pts = np.concatenate([rng.uniform(-1.0, 41.0, size=(50_000, 2)), v, mid, v[fa[:, :3]].mean(1)])
    log(f"degenerate cells width={width} faces={len(f)} points={len(pts)}")
    d_pts = torch.from_numpy(pts).cuda()
    saved = g._ctree.leaf_coords
    try:
        for tol in (None, 1e-9, 0.05):
            want = o.locate_points(pts, tol)
            t = g._default_tolerance() if tol is None else tol
            for packed in (True, False):
                g._ctree.leaf_coords = saved if packed else None
                out = D.empty(len(pts), torch.int64)
                _native.call("ct_locate_points", g.tree_ref, D.ptr(d_pts), len(pts), t, D.ptr(out), D.stream())
                assert np.array_equal(out.cpu().numpy(), want), (tol, packed)
    finally:
        g._ctree.leaf_coords = saved


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_query_order_is_permutation(kind):
    """ct_query_order returns a permutation, also for clustered, outside and NaN queries."""
    import torch

    from numba_celltree_amd import _device as D

    rng = np.random.default_rng(kind)
    n = 300_001
    if kind == 0:
        q = rng.uniform(-10, 110, size=(n, 2))
        q[:1000] = 50.0  # one bucket
        q[1000] = np.nan
    elif kind == 1:
        c = rng.uniform(0, 100, size=(n, 2))
        w = np.exp(rng.uniform(-5, 5, size=(n, 2)))
        q = np.column_stack([c[:, 0] - w[:, 0], c[:, 0] + w[:, 0], c[:, 1] - w[:, 1], c[:, 1] + w[:, 1]])
        q[7] = np.inf
    else:
        q = S.segments(n, 100.0, 1.0, 0.01, 100.0, 3)
    d_q = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    order = D.query_order(kind, d_q, [0.0, 100.0, 0.0, 100.0])
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(n))
    # the same order, plus the queries written in it
    order2, srt = D.query_order(kind, d_q, [0.0, 100.0, 0.0, 100.0], with_sorted=True)
    o2 = order2.cpu().numpy()
    assert np.array_equal(np.sort(o2), np.arange(n))
    rows = q.reshape(n, -1)[o2]
    assert np.array_equal(srt.cpu().numpy().view(np.int64), rows.view(np.int64))


def _point_keys(q, bbox):
    """Host restatement of the 20-bit point key: Morton code of a 1024 x 1024 grid over bbox."""
    def cell(x, lo, hi):
        g = (x - lo) * (1.0 / (hi - lo)) * 1024.0
        g = np.where(g > 0.0, np.where(g < 1023.0, g, 1023.0), 0.0)  # NaN -> 0
        return g.astype(np.uint32)

    def spread(v):
        out = np.zeros_like(v)
        for b in range(10):
            out |= ((v >> b) & 1) << (2 * b)
        return out

    return spread(cell(q[:, 0], bbox[0], bbox[1])) | (spread(cell(q[:, 1], bbox[2], bbox[3])) << 1)


def test_query_order_groups_and_chunks():
    """The point order is grouped by the top 8 key bits and sorted by the full key within
    every chunk of 32768 entries of a group; a clustered set makes groups of many chunks."""
    import torch

    from numba_celltree_amd import _device as D

    rng = np.random.default_rng(5)
    bbox = [0.0, 100.0, 0.0, 100.0]
    q = np.concatenate([
        rng.uniform(-5, 105, size=(150_000, 2)),
        rng.normal(30.0, 0.5, size=(140_000, 2)),  # ~4 chunks in one or two coarse groups
        rng.uniform(0, 100, size=(1, 2)),
    ])
    q[17] = np.nan
    n = len(q)
    keys = _point_keys(q, bbox)
    order = D.query_order(0, torch.from_numpy(q).cuda(), bbox).cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(n))
    k = keys[order]
    coarse = k >> 12
    assert np.all(np.diff(coarse.astype(np.int64)) >= 0)
    n_chunks = 0
    for c in np.unique(coarse):
        lo, hi = np.searchsorted(coarse, c, "left"), np.searchsorted(coarse, c, "right")
        for s in range(lo, hi, 32768):
            seg = k[s:min(s + 32768, hi)]
            assert np.all(np.diff(seg.astype(np.int64)) >= 0), (c, s)
            n_chunks += 1
    assert n_chunks > len(np.unique(coarse))  # some group holds several chunks


def test_locate_boxes_ordered(trees):
    name, v, f, g, o = trees
    L = _extent(v)
    h = L / np.sqrt(len(f))
    bx = S.boxes(20_000, L, h, 0.1, 10.0, 4) + np.array([v[:, 0].min()] * 2 + [v[:, 1].min()] * 2)
    log(f"locate_boxes {name}")
    gi, gj = g.locate_boxes(bx)
    oi, oj = o.locate_boxes(bx)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    gi, gj, ga = g.intersect_boxes(bx)
    oi, oj, oa = o.intersect_boxes(bx)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert_close(ga, oa, h * h)


def test_intersect_edges_ordered(trees):
    name, v, f, g, o = trees
    L = _extent(v)
    h = L / np.sqrt(len(f))
    seg = S.segments(20_000, L, h, 0.1, 20.0, 5) + v.min(0)[None, None, :]
    # segments along mesh edges and through vertices (collinear / touching branches)
    fv = f[:, :3]
    along = np.stack([v[fv[:500, 0]], v[fv[:500, 1]]], axis=1)
    through = np.stack([v[fv[500:1000, 0]] - 0.5 * h, v[fv[500:1000, 0]] + 0.5 * h], axis=1)
    seg = np.concatenate([seg, along, through, along[:, ::-1]])
    log(f"intersect_edges {name} n={len(seg)}")
    gi, gj, gxy = g.intersect_edges(seg)
    oi, oj, oxy = o.intersect_edges(seg)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert_close(gxy, oxy, L)


def test_leaf_bboxes_do_not_change_results(trees):
    """Box and segment queries with and without the leaf-ordered cell boxes: identical rows
    (bit for bit), and the oracle's."""
    name, v, f, g, o = trees
    L = _extent(v)
    h = L / np.sqrt(len(f))
    bx = S.boxes(5_000, L, h, 0.1, 10.0, 8) + np.array([v[:, 0].min()] * 2 + [v[:, 1].min()] * 2)
    seg = S.segments(5_000, L, h, 0.1, 20.0, 9) + v.min(0)[None, None, :]
    assert g._ctree.leaf_bb
    saved = g._ctree.leaf_bb
    got = {}
    try:
        for on in (True, False):
            g._ctree.leaf_bb = saved if on else None
            log(f"leaf bboxes {name} on={on}")
            got[on] = (g.locate_boxes(bx), g.intersect_edges(seg))
    finally:
        g._ctree.leaf_bb = saved
    (bi, bj), (ei, ej, exy) = got[True]
    (ci, cj), (fi, fj, fxy) = got[False]
    assert np.array_equal(bi, ci) and np.array_equal(bj, cj)
    assert np.array_equal(ei, fi) and np.array_equal(ej, fj)
    assert np.array_equal(exy.view(np.int64), fxy.view(np.int64))
    oi, oj = o.locate_boxes(bx)
    assert np.array_equal(bi, oi) and np.array_equal(bj, oj)
    oi, oj, oxy = o.intersect_edges(seg)
    assert np.array_equal(ei, oi) and np.array_equal(ej, oj)
    assert_close(exy, oxy, L)


def test_intersect_faces(trees):
    name, v, f, g, o = trees
    L = _extent(v)
    tv, tf = S.hex_voronoi(40, 40, L, 0.1, 9)
    tv = tv + v.min(0)
    log(f"intersect_faces {name}")
    gi, gj, ga = g.intersect_faces(tv, tf, -1)
    oi, oj, oa = o.intersect_faces(tv, tf, -1)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert_close(ga, oa, (L / 40) ** 2)
    gi, gj = g.locate_faces(tv, tf.astype(np.int64).copy())
    oi, oj = o.locate_faces(tv, tf.astype(np.int64).copy())
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)


def test_barycentric(trees):
    name, v, f, g, o = trees
    lo, hi = v.min(0), v.max(0)
    pts = np.random.default_rng(2).uniform(lo, hi, size=(100_000, 2))
    pts = np.concatenate([pts, v[:1000]])
    log(f"barycentric {name}")
    gf, gw = g.compute_barycentric_weights(pts)
    of, ow = o.compute_barycentric_weights(pts)
    assert np.array_equal(gf, of)
    assert_close(gw, ow, 1.0)


def test_edge_tree(trees):
    import numba_celltree_amd as M

    name, v, f, _, _ = trees
    fv = f[:, :3]
    edges = np.unique(np.sort(np.concatenate([fv[:, [0, 1]], fv[:, [1, 2]]]), axis=1), axis=0)[:20_000]
    log(f"edge tree {name} n={len(edges)}")
    g = M.EdgeCellTree2d(v, edges)
    o = R.OracleEdgeCellTree2d(v, edges)
    same_tree(g, o)
    L = _extent(v)
    mid = 0.5 * (v[edges[:, 0]] + v[edges[:, 1]])
    pts = np.concatenate([mid, v[edges[:, 0]], np.random.default_rng(3).uniform(v.min(0), v.max(0), (50_000, 2))])
    assert np.array_equal(g.locate_points(pts), o.locate_points(pts))
    h = L / np.sqrt(len(f))
    seg = S.segments(5_000, L, h, 0.1, 10.0, 6) + v.min(0)[None, None, :]
    gi, gj, gxy = g.intersect_edges(seg)
    oi, oj, oxy = o.intersect_edges(seg)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert_close(gxy, oxy, L)


def _node_boxes_host(nodes, bb_indices, bb):
    """Tight box of every node's cells, bottom-up over the DFS numbering (children of node
    j have larger indices), NaN-poisoned like the device kernel."""
    n = len(nodes)
    out = np.empty((n, 4))
    for j in range(n - 1, -1, -1):
        c = nodes["child"][j]
        if c == -1:
            b = bb[bb_indices[nodes["ptr"][j]: nodes["ptr"][j] + nodes["size"][j]]]
            box = np.array([b[:, 0].min(), b[:, 1].max(), b[:, 2].min(), b[:, 3].max()])
            if np.isnan(b).any():
                box[:] = np.nan
        else:
            l, r = out[c], out[c + 1]
            box = np.array([min(l[0], r[0]), max(l[1], r[1]), min(l[2], r[2]), max(l[3], r[3])])
            if np.isnan(l).any() or np.isnan(r).any():
                box[:] = np.nan
        out[j] = box
    return out


def test_node_boxes(trees):
    """ct_node_boxes equals the host bottom-up union of the cells' boxes, and every node's
    box contains its children's."""
    name, v, f, g, o = trees
    log(f"node boxes {name}")
    want = _node_boxes_host(o.nodes, o.bb_indices, o.bb_coords)
    got = g.d_node_boxes.cpu().numpy()[1:]
    assert np.array_equal(got, want)
    assert np.isnan(g.d_node_boxes.cpu().numpy()[0]).all()


def test_intersect_edges_pruning_is_exact(trees):
    """Segment-face queries with and without node-box pruning give the oracle's rows, also
    for very long, axis-parallel, degenerate and non-finite segments."""
    name, v, f, g, o = trees
    L = _extent(v)
    lo = v.min(0)
    rng = np.random.default_rng(11)
    n = 4000
    a = lo + rng.uniform(-0.1 * L, 1.1 * L, size=(n, 2))
    b = lo + rng.uniform(-0.1 * L, 1.1 * L, size=(n, 2))  # long: across the mesh
    seg = np.stack([a, b], axis=1)
    horiz = seg[:500].copy()
    horiz[:, 1, 1] = horiz[:, 0, 1]
    vert = seg[500:1000].copy()
    vert[:, 1, 0] = vert[:, 0, 0]
    point = seg[1000:1100].copy()
    point[:, 1] = point[:, 0]
    # segments through the vertices and along the cell edges, end to end across the mesh
    fv = f[:, :3]
    edge_line = np.stack([v[fv[:300, 0]], v[fv[:300, 0]] + 50.0 * (v[fv[:300, 1]] - v[fv[:300, 0]])], axis=1)
    bad = seg[1100:1110].copy()
    bad[0, 0, 0] = np.nan
    bad[1, 1, 1] = np.inf
    bad[2, 0] = -np.inf
    bad[3] = np.nan
    seg = np.concatenate([seg, horiz, vert, point, edge_line, bad])
    log(f"edges pruning {name} n={len(seg)}")
    oi, oj, oxy = o.intersect_edges(seg)
    gi, gj, gxy = g.intersect_edges(seg)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert_close(gxy, oxy, L)
    saved = g._ctree.node_boxes
    try:
        g._ctree.node_boxes = None
        ui, uj, uxy = g.intersect_edges(seg)
    finally:
        g._ctree.node_boxes = saved
    assert np.array_equal(ui, gi) and np.array_equal(uj, gj)
    assert np.array_equal(uxy.view(np.int64), gxy.view(np.int64))


def test_intersect_edges_hit_record_is_exact(trees):
    """The write pass replayed from the count pass's hit record gives the rows of a second
    tree walk, bit for bit: record sizes 1, 3 and 16 mix replayed and re-walked segments
    (the segments range from point-like to across the mesh)."""
    import torch

    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    name, v, f, g, o = trees
    L = _extent(v)
    lo = v.min(0)
    rng = np.random.default_rng(21)
    n = 6000
    a = lo + rng.uniform(-0.05 * L, 1.05 * L, size=(n, 2))
    length = L * np.exp(rng.uniform(np.log(1e-4), 0.0, size=(n, 1)))
    ang = rng.uniform(0, 2 * np.pi, size=(n, 1))
    b = a + length * np.concatenate([np.cos(ang), np.sin(ang)], axis=1)
    seg = torch.from_numpy(np.ascontiguousarray(np.stack([a, b], axis=1))).cuda()
    log(f"edges hit record {name} n={n}")
    st = D.stream()
    order, srt = g._order(2, seg)

    def run(slots):
        counts = D.empty(n, torch.int64)
        hits = D.empty(max(slots, 1) * n, torch.int32) if slots else None
        if slots:
            _native.call("ct_locate_edges_count_rec", g.tree_ref, D.ptr(seg), n, 1, D.ptr(order), D.ptr(srt),
                         D.ptr(counts), D.ptr(hits), slots, st)
        else:
            _native.call("ct_locate_edges_count", g.tree_ref, D.ptr(seg), n, 1, D.ptr(order), D.ptr(srt),
                         D.ptr(counts), st)
        offsets, total = D.exclusive_scan(counts)
        ii = D.empty(total, torch.int64)
        jj = D.empty(total, torch.int64)
        xy = D.empty((total, 2, 2), torch.float64)
        if slots:
            _native.call("ct_locate_edges_write_rec", g.tree_ref, D.ptr(seg), n, 1, D.ptr(order), D.ptr(srt),
                         D.ptr(offsets), D.ptr(hits), slots, D.ptr(ii), D.ptr(jj), D.ptr(xy), st)
        else:
            _native.call("ct_locate_edges_write", g.tree_ref, D.ptr(seg), n, 1, D.ptr(order), D.ptr(srt),
                         D.ptr(offsets), D.ptr(ii), D.ptr(jj), D.ptr(xy), st)
        return counts.cpu().numpy(), ii.cpu().numpy(), jj.cpu().numpy(), xy.cpu().numpy()

    c0, i0, j0, x0 = run(0)
    assert (c0 > 16).any() and (c0 <= 1).any()
    for slots in (1, 3, 16):
        c, i, j, x = run(slots)
        assert np.array_equal(c, c0) and np.array_equal(i, i0) and np.array_equal(j, j0), slots
        assert np.array_equal(x.view(np.int64), x0.view(np.int64)), slots


def test_locate_boxes_hit_record_is_exact(trees):
    """Box pairs copied from the count pass's hit record equal a second tree walk's: record
    sizes 1, 5 and 32 with boxes from a fraction of a cell to a third of the mesh."""
    import torch

    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native

    name, v, f, g, o = trees
    L = _extent(v)
    lo = v.min(0)
    rng = np.random.default_rng(23)
    n = 6000
    c = lo + rng.uniform(-0.05 * L, 1.05 * L, size=(n, 2))
    half = 0.5 * L * np.exp(rng.uniform(np.log(1e-4), np.log(0.3), size=(n, 2)))
    boxes = np.column_stack([c[:, 0] - half[:, 0], c[:, 0] + half[:, 0], c[:, 1] - half[:, 1], c[:, 1] + half[:, 1]])
    d_boxes = torch.from_numpy(np.ascontiguousarray(boxes)).cuda()
    log(f"boxes hit record {name} n={n}")
    st = D.stream()
    order, srt = g._order(1, d_boxes)

    def run(slots):
        counts = D.empty(n, torch.int64)
        hits = D.empty(slots * n, torch.int32) if slots else None
        _native.call("ct_locate_boxes_count_rec", g.tree_ref, D.ptr(d_boxes), n, D.ptr(order), D.ptr(srt),
                     D.ptr(counts), D.ptr(hits), slots, st)
        offsets, total = D.exclusive_scan(counts)
        ii = D.empty(total, torch.int64)
        jj = D.empty(total, torch.int64)
        _native.call("ct_locate_boxes_write_rec", g.tree_ref, D.ptr(d_boxes), n, D.ptr(order), D.ptr(srt),
                     D.ptr(offsets), D.ptr(hits), slots, D.ptr(ii), D.ptr(jj), st)
        return counts.cpu().numpy(), ii.cpu().numpy(), jj.cpu().numpy()

    c0, i0, j0 = run(0)
    assert (c0 > 32).any() and (c0 <= 1).any()
    oi, oj = o.locate_boxes(boxes)
    assert np.array_equal(i0, oi) and np.array_equal(j0, oj)
    for slots in (1, 5, 32):
        c, i, j = run(slots)
        assert np.array_equal(c, c0) and np.array_equal(i, i0) and np.array_equal(j, j0), slots


def test_small_and_degenerate_box_and_segment_batches(trees):
    """Box and segment batches at the edges of the split path, against the oracle: no
    queries, one query (a segment across the mesh and a box over most of it, both split),
    queries that all miss the tree, and non-finite queries among others."""
    name, v, f, g, o = trees
    L = _extent(v)
    lo = v.min(0)
    log(f"small / degenerate batches {name}")
    mid = lo + 0.5 * L
    cases = {
        "empty": (np.zeros((0, 4)), np.zeros((0, 2, 2))),
        "one": (np.array([[lo[0] + 0.05 * L, lo[0] + 0.9 * L, lo[1] + 0.1 * L, lo[1] + 0.8 * L]]),
                np.array([[[lo[0] + 0.01 * L, lo[1] + 0.02 * L], [lo[0] + 0.97 * L, lo[1] + 0.93 * L]]])),
        "outside": (np.array([[lo[0] - 3 * L, lo[0] - 2 * L, lo[1], lo[1] + L]] * 40),
                    np.array([[[lo[0] - 3 * L, lo[1]], [lo[0] - 2 * L, lo[1] + L]]] * 40)),
        "non-finite": (np.array([[np.nan, mid[0], lo[1], mid[1]], [lo[0], np.inf, lo[1], mid[1]],
                                 [lo[0], mid[0], lo[1], mid[1]]] * 10),
                       np.array([[[np.nan, mid[1]], [mid[0], mid[1]]], [[lo[0], lo[1]], [np.inf, mid[1]]],
                                 [[lo[0], lo[1]], [mid[0], mid[1]]]] * 10)),
    }
    for label, (boxes, segs) in cases.items():
        gi, gj = g.locate_boxes(boxes)
        oi, oj = o.locate_boxes(boxes)
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj), label
        gi, gj, gxy = g.intersect_edges(segs)
        oi, oj, oxy = o.intersect_edges(segs)
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj), label
        assert_close(gxy.reshape(-1, 4), oxy.reshape(-1, 4), L)
        if label == "one":
            assert len(gi) > 10


@pytest.mark.parametrize("max_parts, leaf_cells, slots", [(2, 2, 4), (7, 4, 0), (16, 1, 64), (16, 16, 1024)])
def test_split_queries_are_exact(trees, max_parts, leaf_cells, slots):
    """Box and segment queries cut into subqueries by subtree (ct_split_queries) give the
    rows of the whole walks, bit for bit and in the same order, and the oracle's: every
    query is split (no size threshold), with several frontier sizes and hit records from
    none to every row, so both the replayed and the re-walked write paths run; queries from a
    fraction of a cell to across the mesh, some outside it, some along mesh edges."""
    name, v, f, g, o = trees
    L = _extent(v)
    lo = v.min(0)
    h = L / np.sqrt(len(f))
    rng = np.random.default_rng(31)
    n = 5000
    c = lo + rng.uniform(-0.05 * L, 1.05 * L, size=(n, 2))
    wh = L * np.exp(rng.uniform(np.log(1e-4), np.log(0.3), size=(n, 2)))
    boxes = np.column_stack([c[:, 0] - wh[:, 0], c[:, 0] + wh[:, 0], c[:, 1] - wh[:, 1], c[:, 1] + wh[:, 1]])
    length = L * np.exp(rng.uniform(np.log(1e-4), 0.0, size=(n, 1)))
    ang = rng.uniform(0, 2 * np.pi, size=(n, 1))
    seg = np.stack([c, c + length * np.concatenate([np.cos(ang), np.sin(ang)], axis=1)], axis=1)
    fv = f[:, :3]
    along = np.stack([v[fv[:300, 0]], v[fv[:300, 1]]], axis=1)
    seg = np.concatenate([seg, along, along[:, ::-1] + np.array([0.0, 0.3 * h])])
    log(f"split queries {name} parts={max_parts} leaf={leaf_cells} slots={slots}")

    def run(split: bool):
        g.ORDER_MIN = 0  # the split path runs on ordered batches
        g.SPLIT_MIN_CELLS = 0 if split else 10**9
        g.SPLIT_MAX_PARTS = max_parts if split else 1
        g.SPLIT_LEAF_CELLS = g.EDGE_SPLIT_LEAF_CELLS = leaf_cells
        g.BOX_HIT_SLOTS = g.EDGE_HIT_SLOTS = max(slots, 1)
        g.BOX_HIT_BUDGET = g.EDGE_HIT_BUDGET = (slots * 10**6) if slots >= 4 else 0
        # clipped-segment record: as many slots as cells, none, or fewer than the cells (the
        # write pass then copies, clips recorded cells again, or walks again, per subquery)
        g.EDGE_XY_SLOTS = {4: 4, 0: 0, 64: 16, 1024: 1024}[slots]
        try:
            return g.locate_boxes(boxes), g.intersect_edges(seg)
        finally:
            for attr in ("ORDER_MIN", "SPLIT_MIN_CELLS", "SPLIT_MAX_PARTS", "SPLIT_LEAF_CELLS",
                         "EDGE_SPLIT_LEAF_CELLS", "BOX_HIT_SLOTS",
                         "EDGE_HIT_SLOTS", "BOX_HIT_BUDGET", "EDGE_HIT_BUDGET", "EDGE_XY_SLOTS"):
                delattr(g, attr)

    (bi, bj), (ei, ej, exy) = run(True)
    (ui, uj), (fi, fj, fxy) = run(False)
    assert np.array_equal(bi, ui) and np.array_equal(bj, uj)
    assert np.array_equal(ei, fi) and np.array_equal(ej, fj)
    assert np.array_equal(exy.view(np.int64), fxy.view(np.int64))
    oi, oj = o.locate_boxes(boxes)
    assert np.array_equal(bi, oi) and np.array_equal(bj, oj)
    oi, oj, oxy = o.intersect_edges(seg)
    assert np.array_equal(ei, oi) and np.array_equal(ej, oj)
    assert_close(exy, oxy, L)
    # the split really happened: more subqueries than queries
    from numba_celltree_amd import _device as D

    import torch

    from numba_celltree_amd import _native

    d = torch.from_numpy(boxes).cuda()
    order = D.query_order(1, d, g._bbox)
    g.SPLIT_MIN_CELLS, g.SPLIT_MAX_PARTS, g.SPLIT_LEAF_CELLS = 0, max_parts, leaf_cells
    try:
        subs, part_offsets, m = g._split(1, d, order)
    finally:
        del g.SPLIT_MIN_CELLS, g.SPLIT_MAX_PARTS, g.SPLIT_LEAF_CELLS
    assert m > n
    # without the frontier buffer the write call computes the frontiers again: same table
    again = D.empty((m, 2), torch.int32)
    _native.call("ct_split_queries", g.tree_ref, 1, D.ptr(d), n, D.ptr(order), 0, leaf_cells, max_parts, None,
                 D.ptr(part_offsets), D.ptr(again), None, D.stream())
    assert torch.equal(again, subs[:m])


def test_split_edge_tree_is_exact():
    """Segment-edge queries of an EdgeCellTree2d split into subqueries equal the whole walks
    and the oracle."""
    import numba_celltree_amd as M

    v, f = S.jittered_grid(81, 100.0, 0.25, 1.0, 8)
    e = np.unique(np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1), axis=0)
    g = M.EdgeCellTree2d(v, e)
    o = R.OracleEdgeCellTree2d(v, e)
    rng = np.random.default_rng(5)
    a = rng.uniform(-5, 105, size=(6000, 2))
    b = a + 60 * np.exp(rng.uniform(np.log(1e-3), 0, size=(6000, 1))) * rng.normal(size=(6000, 2))
    seg = np.stack([a, b], axis=1)
    g.ORDER_MIN, g.SPLIT_MIN_CELLS, g.SPLIT_MAX_PARTS, g.EDGE_SPLIT_LEAF_CELLS = 0, 0, 12, 2
    gi, gj, gxy = g.intersect_edges(seg)
    oi, oj, oxy = o.intersect_edges(seg)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert_close(gxy, oxy, 100.0)
