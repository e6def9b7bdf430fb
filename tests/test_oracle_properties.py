"""Pins of the CPU oracle beyond the reference's golden vectors.

Each check is fixed by the mathematics rather than by the oracle's own formulas, and is
chosen so that a plausible slip (a dropped term, a swapped operand, a wrong index or push
order) fails it:

* tree build: structural invariants (bounds validity, permutation, leaf sizes, split
  planes equal to the extreme child bounds, DFS child numbering);
* locate_boxes: the result set equals brute force over all (box, cell) pairs;
* locate_points: equals a brute-force exact-orientation search for points well away from
  every edge, and every located point lies in its cell;
* Sutherland-Hodgman: areas equal an exact rational clip (fractions.Fraction) to 1e-12
  relative, and the overlaps of a target with all cells sum to the target's area;
* Cyrus-Beck: clipped pieces of a segment crossing a conforming mesh sum to its length;
* Wachspress / triangle weights: partition of unity, linear precision, Lagrange property.
"""

from fractions import Fraction

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R


@pytest.fixture(scope="module")
def small_mixed():
    v, f = S.mixed_mesh_1m(nv=41, seed=11)
    return v, f, R.OracleCellTree2d(v, f, -1)


def polys_of(v, f):
    return [v[row[row >= 0]] for row in f]


# ------------------------------------------------------------------ tree build
@pytest.mark.parametrize("nb,cpl", [(4, 2), (2, 1), (8, 3)])
def test_tree_invariants(nb, cpl):
    v, f = S.mixed_mesh_1m(nv=61, seed=2)
    t = R.OracleCellTree2d(v, f, -1, n_buckets=nb, cells_per_leaf=cpl)
    nodes, bbi, bb = t.nodes, t.bb_indices, t.bb_coords
    assert np.array_equal(np.sort(bbi), np.arange(len(f)))
    assert t.validate_node_bounds().all()
    internal = nodes["child"] >= 0
    # DFS numbering: the k-th internal node in preorder has children 1 + 2k, 2 + 2k.
    order = []
    stack = [0]
    while stack:
        x = stack.pop()
        if nodes[x]["child"] >= 0:
            order.append(x)
            stack.append(nodes[x]["child"] + 1)
            stack.append(nodes[x]["child"])
    assert np.array_equal(nodes["child"][order], 1 + 2 * np.arange(len(order)))
    assert len(nodes) == 1 + 2 * internal.sum()
    for x in np.flatnonzero(internal):
        nd = nodes[x]
        l, r = nodes[nd["child"]], nodes[nd["child"] + 1]
        d = 1 if nd["dim"] else 0
        assert l["ptr"] == nd["ptr"] and r["ptr"] == l["ptr"] + l["size"]
        assert l["size"] + r["size"] == nd["size"]
        lcells = bbi[l["ptr"]: l["ptr"] + l["size"]]
        rcells = bbi[r["ptr"]: r["ptr"] + r["size"]]
        if cpl == 1 and nd["size"] == 2:
            # creation.py:312-320: the node keeps the range of both cells
            both = np.concatenate([lcells, rcells])
            assert nd["Lmax"] == bb[both, 2 * d + 1].max()
            assert nd["Rmin"] == bb[both, 2 * d].min()
        else:
            assert nd["Lmax"] == bb[lcells, 2 * d + 1].max()
            assert nd["Rmin"] == bb[rcells, 2 * d].min()
    leaves = nodes[~internal]
    degenerate = leaves["Lmax"] == -1.0
    assert (leaves["size"][degenerate] >= 0).all()
    big = leaves["size"] > cpl
    # a leaf larger than cells_per_leaf only arises from the twice-failed bucketing
    assert np.all(~big | ((leaves["Lmax"] == -1.0) & (leaves["Rmin"] == -1.0)))


# ------------------------------------------------------------------ boxes
def test_locate_boxes_equals_brute_force(small_mixed):
    v, f, t = small_mixed
    bx = S.boxes(3000, 1000.0, 25.0, 0.1, 4.0, 3)
    ii, jj = t.locate_boxes(bx)
    bb = t.bb_coords
    a = bx[:, None, :]
    b = bb[None, :, :]
    hit = (a[..., 0] < b[..., 1]) & (b[..., 0] < a[..., 1]) & (a[..., 2] < b[..., 3]) & (b[..., 2] < a[..., 3])
    bi, bj = np.nonzero(hit)
    got = set(zip(ii.tolist(), jj.tolist()))
    want = set(zip(bi.tolist(), bj.tolist()))
    assert got == want
    assert np.all(np.diff(ii) >= 0)  # grouped in box order


# ------------------------------------------------------------------ points
def _orient(a, b, p):
    return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0])


def test_locate_points_brute_force(small_mixed):
    """Exact-orientation brute force for points farther than 1e-6 h from every cell edge."""
    v, f, t = small_mixed
    ccw = t.faces  # counter-clockwise rows
    rng = np.random.default_rng(5)
    pts = rng.uniform(-20, 1020, size=(4000, 2))
    got = t.locate_points(pts)
    h = 1000.0 / 40
    # pad triangles by repeating their first vertex: the extra edge has zero length
    nv = (ccw >= 0).sum(1)
    rows = np.where(ccw >= 0, ccw, ccw[:, :1])
    P = v[rows]  # (F, 4, 2)
    A = P
    B = np.roll(P, -1, axis=1)
    ex = B - A
    ln = np.hypot(ex[..., 0], ex[..., 1])
    real = ln > 0
    for chunk in np.array_split(np.arange(len(pts)), 8):
        p = pts[chunk][:, None, None, :]
        o = ex[None, ..., 0] * (p[..., 1] - A[None, ..., 1]) - ex[None, ..., 1] * (p[..., 0] - A[None, ..., 0])
        inside = np.all((o > 0) | ~real[None], axis=2)
        dist = np.where(real[None], np.abs(o) / np.where(real, ln, 1.0)[None], np.inf)
        near = np.any((dist.min(axis=2) < 1e-6 * h) & np.all((o > -1e-6 * h * ln[None]) | ~real[None], axis=2), axis=1)
        for q, k in enumerate(chunk):
            if near[q]:
                continue
            hits = np.flatnonzero(inside[q])
            assert len(hits) <= 1
            assert got[k] == (hits[0] if len(hits) else -1), (pts[k], got[k], hits)
    assert (nv >= 3).all()


# ------------------------------------------------------------------ clipping
def _exact_clip_area(subject, clipper):
    """Convex polygon intersection area in exact rational arithmetic."""
    poly = [(Fraction(x), Fraction(y)) for x, y in subject]
    cl = [(Fraction(x), Fraction(y)) for x, y in clipper]
    for i in range(len(cl)):
        r, s = cl[i - 1], cl[i]
        ux, uy = s[0] - r[0], s[1] - r[1]
        if ux == 0 and uy == 0:
            continue
        side = lambda p: ux * (p[1] - r[1]) - uy * (p[0] - r[0])  # noqa: E731
        out = []
        for j in range(len(poly)):
            a, b = poly[j - 1], poly[j]
            sa, sb = side(a), side(b)
            if sb >= 0:
                if sa < 0:
                    t = sa / (sa - sb)
                    out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
                out.append(b)
            elif sa >= 0:
                t = sa / (sa - sb)
                out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
        poly = out
        if len(poly) < 3:
            return Fraction(0)
    area = Fraction(0)
    for i in range(len(poly)):
        area += poly[i - 1][0] * poly[i][1] - poly[i][0] * poly[i - 1][1]
    return abs(area) / 2


def _random_convex(rng, n, center, radius):
    """CCW convex polygon: n points on a circle, in angle order."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    return np.column_stack([center[0] + radius * np.cos(ang), center[1] + radius * np.sin(ang)])


def test_sutherland_hodgman_vs_exact_rational():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(300):
        a = _random_convex(rng, rng.integers(3, 7), rng.uniform(0, 1, 2), rng.uniform(0.2, 0.6))
        b = _random_convex(rng, rng.integers(3, 7), rng.uniform(0, 1, 2), rng.uniform(0.2, 0.6))
        got = R.polygon_polygon_clip_area(a, b)
        want = float(_exact_clip_area(a, b))
        assert abs(got - want) <= 1e-12 * max(want, 1e-3), (got, want)
        checked += want > 0
    assert checked > 100


def test_overlap_areas_conserve_target_area(small_mixed):
    v, f, t = small_mixed
    tv, tf = S.hex_voronoi(12, 12, 1000.0, 0.1, 4)
    ii, jj, area = t.intersect_faces(tv, tf, -1)
    target_area = np.array([R.polygon_area(tv[row[row >= 0]]) for row in tf])
    sums = np.bincount(ii, weights=area, minlength=len(tf))
    assert np.allclose(sums, target_area, rtol=1e-12, atol=0)


def test_clipped_lengths_sum_to_segment_length(small_mixed):
    v, f, t = small_mixed
    rng = np.random.default_rng(8)
    a = rng.uniform(100, 900, size=(500, 2))
    th = rng.uniform(0, 2 * np.pi, 500)
    seg = np.stack([a, a + 60.0 * np.column_stack([np.cos(th), np.sin(th)])], axis=1)
    ii, jj, xy = t.intersect_edges(seg)
    lengths = np.linalg.norm(xy[:, 1] - xy[:, 0], axis=1)
    tot = np.bincount(ii, weights=lengths, minlength=len(seg))
    # a segment may touch a vertex or run along an edge; those add zero-length rows or
    # duplicate an edge-collinear piece, so compare with a one-piece slack
    assert np.all(tot >= 60.0 * (1 - 1e-12))
    assert np.median(np.abs(tot - 60.0)) < 1e-9


# ------------------------------------------------------------------ barycentric
def test_barycentric_identities(small_mixed):
    v, f, t = small_mixed
    rng = np.random.default_rng(9)
    pts = rng.uniform(0, 1000, size=(5000, 2))
    fi, w = t.compute_barycentric_weights(pts)
    ok = fi >= 0
    assert ok.mean() > 0.99
    faces = t.faces[fi[ok]]
    verts = np.where(faces[..., None] >= 0, v[np.maximum(faces, 0)], 0.0)
    assert np.allclose(w[ok].sum(1), 1.0, atol=1e-12)
    recon = (w[ok][..., None] * verts).sum(1)
    assert np.allclose(recon, pts[ok], rtol=0, atol=1e-9)
    assert np.all(w[ok][faces < 0] == 0.0)
    # Lagrange property at the vertices of a few cells
    for k in range(20):
        row = t.faces[k]
        row = row[row >= 0]
        for m, vi in enumerate(row):
            wk = R.wachspress_weights(v[row], v[vi], 4, 1e-9)
            e = np.zeros(4)
            e[m] = 1.0
            assert np.allclose(wk, e, atol=1e-12)
    tri = np.array([[0.0, 0.0], [3.0, 0.5], [1.0, 2.0]])
    p = np.array([1.2, 0.8])
    wt = R.triangle_weights(tri, p)
    assert np.isclose(wt.sum(), 1.0) and np.allclose(wt @ tri, p)
