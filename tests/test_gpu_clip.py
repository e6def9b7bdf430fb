"""The Sutherland-Hodgman clip (area_of_intersection, box_area_of_intersection; the kernels of
intersect_faces and the regrid forward) against the oracle on shapes the meshes of the other
tests rarely produce.

The clip keeps its input polygon in registers and appends each stage's output to the lane's
row in LDS; the vertex classes select the row capacity (subject + clipper <= 16 slots) or the
plain indexed form beyond that.  The pairs here cover, for triangles to 12-gons on both sides:
repeated vertices (zero-length edges, skipped by both the subject loop and the clip-edge loop),
three collinear vertices, identical polygons (every edge on a clip line: the parallel rule of
sutherland_hodgman.py:133-137), polygons sharing one edge from outside (area 0), nested
polygons, disjoint ones, and random overlapping convex pairs; boxes against faces likewise.
Areas must equal the oracle's bit for bit (the same operations in the same order, no
contraction).
"""

import sys

import numpy as np
import pytest

from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def _ngon(n, cx, cy, r, rot=0.0, sx=1.0, sy=1.0):
    a = rot + 2 * np.pi * np.arange(n) / n
    return np.column_stack([cx + r * sx * np.cos(a), cy + r * sy * np.sin(a)])


def _pairs(seed):
    """(polygon a, polygon b) lists; every polygon counter-clockwise."""
    rng = np.random.default_rng(seed)
    out = []
    for na in (3, 4, 6, 8, 12):
        for nb in (3, 4, 6, 8, 12):
            a = _ngon(na, 0.0, 0.0, 1.0, rng.uniform(0, 6.3), 1.0, rng.uniform(0.5, 1.0))
            out.append((a, _ngon(nb, rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8), rng.uniform(0.4, 1.5),
                                 rng.uniform(0, 6.3))))  # overlapping
            out.append((a, a.copy()))  # identical
            out.append((a, _ngon(nb, 0.0, 0.0, 0.3, rng.uniform(0, 6.3))))  # b inside a
            out.append((_ngon(na, 0.0, 0.0, 0.3), _ngon(nb, 0.0, 0.0, 2.0, rng.uniform(0, 6.3))))  # a inside b
            out.append((a, _ngon(nb, 5.0, 5.0, 1.0)))  # disjoint
    sq = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    out.append((sq, sq + [1.0, 0.0]))  # shared edge from outside
    out.append((sq, sq + [1.0, 1.0]))  # shared corner
    out.append((sq, sq + [0.5, 0.0]))  # half overlap, two edges on clip lines
    tri = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 2.0]])
    out.append((tri, sq))
    # repeated vertices (zero-length edges) in the subject, the clipper, both
    rep = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0], [0.0, 1.0]])
    out.append((rep, sq + 0.25))
    out.append((sq + 0.25, rep))
    out.append((rep, rep + 0.5))
    # three collinear vertices
    col = np.array([[0.0, 0.0], [0.5, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    out.append((col, sq + [0.3, -0.2]))
    out.append((sq + [0.3, -0.2], col))
    out.extend(_full_rows())
    return out


def _full_rows():
    """Pairs whose clipped polygon fills the clip row past its LDS slots: a regular octagon and
    a square that cuts off every other corner (12 vertices), both ways round, and two octagons
    turned by 22.5 degrees (16 vertices)."""
    octagon = _ngon(8, 0.0, 0.0, 1.0)
    square = np.array([[-0.95, -0.95], [0.95, -0.95], [0.95, 0.95], [-0.95, 0.95]])
    return [(octagon, square), (square, octagon), (octagon, _ngon(8, 0.0, 0.0, 1.0, np.pi / 8))]


def _as_mesh(polys, width):
    verts, faces = [], []
    off = 0
    for p in polys:
        verts.append(p)
        row = np.full(width, -1, dtype=np.int64)
        row[: len(p)] = off + np.arange(len(p))
        faces.append(row)
        off += len(p)
    return np.concatenate(verts), np.array(faces)


@pytest.mark.parametrize("seed", [0, 1])
def test_area_of_intersection_matches_oracle(seed):
    from numba_celltree_amd import algorithms as A

    pairs = _pairs(seed)
    wa = max(len(a) for a, _ in pairs)
    wb = max(len(b) for _, b in pairs)
    for cut_a, cut_b in ((4, 4), (8, 4), (4, 8), (8, 8), (wa, wb)):
        # the same pairs in every vertex class that holds them (faces padded to the class width)
        sel = [k for k, (a, b) in enumerate(pairs) if len(a) <= cut_a and len(b) <= cut_b]
        va, fa = _as_mesh([pairs[k][0] for k in sel], cut_a)
        vb, fb = _as_mesh([pairs[k][1] for k in sel], cut_b)
        i = np.arange(len(sel))
        got = A.area_of_intersection(va, vb, fa, fb, i, i)
        want = R.area_of_intersection(va, vb, fa, fb, i, i)
        diff = np.abs(got - want)
        print(f"[clip] seed {seed} widths {cut_a}x{cut_b}: {len(sel)} pairs, max diff {diff.max():.3e}, "
              f"bit-identical {np.array_equal(got, want)}", file=sys.stderr, flush=True)
        np.testing.assert_array_equal(got, want)  # exact values (a zero of either sign is zero)
        shared_edge = np.array([pairs[k][0].shape == pairs[k][1].shape and
                                np.array_equal(pairs[k][1], pairs[k][0] + [1.0, 0.0]) for k in sel])
        assert shared_edge.sum() == 1 and got[shared_edge][0] == 0.0
    assert len(pairs) > 100


def test_intersect_faces_full_rows_match_oracle():
    """The same 12- and 16-vertex intersections through intersect_faces (the pair kernel with
    its SAT pre-test), several copies of each polygon shifted so every pair class holds some."""
    import numba_celltree_amd as M

    pairs = _full_rows()
    shifts = [np.array([3.0 * i, 0.0]) for i in range(4)]  # copies 3 apart overlap only within a copy
    for wa, wb in ((8, 4), (4, 8), (8, 8)):
        sel = [(a, b) for a, b in pairs if len(a) <= wa and len(b) <= wb]
        va, fa = _as_mesh([a + d for a, _ in sel for d in shifts], wa)
        vb, fb = _as_mesh([b + d for _, b in sel for d in shifts], wb)
        gi, gj, ga = M.CellTree2d(va, fa, -1).intersect_faces(vb, fb, -1)
        oi, oj, oa = R.OracleCellTree2d(va, fa, -1).intersect_faces(vb, fb, -1)
        print(f"[clip] intersect_faces widths {wa}x{wb}: {len(gi)} pairs, max diff "
              f"{np.abs(ga - oa).max():.3e}", file=sys.stderr, flush=True)
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
        assert len(gi) >= len(sel) * len(shifts)
        np.testing.assert_allclose(ga, oa, rtol=1e-12, atol=0.0)  # the parity tests' bar


def test_box_area_of_intersection_matches_oracle():
    from numba_celltree_amd import algorithms as A

    rng = np.random.default_rng(3)
    polys = [_ngon(n, rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.2, 1.5), rng.uniform(0, 6.3))
             for n in (3, 4, 5, 6, 7, 8, 12) for _ in range(20)]
    polys.append(np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]))  # repeated vertex
    for width in (8, 12):
        keep = [p for p in polys if len(p) <= width]
        v, f = _as_mesh(keep, width)
        lo = rng.uniform(-1.5, 1.0, (40, 2))
        boxes = np.column_stack([lo[:, 0], lo[:, 0] + rng.uniform(0.05, 2.0, 40),
                                 lo[:, 1], lo[:, 1] + rng.uniform(0.05, 2.0, 40)])
        boxes[0] = [0.0, 1.0, 0.0, 1.0]  # a box equal to the repeated-vertex square
        ib = np.repeat(np.arange(len(boxes)), len(keep))
        iface = np.tile(np.arange(len(keep)), len(boxes))
        got = A.box_area_of_intersection(boxes, v, f, ib, iface)
        want = R.box_area_of_intersection(boxes, v, f, ib, iface)
        print(f"[clip] boxes width {width}: {len(ib)} pairs, max diff {np.abs(got - want).max():.3e}",
              file=sys.stderr, flush=True)
        np.testing.assert_array_equal(got, want)  # exact values (a zero of either sign is zero)
        assert (want > 0).sum() > 100
