"""The oracle's FP64 operation counter (ref_face_pair_flops) against an independent count.

The roofline of the clip kernels divides the FP64 operations the reference executes by the
kernel time, so the counter has to be right.  Here the reference's separating_axis.py:17-75,
sutherland_hodgman.py:56-148 and geometry_utils.py:82-95 are restated in plain Python over a
float type that counts its own adds, subtracts, multiplies and divides (compares, abs, min /
max and negation are free, as in the C++ counter).  Both must agree pair by pair -- on the
count, the SAT verdict and the area bit for bit -- on candidate pairs of a small version of
the overlap workload, on touching and shared-edge pairs, and on polygons with repeated
vertices (the zero-length-edge branches).  A dropped or doubled term in either counter
fails.  The committed full-size counts (tests/golden/flops_overlap.json) are recomputed
here from the same generator.
"""

import json
import os

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "flops_overlap.json")

FLOAT_MAX = np.finfo(np.float64).max
FLOAT_MIN = np.finfo(np.float64).min


class Ops:
    n = 0


class CF:
    """A float that counts + - * / (one each)."""

    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v)

    def _b(self, o):
        return o.v if isinstance(o, CF) else float(o)

    def __add__(self, o):
        Ops.n += 1
        return CF(self.v + self._b(o))

    def __radd__(self, o):
        Ops.n += 1
        return CF(self._b(o) + self.v)

    def __sub__(self, o):
        Ops.n += 1
        return CF(self.v - self._b(o))

    def __rsub__(self, o):
        Ops.n += 1
        return CF(self._b(o) - self.v)

    def __mul__(self, o):
        Ops.n += 1
        return CF(self.v * self._b(o))

    def __rmul__(self, o):
        Ops.n += 1
        return CF(self._b(o) * self.v)

    def __truediv__(self, o):
        Ops.n += 1
        return CF(self.v / self._b(o))

    def __neg__(self):
        return CF(-self.v)

    def __abs__(self):
        return CF(abs(self.v))

    def __lt__(self, o):
        return self.v < self._b(o)

    def __gt__(self, o):
        return self.v > self._b(o)

    def __le__(self, o):
        return self.v <= self._b(o)

    def __ge__(self, o):
        return self.v >= self._b(o)

    def __eq__(self, o):
        return self.v == self._b(o)

    def __ne__(self, o):
        return self.v != self._b(o)


# ------------------------------------------------- separating_axis.py:17-55
def extrema_projected(norm, poly):
    mn, mx = CF(FLOAT_MAX), CF(FLOAT_MIN)
    for x, y in poly:
        proj = x * norm[0] + y * norm[1]  # dot_product(as_point(polygon[i]), norm)
        mn = proj if proj < mn else mn
        mx = proj if proj > mx else mx
    return mn, mx


def is_separating_axis(norm, a, b):
    mina, maxa = extrema_projected(norm, a)
    minb, maxb = extrema_projected(norm, b)
    return not (maxa > minb and maxb > mina)


def separating_axes(a, b):
    p = a[-1]
    for q in a:
        norm = (p[1] - q[1], q[0] - p[0])
        p = q
        if norm[0] == 0.0 and norm[1] == 0.0:
            continue
        if is_separating_axis(norm, a, b):
            return False
    return True


# ---------------------------------------------- sutherland_hodgman.py:56-148
def inside(p, r, U):
    return U[0] * (p[1] - r[1]) > U[1] * (p[0] - r[0])


def intersection(a, V, r, N):
    W = (r[0] - a[0], r[1] - a[1])
    nw = N[0] * W[0] + N[1] * W[1]
    nv = N[0] * V[0] + N[1] * V[1]
    if nv != 0:
        t = nw / nv
        return True, (a[0] + t * V[0], a[1] + t * V[1])
    return False, None


def polygon_area(poly):  # geometry_utils.py:82-95
    area = CF(0.0)
    a, b = poly[0], poly[1]
    U = (b[0] - a[0], b[1] - a[1])
    for c in poly[2:]:
        V = (a[0] - c[0], a[1] - c[1])
        area = area + abs(U[0] * V[1] - U[1] * V[0])
        U = V
    return 0.5 * area


def clip_area(polygon, clipper):
    output = list(polygon)
    r = clipper[-1]
    for s in clipper:
        U = (s[0] - r[0], s[1] - r[1])
        if U[0] == 0 and U[1] == 0:
            continue
        N = (-U[1], U[0])
        subject = output
        output = []
        a = subject[-1]
        a_inside = inside(a, r, U)
        for b in subject:
            V = (b[0] - a[0], b[1] - a[1])
            if V[0] == 0 and V[1] == 0:
                continue
            b_inside = inside(b, r, U)
            if b_inside:
                if not a_inside:
                    ok, pt = intersection(a, V, r, N)
                    if ok:
                        output.append(pt)
                output.append(b)
            elif a_inside:
                ok, pt = intersection(a, V, r, N)
                if ok:
                    output.append(pt)
                else:
                    b_inside = True
                    output.append(b)
            a = b
            a_inside = b_inside
        if len(output) < 3:
            return CF(0.0)
        r = s
    return polygon_area(output)


def _poly(vertices, row):
    n = len(row)
    for i in range(3, len(row)):
        if row[i] == -1:
            n = i
            break
    return [(CF(vertices[k, 0]), CF(vertices[k, 1])) for k in row[:n]]


def python_counts(va, vb, fa, fb, ia, ib):
    out = []
    for i, j in zip(ia, ib):
        a, b = _poly(va, fa[i]), _poly(vb, fb[j])
        Ops.n = 0
        ok = separating_axes(a, b) and separating_axes(b, a)
        sat = Ops.n
        Ops.n = 0
        area = clip_area(a, b).v if ok else 0.0
        out.append((sat, ok, Ops.n, area))
    return out


def _check(va, vb, fa, fb, ia, ib):
    sat, passed, clip, area = R.face_pair_flops(va, vb, fa, fb, ia, ib)
    want = python_counts(va, vb, np.asarray(fa), np.asarray(fb), ia, ib)
    for k, (s, ok, c, a) in enumerate(want):
        assert (sat[k], bool(passed[k]), clip[k]) == (s, ok, c), (k, sat[k], passed[k], clip[k], s, ok, c)
        assert area[k] == a or (np.isnan(area[k]) and np.isnan(a)), (k, area[k], a)
    # the counter's areas are the oracle's own areas
    kept = passed.astype(bool)
    np.testing.assert_array_equal(
        area[kept], R.area_of_intersection(va, vb, fa, fb, np.asarray(ia)[kept], np.asarray(ib)[kept])
    )
    return sat, passed, clip


def test_counts_match_python_restatement_on_overlap_pairs():
    # A small version of the overlap workload: jittered triangles x hex-Voronoi cells.
    sv, sf = S.jittered_grid(41, 100.0, 0.25, 1.0, 0)
    tv, tf = S.hex_voronoi(12, 12, 100.0, 0.1, 7)
    tf = tf.astype(np.int64)
    R.counter_clockwise(tv, tf)
    tree = R.OracleCellTree2d(sv, sf, -1)
    bb = R.build_face_bboxes(tf, tv)
    ii, jj = tree.locate_boxes(bb)
    sel = np.random.default_rng(0).choice(len(ii), size=min(len(ii), 4000), replace=False)
    sat, passed, clip = _check(tv, sv, tf, tree.faces, ii[sel], jj[sel])
    assert passed.any() and (~passed).any() and (clip > 0).any()


def test_counts_on_degenerate_pairs():
    """Touching squares, shared edges, nested and identical polygons, repeated vertices
    (zero-length clipper and subject edges), a parallel-edge intersection, and the shapely
    golden pairs of the reference's Sutherland-Hodgman tests."""
    from tests.golden import cases as C

    v = np.array(
        [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0], [2.0, 0.0], [2.0, 1.0], [0.5, 0.5], [0.25, 0.25],
         [0.75, 0.25], [3.0, 3.0], [4.0, 3.0], [4.0, 4.0]]
    )
    f = np.array(
        [[0, 1, 2, 3], [1, 4, 5, 2], [7, 8, 6, -1], [0, 1, 1, 2], [0, 0, 1, 2], [9, 10, 11, -1], [0, 4, 5, 3],
         [0, 1, 2, 3]]
    )
    n = len(f)
    ia, ib = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    _check(v, v, f, f, ia.ravel(), ib.ravel())
    va = C.SH_A.reshape(-1, 2)
    vb = C.SH_B.reshape(-1, 2)
    fa = np.arange(len(va)).reshape(-1, 3)
    fb = np.arange(len(vb)).reshape(-1, 3)
    idx = np.arange(len(fa))
    _check(va, vb, fa, fb, idx, idx)


def test_hand_count_of_one_pair():
    """A unit square (subject) fully inside a larger square (clipper): SAT tests 4 + 4 edge
    normals, none separating (each: 2 subs for the normal, 3 ops per projected vertex of
    both polygons = 2 + 24); the clip keeps all 4 vertices through 4 clip edges (each: 2
    for U, 4 for a_inside, then per vertex 2 for V and 4 for b_inside = 2 + 4 + 24) and the
    area of 4 vertices costs 2 + 2 * 6 + 1."""
    v = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0], [-1.0, -1.0], [2.0, -1.0], [2.0, 2.0],
                  [-1.0, 2.0]])
    f = np.array([[0, 1, 2, 3], [4, 5, 6, 7]])
    sat, passed, clip, area = R.face_pair_flops(v, v, f, f, np.array([0]), np.array([1]))
    assert passed[0] and area[0] == 1.0
    assert sat[0] == 8 * (2 + 24)
    assert clip[0] == 4 * (2 + 4 + 24) + (2 + 2 * 6 + 1)


def test_fixture_counts_are_reproducible():
    """The committed full-size counts equal a fresh count from the generator (CPU oracle,
    a few seconds)."""
    from tests.golden import make_flops_fixture as mk

    want = json.load(open(FIXTURE))
    got = mk.overlap_counts()
    for key in ("candidates", "sat_passed", "kept", "sat_flops", "clip_flops", "clip_flops_kept"):
        assert got[key] == want["intersect_faces"][key], key
    got = mk.regrid_counts()
    for key in ("kept", "clip_flops_kept", "backward_pair_flops", "backward_ga_flops_16f"):
        assert got[key] == want["regrid"][key], key
