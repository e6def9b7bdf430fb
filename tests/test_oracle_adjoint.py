"""The oracle's overlap-area adjoint (ref_pair_area_gradient) against the mathematics.

The regrid backward's roofline divides the adjoint's FP64 operations by the kernel time, and
the GPU adjoint is checked against this restatement, so both its values and its operation
counter are pinned here, independently of the GPU code:

* values: the derivative of the oracle's own clip area (area_of_intersection,
  sutherland_hodgman.py:84-148) by central finite differences, vertex coordinate by vertex
  coordinate, on random overlapping convex pairs (triangles to hexagons, both roles); a
  dropped term, a wrong sign or a swapped end weight fails by orders of magnitude;
* the identities it must satisfy: translating both polygons together leaves the area
  unchanged (the gradients sum to zero), and scaling both by s multiplies it by s^2 (the
  gradients dotted with the vertices sum to twice the area);
* the operation counter: a closed form for two offset axis-parallel squares, counted by
  hand from the restated operations (see test_op_count_closed_form).
"""

import numpy as np
import pytest

from oracle import refcpu as R


def _convex(rng, n, center, radius):
    """n points on an ellipse at sorted angles: a convex polygon, counter-clockwise."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    sx, sy = radius * rng.uniform(0.5, 1.0, 2)
    return np.column_stack([center[0] + sx * np.cos(ang), center[1] + sy * np.sin(ang)])


def _pairs(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        na, nb = rng.integers(3, 7, 2)
        a = _convex(rng, na, rng.uniform(-1, 1, 2), rng.uniform(0.5, 2.0))
        b = _convex(rng, nb, rng.uniform(-1, 1, 2), rng.uniform(0.5, 2.0))
        out.append((a, b))
    return out


def _as_faces(a, b):
    va = a.copy()
    vb = b.copy()
    fa = np.full((1, 8), -1, dtype=np.int64)
    fb = np.full((1, 8), -1, dtype=np.int64)
    fa[0, : len(a)] = np.arange(len(a))
    fb[0, : len(b)] = np.arange(len(b))
    return va, vb, fa, fb


def _area(a, b):
    va, vb, fa, fb = _as_faces(a, b)
    return float(R.area_of_intersection(va, vb, fa, fb, np.zeros(1), np.zeros(1))[0])


def _grad(a, b, g=1.0):
    va, vb, fa, fb = _as_faces(a, b)
    ga, gb, ops = R.pair_area_gradient(va, vb, fa, fb, np.zeros(1), np.zeros(1), g)
    return ga[0, : len(a)], gb[0, : len(b)], int(ops[0])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_adjoint_matches_finite_differences(seed):
    checked = 0
    for a, b in _pairs(seed, 60):
        A = _area(a, b)
        if A < 1e-3:
            continue
        ga, gb, _ = _grad(a, b)
        h = 1e-6
        for poly, grad, which in ((a, ga, 0), (b, gb, 1)):
            fd = np.zeros_like(poly)
            for i in range(len(poly)):
                for k in range(2):
                    p = poly.copy()
                    m = poly.copy()
                    p[i, k] += h
                    m[i, k] -= h
                    args_p = (p, b) if which == 0 else (a, p)
                    args_m = (m, b) if which == 0 else (a, m)
                    fd[i, k] = (_area(*args_p) - _area(*args_m)) / (2 * h)
            scale = max(1.0, np.abs(fd).max())
            np.testing.assert_allclose(grad, fd, rtol=0, atol=1e-6 * scale)
        checked += 1
    assert checked >= 20


@pytest.mark.parametrize("seed", [4, 5])
def test_adjoint_identities(seed):
    for a, b in _pairs(seed, 40):
        A = _area(a, b)
        if A < 1e-3:
            continue
        ga, gb, _ = _grad(a, b, 1.0)
        total = ga.sum(0) + gb.sum(0)  # translation invariance
        np.testing.assert_allclose(total, 0.0, atol=1e-12 * max(1.0, np.abs(ga).max()))
        euler = (ga * a).sum() + (gb * b).sum()  # A is homogeneous of degree 2
        np.testing.assert_allclose(euler, 2.0 * A, rtol=1e-10)
        ga2, gb2, _ = _grad(a, b, 2.5)  # linear in the scale gA (up to the products' rounding)
        np.testing.assert_allclose(ga2, 2.5 * ga, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(gb2, 2.5 * gb, rtol=1e-13, atol=1e-15)


def test_op_count_closed_form():
    """Subject [0,2]^2, clipper [1,3]^2, both CCW from the lower-left corner.

    Per edge: its direction (2), then per half-plane of the other square the edge u (2), c0
    (5), c1 (3), plus 2 for the bound comparison where the edge crosses the line (c1 != 0:
    two of the four half-planes; the other two are parallel, c1 == 0, and not collinear),
    then the two end divisions (2): 2 + 4 * 10 + 2 * 2 + 2 = 48.  An edge with a non-empty
    piece adds the weights (4 + 2) and the two end gradients (4 + 4), and 4 adds into the
    vertex sums: 48 + 14 + 4 = 66.  The subject's right and top edges and the clipper's
    bottom and left edges have pieces: 4 * 48 + 4 * 66 = 456.

    Values, by hand from the overlap [1,2]^2: the subject's right edge (2,0) -> (2,2) lies in
    the clipper for t in [1/2, 1], weights w1 = (1 - 1/4) / 2 = 3/8, w0 = 1/2 - 3/8 = 1/8,
    outward normal times length (2, 0): (1/4, 0) to (2,0) and (3/4, 0) to (2,2); the top edge
    (2,2) -> (0,2), t in [0, 1/2]: (0, 3/4) to (2,2) and (0, 1/4) to (0,2).  The clipper's
    bottom and left edges mirror this with inward signs."""
    a = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    b = a + 1.0
    ga, gb, ops = _grad(a, b)
    assert ops == 456
    assert _area(a, b) == 1.0
    np.testing.assert_array_equal(ga, [[0, 0], [0.25, 0], [0.75, 0.75], [0, 0.25]])
    np.testing.assert_array_equal(gb, [[-0.75, -0.75], [0, -0.25], [0, 0], [-0.25, 0]])


def test_zero_scale_counts_nothing():
    a = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    ga, gb, ops = _grad(a, a + 1.0, 0.0)
    assert ops == 0 and not ga.any() and not gb.any()
