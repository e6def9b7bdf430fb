"""The packed-leaf cell test's rewrite of point_in_polygon_or_on_edge, checked on the CPU.

The device kernel tests every packed cell as a fixed-length row padded with its last vertex,
without the reference's repeated-vertex skip and without its early return (on-edge hits are
OR-ed with the crossing parity).  This restates that loop in Python and compares it with the
oracle's point_in_polygon_or_on_edge (geometry_utils.py:195-223) on rows with repeated
vertices, fill-padded triangles, and points on vertices, on edges, a hair off edges and at
random.  The kernel itself is compared with the oracle in test_gpu_parity.py.
"""

import numpy as np

from oracle import refcpu as R


def _within_perp(A, wx, wy, tol):
    return (A * A) < ((tol * (wx * wx + wy * wy)) * tol)


def _in_bounds(p, a, b):
    xmin, xmax = (a[0], b[0]) if a[0] < b[0] else (b[0], a[0])
    ymin, ymax = (a[1], b[1]) if a[1] < b[1] else (b[1], a[1])
    if abs(xmax - xmin) >= abs(ymax - ymin):
        return xmin <= p[0] <= xmax
    return ymin <= p[1] <= ymax


def _padded_cell(p, row, tol):
    v0 = row[-1]
    ux, uy = v0[0] - p[0], v0[1] - p[1]
    c = on = False
    for v1 in row:
        vx, vy = v1[0] - p[0], v1[1] - p[1]
        A = ux * vy - uy * vx
        wx, wy = v1[0] - v0[0], v1[1] - v0[1]
        if _within_perp(A, wx, wy, tol):
            on = on or _in_bounds(p, v0, v1)
        straddle = (v0[1] > p[1]) != (v1[1] > p[1])
        with np.errstate(all="ignore"):
            xi = np.float64(wx) * np.float64(p[1] - v0[1]) / np.float64(wy) + v0[0]
        c = c != (straddle and p[0] < xi)
        v0, ux, uy = v1, vx, vy
    return on or c


def _polygons(rng):
    g = np.stack(np.meshgrid(np.arange(5.0), np.arange(5.0)), -1).reshape(-1, 2)
    v = g + rng.uniform(-0.2, 0.2, size=g.shape)
    for j in range(4):
        for i in range(4):
            a = j * 5 + i
            b, c, d = a + 1, a + 6, a + 5
            for idx in ([a, b, c, d], [a, b, c], [a, b, b, c], [a, a, b, c], [c, d, a, c], [a, b, c, c, d]):
                yield [tuple(v[k]) for k in idx]


def test_padded_rows_match_reference_loop():
    rng = np.random.default_rng(0)
    checked = 0
    for poly in _polygons(rng):
        n = len(poly)
        row = poly + [poly[-1]] * (8 - n)
        pts = [np.array(q) for q in poly]
        pts += [0.5 * (np.array(poly[i]) + np.array(poly[(i + 1) % n])) for i in range(n)]
        pts += [q + rng.normal(0.0, 1e-10, 2) for q in pts[: 2 * n]]
        pts += [np.array(poly[0]) + rng.uniform(-1.0, 1.0, 2) for _ in range(10)]
        for p in pts:
            p = (float(p[0]), float(p[1]))
            for tol in (1e-12, 1e-9, 0.05):
                want = bool(R.point_in_polygon_or_on_edge(p, np.array(poly), tol))
                assert _padded_cell(p, row, tol) == want, (poly, p, tol)
                checked += 1
    assert checked > 5000
