"""The CPU oracle against every golden vector of the reference's own test-suite.

These tests pin the oracle (oracle/refcpu.cpp) before it is trusted as the checker
of the HIP library: the ordered disk-mesh goldens pin the tree build, its DFS
numbering and every traversal order; the algorithm cases pin the clippers,
predicates and barycentric weights, including their degenerate branches.
"""

import types

import numpy as np
import pytest

from oracle import refcpu as R
from tests.golden import cases as C

IMPL = types.SimpleNamespace(
    CellTree2d=R.OracleCellTree2d,
    EdgeCellTree2d=R.OracleEdgeCellTree2d,
    area_of_intersection=R.area_of_intersection,
    box_area_of_intersection=R.box_area_of_intersection,
    polygons_intersect=R.polygons_intersect,
    barycentric_triangle_weights=R.barycentric_triangle_weights,
    barycentric_wachspress_weights=R.barycentric_wachspress_weights,
    counter_clockwise=R.counter_clockwise,
    build_face_bboxes=R.build_face_bboxes,
    build_edge_bboxes=R.build_edge_bboxes,
    sort_intersections_by_edge=R.sort_intersections_by_edge,
    bbox_tree=R.bbox_tree,
    bbox_distances=R.bbox_distances,
    default_tolerance=R.default_tolerance,
)


@pytest.mark.parametrize("case", C.TREE_CASES + C.MODULE_CASES + C.UTIL_CASES, ids=lambda f: f.__name__)
def test_reference_case(case):
    case(IMPL)


def test_shape_and_bounds_errors():  # test_celltree.py:134-164
    with pytest.raises(ValueError):
        R.OracleCellTree2d([(1, 2, 3), (3, 4, 5), (4, 5, 6)], C.faces2, -1)
    with pytest.raises(ValueError):
        R.OracleCellTree2d(C.nodes2, [0, 1, 2, 1, 3, 2], -1)
    with pytest.raises(ValueError):
        R.OracleCellTree2d(C.nodes, C.faces, C.fill_value, cells_per_leaf=-1)
    with pytest.raises(ValueError):
        R.OracleCellTree2d(C.nodes, C.faces, C.fill_value, n_buckets=0)


# ---------------------------------------------------------- test_sutherland_hodgman.py
def test_sh_intersection():  # :93-110
    ok, p = R.sh_intersection([0.0, 0.0], [1.0, 1.0], [1.0, 0.0], [1.0, 1.0])
    assert ok and np.allclose(p, [0.5, 0.5])
    U = np.array([2.0, 1.0]) - np.array([1.0, 0.0])
    ok, _ = R.sh_intersection([0.0, 0.0], [1.0, 1.0], [1.0, 0.0], [-U[1], U[0]])
    assert not ok


def test_clip_area():  # :113-130
    for a, b, e in zip(C.SH_A, C.SH_B, C.SH_EXPECTED):
        assert np.allclose(R.polygon_polygon_clip_area(a, b), e)
    a = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    assert np.allclose(R.polygon_polygon_clip_area(a, a + 2.0), 0)


def test_clip_area_repeated_vertex():  # :133-156
    a = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    assert np.allclose(R.polygon_polygon_clip_area(a, a + 2.0), 0)
    b = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [0.0, 1.0]])
    # Closed form: the triangle (0,0),(1,0),(1,1) clipped by (0,0),(1,0),(0,1) is
    # the triangle (0,0),(1,0),(0.5,0.5) of area 0.25.
    assert np.isclose(R.polygon_polygon_clip_area(a, b), 0.25, rtol=1e-15)


def test_clip_area_epsilon():  # :159-187
    EPS = np.finfo(np.float64).eps
    a = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0]])
    for eps in (EPS, -EPS):
        b = np.array([[-1.0 - eps, -1.0 - eps], [1.0 + eps, -1.0 - eps], [1.0 + eps, 1.0 + eps]])
        assert np.allclose(R.polygon_polygon_clip_area(a, b), 2.0)


def test_polygon_area():  # test_geometry_utils.py:95-127
    assert np.allclose(R.polygon_area([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]), 1.0)
    assert np.allclose(R.polygon_area([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]]), 0.5)
    p = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.5, 2.0], [0.0, 1.0]])
    assert np.allclose(R.polygon_area(p), 1.5)
    assert np.allclose(R.polygon_area(p[::-1]), 1.5)


# ------------------------------------------------------------- test_line_box_clip.py
BOX = np.array([0.0, 2.0, 0.0, 2.0])
POLY = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])


def _cs(a, b, box):
    return R.cohen_sutherland_line_box_clip(a, b, box)


def _cb(a, b, poly):
    return R.cyrus_beck_line_polygon_clip(a, b, poly, 1e-9)


LINE_BOX_CASES = [  # test_line_box_clip.py:39-127: (a, b, intersects, c, d)
    ((-1.0, 0.0), (2.0, 3.0), True, (0.0, 1.0), (1.0, 2.0)),
    ((0.0, -0.1), (0.0, -0.1), False, (np.nan, np.nan), (np.nan, np.nan)),
    ((-1.0, 1.0), (3.0, 1.0), True, (0.0, 1.0), (2.0, 1.0)),
    ((1.0, -3.0), (1.0, 3.0), True, (1.0, 0.0), (1.0, 2.0)),
    ((1.0, -3.0), (1.0, 1.0), True, (1.0, 0.0), (1.0, 1.0)),
    ((1.0, 1.0), (1.0, 3.0), True, (1.0, 1.0), (1.0, 2.0)),
    ((-1.0, 3.0), (3.0, 3.0), False, None, None),
    ((-1.0, 1.0), (1.0, 1.0), True, (0.0, 1.0), (1.0, 1.0)),
    ((0.5, 0.5), (1.5, 1.5), True, (0.5, 0.5), (1.5, 1.5)),
    ((-1.5, 0.0), (-0.5, 1.0), False, None, None),
    ((2.5, 0.0), (3.5, 1.0), False, None, None),
]

DEGENERACY_CASES = [  # test_line_box_clip.py:130-304
    ((-1.0, -1.0), (3.0, 3.0), True, (0.0, 0.0), (2.0, 2.0)),
    ((-1.0, -1.0), (-1.0, -1.0), False, (np.nan, np.nan), (np.nan, np.nan)),
    ((-1.0, 0.0), (3.0, 0.0), True, (0.0, 0.0), (2.0, 0.0)),
    ((-1.0, 2.0), (3.0, 2.0), True, (0.0, 2.0), (2.0, 2.0)),
    ((-1.0, 0.0), (1.0, 0.0), True, (0.0, 0.0), (1.0, 0.0)),
    ((-1.0, 2.0), (1.0, 2.0), True, (0.0, 2.0), (1.0, 2.0)),
    ((0.5, 0.0), (1.5, 0.0), True, (0.5, 0.0), (1.5, 0.0)),
    ((0.5, 2.0), (1.5, 2.0), True, (0.5, 2.0), (1.5, 2.0)),
    ((0.0, 0.0), (2.0, 0.0), True, (0.0, 0.0), (2.0, 0.0)),
    ((0.0, 2.0), (2.0, 2.0), True, (0.0, 2.0), (2.0, 2.0)),
    ((0.0, 0.0), (0.0, 2.0), True, (0.0, 0.0), (0.0, 2.0)),
    ((2.0, 0.0), (2.0, 2.0), True, (2.0, 0.0), (2.0, 2.0)),
    ((0.0, 0.5), (0.0, 1.5), True, (0.0, 0.5), (0.0, 1.5)),
    ((2.0, 0.5), (2.0, 1.5), True, (2.0, 0.5), (2.0, 1.5)),
    ((-1.0, 1.0), (0.0, 2.0), False, (np.nan, np.nan), (np.nan, np.nan)),
]


@pytest.mark.parametrize("clip,shape", [(_cs, BOX), (_cb, POLY)], ids=["cohen_sutherland", "cyrus_beck"])
def test_line_box_clip(clip, shape):
    for a, b, inter, c, d in LINE_BOX_CASES:
        ok, cc, dd = clip(np.array(a), np.array(b), shape)
        assert ok is inter, (a, b)
        if c is not None:
            assert np.allclose(cc, c, equal_nan=True)
            assert np.allclose(dd, d, equal_nan=True)


@pytest.mark.parametrize(
    "clip,shape", [(_cs, BOX), (_cb, POLY), (_cb, POLY[::-1])], ids=["cs", "cb", "cb_reversed"]
)
def test_line_box_clip_degeneracy(clip, shape):
    for a, b, inter, c, d in DEGENERACY_CASES:
        ok, cc, dd = clip(np.array(a), np.array(b), shape)
        assert ok is inter, (a, b)
        assert np.allclose(cc, c, equal_nan=True)
        assert np.allclose(dd, d, equal_nan=True)


# --------------------------------------------------------- test_line_polygon_clip.py
def _flip_equal(a, b, poly):
    ok1, c1, d1 = _cb(a, b, poly)
    ok2, c2, d2 = _cb(b, a, poly)
    return ok1 == ok2 and np.array_equal(c1, d2, equal_nan=True) and np.array_equal(d1, c2, equal_nan=True)


def test_line_polygon_clip():  # :21-84
    poly = np.array([[1.0, 3.0], [4.0, 3.0], [4.0, 5.0], [1.0, 5.0]])
    a, b = np.array([0.0, 0.0]), np.array([4.0, 6.0])
    ok, c, d = _cb(a, b, poly)
    assert ok and np.allclose(c, [2.0, 3.0]) and np.allclose(d, [3.3333333333333, 5.0])
    assert _flip_equal(a, b, poly)
    a = b = np.array([0.0, 0.1])
    ok, c, d = _cb(a, b, poly)
    assert not ok and np.isnan(c).all() and np.isnan(d).all()
    a, b = np.array([0.0, 4.0]), np.array([5.0, 4.0])
    ok, c, d = _cb(a, b, poly)
    assert ok and np.allclose(c, [1.0, 4.0]) and np.allclose(d, [4.0, 4.0])
    assert _flip_equal(a, b, poly)
    for a, b, ec, ed in [
        ((1.0, -3.0), (1.0, 3.0), (1.0, 0.0), (1.0, 2.0)),
        ((1.0, -3.0), (1.0, 1.0), (1.0, 0.0), (1.0, 1.0)),
        ((1.0, 1.0), (1.0, 3.0), (1.0, 1.0), (1.0, 2.0)),
    ]:
        a, b = np.array(a), np.array(b)
        ok, c, d = _cb(a, b, POLY)
        assert ok and np.allclose(c, ec) and np.allclose(d, ed)
        assert _flip_equal(a, b, POLY)


def test_line_polygon_clip_through_vertex():  # :87-105
    ok, c, d = _cb(np.array([2.0, 2.0]), np.array([1.0, 4.0]), np.array([[2.0, 4.0], [1.0, 4.0], [1.0, 3.0], [2.0, 3.0]]))
    assert ok and np.allclose(c, [1.5, 3.0]) and np.allclose(d, [1.0, 4.0])
    a, b = np.array([2.0, 0.0]), np.array([2.0, 2.0])
    assert not _cb(a, b, np.array([[2.0, 3.0], [1.0, 3.0], [1.0, 2.0], [2.0, 2.0]]))[0]
    assert not _cb(a, b, np.array([[3.0, 3.0], [2.0, 3.0], [2.0, 2.0], [3.0, 2.0]]))[0]


def test_line_triangle_clip():  # :108-161
    poly = np.array([[0.0, 0.5], [2.0, 0.0], [2.0, 2.0]])
    a, b = np.array([1.0, 1.0]), np.array([3.0, 1.0])
    ok, c, d = _cb(a, b, poly)
    assert ok and np.allclose(c, [1.0, 1.0]) and np.allclose(d, [2.0, 1.0])
    assert _flip_equal(a, b, poly)
    poly = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0]])
    for a, b in [((0.0, 0.0), (2.0, 0.0)), ((2.0, 0.0), (2.0, 2.0)), ((0.0, 0.0), (2.0, 2.0)),
                 ((-1.0, -1.0), (2.0, 2.0)), ((-1.0, -1.0), (3.0, 3.0)), ((0.0, 0.0), (3.0, 3.0))]:
        assert _cb(np.array(a), np.array(b), poly)[0]
        assert _cb(np.array(b), np.array(a), poly)[0]


def test_line_box_degeneracies():  # :164-215
    n = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, 1.0], [0.0, 2.0], [1.0, 2.0], [2.0, 2.0]])
    f = np.array([[3, 4, 6, 7], [4, 5, 8, 7], [0, 1, 4, 3], [1, 2, 5, 4]])
    a, b = np.array([1.0, 1.0]), np.array([1.5, 1.25])
    ok, c, d = _cb(a, b, n[f[1]])
    assert ok and np.array_equal(c, a) and np.array_equal(d, b)
    for i in [0, 2, 3]:
        assert not _cb(a, b, n[f[i]])[0]
    a, b = np.array([0.5, 0.75]), np.array([1.0, 1.0])
    ok, c, d = _cb(a, b, n[f[2]])
    assert ok and np.array_equal(c, a) and np.array_equal(d, b)


# ------------------------------------------------------------ test_separating_axis.py
def test_separating_axes():  # :9-137, 183-223
    sa = R.separating_axes
    a = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    assert sa(a, a)
    b = np.array([[2.0, 0.0], [3.0, 1.0], [2.0, 1.0]])
    assert not sa(a, b) and not sa(b, a)
    b = np.array([[1.0, 0.0], [2.0, 1.0], [1.0, 1.0]])
    assert not sa(a, b) and not sa(b, a)
    a = np.array([[0.0, 0.0], [4.0, 0.0], [0.0, 4.0]])
    b = np.array([[1.0, 1.0], [2.0, 1.0], [1.0, 2.0]])
    assert sa(a, b) and sa(b, a)
    a = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    b = np.array([[1.0, 0.0], [2.0, 0.0], [2.0, 1.0]])
    assert not sa(a, b) and not sa(b, a)
    a = np.array([[5.0, 3.0], [7.0, 3.0], [7.0, 5.0]])
    b = np.array([[6.9, 5.6], [8.0, 4.75], [7.85, 5.9]])
    assert sa(a, b) and not sa(b, a)
    box = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    t = np.array([[1.0, 1.0], [3.0, 0.0], [3.0, 1.0]])
    assert sa(box, t) and sa(t, box)
    t = np.array([[2.0, 1.0], [3.0, 0.0], [3.0, 1.0]])
    assert not sa(box, t) and not sa(t, box)
    t = np.array([[0.25, 0.25], [0.75, 0.75], [0.6, 0.6]])
    assert sa(box, t) and sa(t, box)
    a = np.array([[0.0, 0.0], [0.5, 0.5], [0.0, 1.0], [1.0, 1.0]])
    assert sa(a, a)
    b = np.array([[2.0, 0.0], [3.0, 1.0], [2.5, 1.0], [2.0, 1.0]])
    assert not sa(a, b) and not sa(b, a)
    a = np.array([[0.0, 0.0], [0.5, 0.5], [0.5, 0.5], [0.0, 1.0], [1.0, 1.0]])
    assert sa(a, a)


# --------------------------------------------------------------- test_barycentric.py
def test_interp_edge_case():  # :11-22
    for p, e in [((0.0, 0.0), [1, 0, 0, 0]), ((1.0, 0.0), [0, 1, 0, 0]), ((0.25, 0.0), [0.75, 0.25, 0, 0]),
                 ((0.75, 0.0), [0.25, 0.75, 0, 0])]:
        assert np.allclose(R.interp_edge_case([0.0, 0.0], [1.0, 0.0], p, 4, 0, 1), e)


def test_compute_weights_triangle():  # :25-44
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    for p, e in [((0.0, 0.0), [1, 0, 0]), ((1.0, 0.0), [0, 1, 0]), ((1.0, 1.0), [0, 0, 1]),
                 ((0.5, 0.0), [0.5, 0.5, 0]), ((1.0, 0.5), [0, 0.5, 0.5]), ((0.5, 0.5), [0.5, 0, 0.5])]:
        assert np.allclose(R.triangle_weights(tri, p), e)


def test_compute_weights_wachspress():  # :95-115
    sq = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    for p, e in [((0.0, 0.0), [1, 0, 0, 0]), ((1.0, 0.0), [0, 1, 0, 0]), ((1.0, 1.0), [0, 0, 1, 0]),
                 ((0.0, 1.0), [0, 0, 0, 1]), ((0.25, 0.0), [0.75, 0.25, 0, 0]), ((0.25, 1.0), [0, 0, 0.25, 0.75]),
                 ((0.5, 0.5), [0.25, 0.25, 0.25, 0.25])]:
        assert np.allclose(R.wachspress_weights(sq, p, 4, 1e-9), e)


# ---------------------------------------------------------- test_geometry_utils.py
OFFSET = 2e-9
LINE_CASES = [  # :351-401: (p, q, intersects, point)
    ((3.0, 2.0), (3.0, 1.0), False, (np.nan, np.nan)),
    ((3.0, 3.0 - OFFSET), (3.0, 1.0), False, (np.nan, np.nan)),
    ((-1.0, 1.0), (1.0, -1.0), True, (0.0, 0.0)),
    ((1.0, 1.0), (3.0, 3.0), True, (2.0, 2.0)),
    ((1.0, 3.0), (3.0, 1.0), True, (2.0, 2.0)),
    ((0.0, 0.0), (-1.0, -1.0), True, (0.0, 0.0)),
    ((-OFFSET, -OFFSET), (-1.0, -1.0), False, (np.nan, np.nan)),
]


@pytest.mark.parametrize("p,q,inter,pt", LINE_CASES)
def test_lines_intersect(p, q, inter, pt):  # :404-419
    a, b = (0.0, 0.0), (4.0, 4.0)
    ok, x, y = R.lines_intersect(a, b, p, q)
    assert ok == inter
    np.testing.assert_allclose([x, y], pt)
    ok, x, y = R.lines_intersect(p, q, a, b)
    assert ok == inter
    np.testing.assert_allclose([x, y], pt)


def test_point_in_polygon_or_on_edge():
    # A point strictly inside, strictly outside, and within tolerance of an edge.
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    assert R.point_in_polygon_or_on_edge([0.5, 0.25], tri, 1e-9)
    assert not R.point_in_polygon_or_on_edge([1.5, 0.25], tri, 1e-9)
    assert R.point_in_polygon_or_on_edge([0.5, -1e-10], tri, 1e-9)
    assert not R.point_in_polygon_or_on_edge([0.5, -1e-8], tri, 1e-9)
    assert R.point_on_edge([0.5, 1e-10], [0.0, 0.0, 1.0, 0.0], 1e-9)
    assert not R.point_on_edge([1.5, 0.0], [0.0, 0.0, 1.0, 0.0], 1e-9)


def _no_sqrt_distances(bb):
    d = np.empty((bb.shape[0], 3))
    d[:, 0] = bb[:, 1] - bb[:, 0]
    d[:, 1] = bb[:, 3] - bb[:, 2]
    d[:, 2] = d[:, 0] ** 2 + d[:, 1] ** 2
    return d


def _min_tolerance(bb_diagonal):
    return max(C.MIN_TOLERANCE, C.TOLERANCE_FACTOR * min(bb_diagonal))


@pytest.mark.parametrize(
    "attr, planted", [("bbox_distances", _no_sqrt_distances), ("default_tolerance", _min_tolerance)]
)
def test_util_cases_catch_planted_slips(monkeypatch, attr, planted):
    """A dropped sqrt in bbox_distances, or min for max in default_tolerance, planted in the
    oracle, fails the utility cases (so those functions are pinned, not just exercised)."""
    monkeypatch.setattr(R, attr, planted)
    impl = types.SimpleNamespace(**{**vars(IMPL), attr: planted})
    failed = []
    for case in C.UTIL_CASES:
        try:
            case(impl)
        except AssertionError:
            failed.append(case.__name__)
    assert failed, f"no utility case noticed the planted {attr}"
    assert "check_default_tolerance_of_tree" in failed
