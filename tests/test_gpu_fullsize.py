"""The plain GPU workloads at their full size against the oracle: the 999,941-cell mixed
triangle/quad mesh (708 x 708 jittered grid on [0, 1000]^2) with 1e7 points (5% outside the
bbox), 1e6 boxes and 1e6 segments, and intersect_faces of 999,698 triangles against 250,000
Voronoi cells, exactly as the benchmark generates them.

Bar: the tree byte-identical; every point's cell, every (box, cell) pair and every (segment,
cell) row identical, in the reference's order; intersection points within 1e-12 relative of
the domain size.  The oracle takes about 15-30 s of host time for all of it.
"""

import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[gpu fullsize] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def full():
    import torch

    import numba_celltree_amd as M

    log("generating the mesh and queries")
    v, f = S.mixed_mesh_1m()
    pts, bx, sg = S.mixed_mesh_queries(10_000_000, 1_000_000, 1_000_000)
    log("building the trees")
    g = M.CellTree2d(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda(), -1)
    o = R.OracleCellTree2d(v, f, -1)
    return v, f, pts, bx, sg, g, o


def test_full_tree_identical(full):
    v, f, pts, bx, sg, g, o = full
    gn, on = g.nodes, o.nodes
    assert gn.shape == on.shape
    for field in ("child", "ptr", "size", "dim"):
        assert np.array_equal(gn[field], on[field]), field
    for field in ("Lmax", "Rmin"):
        assert np.array_equal(gn[field].view(np.int64), on[field].view(np.int64)), field
    assert np.array_equal(g.bb_indices, o.bb_indices)


def test_full_locate_points(full):
    import torch

    v, f, pts, bx, sg, g, o = full
    log("locate_points 1e7")
    got = g.locate_points(torch.from_numpy(pts).cuda()).cpu().numpy()
    want = o.locate_points(pts)
    assert np.array_equal(got, want)
    assert 0.9 < (want >= 0).mean() < 1.0  # the outside points are there


def test_full_locate_boxes(full):
    import torch

    v, f, pts, bx, sg, g, o = full
    log("locate_boxes 1e6")
    gi, gj = (x.cpu().numpy() for x in g.locate_boxes(torch.from_numpy(bx).cuda()))
    oi, oj = o.locate_boxes(bx)
    assert len(gi) == len(oi) > 10_000_000
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)


def test_full_intersect_edges(full):
    import torch

    v, f, pts, bx, sg, g, o = full
    log("intersect_edges 1e6")
    gi, gj, gxy = (x.cpu().numpy() for x in g.intersect_edges(torch.from_numpy(sg).cuda()))
    oi, oj, oxy = o.intersect_edges(sg)
    assert len(gi) == len(oi) > 1_000_000
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    err = np.abs(gxy - oxy)
    assert np.all(err <= 1e-12 * np.maximum(np.abs(oxy), 1000.0))


def test_full_intersect_faces():
    """intersect_faces at full size: 999,698 jittered-grid triangles against 250,000 hex
    Voronoi cells (3 to 7 vertices): the tree identical, every (target, source) pair
    identical in the reference's order, overlap areas within 1e-12 relative."""
    import torch

    import numba_celltree_amd as M

    log("intersect_faces 999,698 x 250,000")
    sv, sf, tv, tf, _ = S.overlap_meshes()
    g = M.CellTree2d(torch.from_numpy(sv).cuda(), torch.from_numpy(sf).cuda(), -1)
    o = R.OracleCellTree2d(sv, sf, -1)
    assert np.array_equal(g.bb_indices, o.bb_indices)
    gi, gj, ga = (x.cpu().numpy() for x in g.intersect_faces(torch.from_numpy(tv).cuda(),
                                                             torch.from_numpy(tf).cuda(), -1))
    oi, oj, oa = o.intersect_faces(tv, tf, -1)
    assert len(gi) == len(oi) > 2_000_000
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    h2 = (1000.0 / 707) ** 2  # a source cell's area scale
    assert np.all(np.abs(ga - oa) <= 1e-12 * np.maximum(np.abs(oa), h2))


def test_full_regrid_forward(full):
    """The regridding forward at full size (16 fields from the 999,941-cell mesh to 250,000
    Voronoi cells): pairs identical to the oracle's intersect_faces, and y = sum(area * F) /
    target area within 1e-12 of that sum over the oracle's areas."""
    import torch

    from numba_celltree_amd.regrid import Regridder

    v, f, pts, bx, sg, g, o = full
    log("regrid forward 16 fields")
    tv, tf = S.hex_voronoi(500, 500, 1000.0, 0.1, 7)
    F = S.sinusoid_fields(v[np.maximum(f, 0)].mean(1), 16, 1)
    rg = Regridder(g, torch.from_numpy(tv).cuda(), torch.from_numpy(tf).cuda(), -1)
    y = rg(torch.from_numpy(F).cuda()).cpu().numpy()
    i, j, area = o.intersect_faces(tv, tf, -1)
    assert np.array_equal(rg.topology.tt.cpu().numpy(), i) and np.array_equal(rg.topology.ss.cpu().numpy(), j)
    ccw = tf.astype(np.int64).copy()
    R.counter_clockwise(tv, ccw)
    tarea = np.array([R.polygon_area(tv[row[row >= 0]]) for row in ccw])
    y_ref = np.zeros((len(tf), F.shape[1]))
    np.add.at(y_ref, i, area[:, None] * F[j])
    y_ref /= tarea[:, None]
    assert np.allclose(y, y_ref, rtol=1e-12, atol=1e-12 * np.abs(F).max())


def test_full_barycentric_weights():
    """compute_barycentric_weights for 1e6 points on the 20,000-triangle mesh (uniform over
    [-0.5, 10.5]^2, so some fall outside): cells identical, weights within 1e-12."""
    import numba_celltree_amd as M

    log("barycentric 1e6 points")
    v, f, _ = S.triangle_grid_20k()
    pts = np.random.default_rng(4).uniform(-0.5, 10.5, size=(1_000_000, 2))
    gf, gw = M.CellTree2d(v, f, -1).compute_barycentric_weights(pts)
    of, ow = R.OracleCellTree2d(v, f, -1).compute_barycentric_weights(pts)
    assert np.array_equal(gf, of)
    assert (of == -1).any() and (of >= 0).mean() > 0.8
    assert np.all(np.abs(gw - ow) <= 1e-12 * np.maximum(np.abs(ow), 1.0))


def test_full_barycentric_adjoint():
    """The barycentric adjoint at its full size, 1e6 points on the 20,000-triangle mesh: the
    differentiable forward equals the oracle's weights (cells identical, weights within
    1e-12), every gradient is finite, points outside get zero gradient, and directional
    derivatives of sum(w * G) in the points and in the vertices match fourth-order central
    differences within 1e-8 over the points that are clear of the triangle edges (where the
    weights' |.| has a kink)."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd.barycentric import BarycentricFunction, barycentric_weights

    log("barycentric adjoint 1e6 points")
    v, f, _ = S.triangle_grid_20k()
    tree = M.CellTree2d(v, f, -1)
    rng = np.random.default_rng(4)
    pts_h = rng.uniform(-0.5, 10.5, size=(1_000_000, 2))
    pts = torch.from_numpy(pts_h).cuda().requires_grad_(True)
    vv = tree.d_vertices.clone().requires_grad_(True)
    face, w = barycentric_weights(tree, pts, vv)
    of, ow = R.OracleCellTree2d(v, f, -1).compute_barycentric_weights(pts_h)
    assert np.array_equal(face.cpu().numpy(), of)
    assert np.all(np.abs(w.detach().cpu().numpy() - ow) <= 1e-12 * np.maximum(np.abs(ow), 1.0))
    G = torch.from_numpy(rng.normal(size=(1_000_000, 3))).cuda()
    (w * G).sum().backward()
    assert bool(torch.isfinite(pts.grad).all()) and bool(torch.isfinite(vv.grad).all())
    assert bool((pts.grad[face < 0] == 0).all())
    keep = (face >= 0) & (w.detach().min(1).values > 1e-3)
    P = pts.detach()[keep].contiguous()
    fk = face[keep].contiguous()
    Gk = G[keep].contiguous()
    tol = tree._default_tolerance()
    p2 = P.clone().requires_grad_(True)
    v2 = tree.d_vertices.clone().requires_grad_(True)
    (BarycentricFunction.apply(p2, v2, tree.d_elements, fk, tol, False) * Gk).sum().backward()

    def loss(px, vx):
        with torch.no_grad():
            return float((BarycentricFunction.apply(px, vx, tree.d_elements, fk, tol, False) * Gk).sum())

    # h = 1e-6: kept points are >= ~7e-5 from their triangle's edge lines (weights > 1e-3 at
    # heights ~0.07), beyond the +-2h |d| perturbations (a larger h lets points cross a line,
    # the kink of |.|).  The rounding of the 9e5-term sums is a random walk, ~1e-13 per
    # evaluation, far below the bound after division by h.
    h = 1e-6
    for name, base, grad, k in (("points", P, p2.grad, 0), ("vertices", tree.d_vertices, v2.grad, 1)):
        d = torch.from_numpy(rng.normal(size=tuple(base.shape))).cuda()

        def at(s):
            args = [P, tree.d_vertices]
            args[k] = base + s * h * d
            return loss(*args)

        fd = (8.0 * (at(1) - at(-1)) - (at(2) - at(-2))) / (12.0 * h)
        an = float((grad * d).sum())
        rel = abs(fd - an) / max(abs(an), abs(fd))
        log(f"barycentric adjoint fd {name}: analytic {an:.15e} fd {fd:.15e} rel {rel:.2e}")
        assert rel < 1e-8, (name, an, fd, rel)
