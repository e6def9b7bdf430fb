"""The Wachspress adjoint (ct_barycentric_wachspress_backward) on the GPU.

Each point's face and the forward's branch are held fixed (barycentric_wachspress.py:38-85:
interior weights, or the linear interpolation along the first edge within tolerance).  Checks:
* the forward equals compute_barycentric_weights bit for bit;
* gradients in the points and the vertices match fourth-order central differences within
  1e-8, on interior points of a mixed triangle / quad mesh and of a hexagon-Voronoi mesh, and
  on points inside the edge-interpolation band (a large explicit tolerance, so that the finite
  difference steps stay in that branch);
* torch.autograd.gradcheck on a small problem.
"""

import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _meshes():
    v, f = S.jittered_grid(41, 10.0, 0.25, 1.0 / 3.0, 3)  # triangles and quads, 4 columns
    yield "mixed", v, f
    v, f = S.hex_voronoi(14, 14, 10.0, 0.1, 5)  # 3-8 vertex polygons
    yield "voronoi", v, f


def _fd_check(tree, pts, face, tol, G, rng, label):
    import torch

    from numba_celltree_amd.barycentric import BarycentricFunction, barycentric_weights

    p = pts.clone().requires_grad_(True)
    vv = tree.d_vertices.clone().requires_grad_(True)
    f2, w = barycentric_weights(tree, p, vv, tolerance=tol)
    assert torch.equal(f2, face)
    (w * G).sum().backward()

    def loss(px, vx):
        with torch.no_grad():
            return float((BarycentricFunction.apply(px, vx, tree.d_elements, face, tol, True) * G).sum())

    h = 1e-6
    for name, base, grad, k in (("points", pts, p.grad, 0), ("vertices", tree.d_vertices, vv.grad, 1)):
        for trial in range(2):
            d = torch.from_numpy(rng.normal(size=tuple(base.shape))).cuda()

            def at(s):
                args = [pts, tree.d_vertices]
                args[k] = base + s * h * d
                return loss(*args)

            fd = (8.0 * (at(1) - at(-1)) - (at(2) - at(-2))) / (12.0 * h)
            an = float((grad * d).sum())
            rel = abs(fd - an) / max(abs(an), abs(fd))
            print(f"[wachspress fd] {label} {name} {trial}: analytic {an:.15e} fd {fd:.15e} rel {rel:.2e}",
                  file=sys.stderr, flush=True)
            assert rel < 1e-8, (label, name, an, fd, rel)


@pytest.mark.parametrize("name, v, f", list(_meshes()), ids=lambda x: x if isinstance(x, str) else "")
def test_wachspress_adjoint_interior(name, v, f):
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd.barycentric import barycentric_weights

    tree = M.CellTree2d(v, f, -1)
    assert int(tree.d_elements.shape[1]) > 3
    rng = np.random.default_rng(11)
    pts = torch.from_numpy(rng.uniform(0.3, 9.7, size=(20_000, 2))).cuda()
    face, w = barycentric_weights(tree, pts)
    ref_face, ref_w = tree.compute_barycentric_weights(pts)
    assert torch.equal(face, ref_face) and torch.equal(w, ref_w)
    # Wachspress weights vanish on the edges a vertex is not on: keep points whose weights are
    # all clear of 0 (the interior branch, away from the kinks of |.|).
    nv = (tree.d_elements >= 0).sum(1)[face.clamp(min=0)]
    k = torch.arange(w.shape[1], device=w.device)[None, :]
    wmin = torch.where(k < nv[:, None], w, torch.ones_like(w)).min(1).values
    keep = (face >= 0) & (wmin > 2e-3)
    pts, face = pts[keep].contiguous(), face[keep].contiguous()
    assert len(pts) > 5000
    G = torch.from_numpy(rng.normal(size=(len(pts), w.shape[1]))).cuda()
    _fd_check(tree, pts, face, tree._default_tolerance(), G, rng, name)


def test_wachspress_adjoint_edge_band():
    """Points inside the edge-interpolation band of a large tolerance (0.05 x the edge
    length squared, in within_perpendicular_distance's units): the weights are the linear
    interpolation along the edge, and so are their derivatives."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import _device as D

    v, f = S.jittered_grid(21, 10.0, 0.25, 0.0, 4)  # quads only
    tree = M.CellTree2d(v, f, -1)
    tol = 0.05
    rng = np.random.default_rng(5)
    faces = tree.faces
    rows = rng.integers(0, len(faces), 3000)
    k = rng.integers(0, 4, 3000)
    a = v[faces[rows, k]]
    b = v[faces[rows, (k + 1) % 4]]
    u = b - a
    nrm = np.column_stack([-u[:, 1], u[:, 0]])
    # inside the cell, a quarter of the band's half-width from the edge line, away from corners
    t = rng.uniform(0.2, 0.8, 3000)
    off = 0.25 * tol * np.linalg.norm(u, axis=1)
    pts_h = a + t[:, None] * u + (off / np.linalg.norm(u, axis=1))[:, None] * nrm
    pts = torch.from_numpy(pts_h).cuda()
    face = D.to_dev(rows, torch.int64)
    # the tree's own location must agree (first containing cell in DFS order may differ for
    # points inside the band of a shared edge): keep the points located in the chosen cell
    located = tree._locate_points_dev(pts, tol)
    keep = located == face
    pts, face = pts[keep].contiguous(), face[keep].contiguous()
    assert len(pts) > 500
    from numba_celltree_amd.barycentric import BarycentricFunction

    w = BarycentricFunction.apply(pts, tree.d_vertices, tree.d_elements, face, tol, True)
    # edge branch: exactly two nonzero weights summing to 1
    assert bool(((w != 0).sum(1) == 2).all())
    G = torch.from_numpy(rng.normal(size=(len(pts), 4))).cuda()
    _fd_check(tree, pts, face, tol, G, rng, "edge band")


def test_wachspress_gradcheck():
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd.barycentric import BarycentricFunction

    v, f = S.hex_voronoi(4, 4, 4.0, 0.1, 2)
    tree = M.CellTree2d(v, f, -1)
    tol = tree._default_tolerance()
    rng = np.random.default_rng(3)
    pts = torch.from_numpy(rng.uniform(0.5, 3.5, size=(40, 2))).cuda()
    face = tree._locate_points_dev(pts, tol)
    keep = face >= 0
    pts, face = pts[keep].contiguous().requires_grad_(True), face[keep].contiguous()
    vs = tree.d_vertices.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(
        lambda a, b: BarycentricFunction.apply(a, b, tree.d_elements, face, tol, True), (pts, vs),
        eps=1e-6, atol=1e-8, rtol=1e-6)
