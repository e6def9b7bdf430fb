"""Regrid forward and hand-written adjoint (numba_celltree_amd.regrid) on the GPU.

Checks:
* forward: y equals sum(area * F) / target_area from the oracle's intersect_faces areas;
* gradients: directional central finite differences in FP64 on a 2,048-cell source /
  144-cell target subproblem, for F, the source vertices and the target vertices, within
  1e-8 relative; torch.autograd.gradcheck on a small problem;
* an independent reference: the Sutherland-Hodgman clip restated with torch operations
  (sutherland_hodgman.py:56-148, geometry_utils.py:82-95) and differentiated by autograd,
  on the same fixed topology.
"""

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def problem(nv, nt, length, seed=0, n_fields=16, inset=0.0):
    """Source grid and target hex-Voronoi mesh on [0, length]^2.  With inset > 0 the target
    mesh is scaled into [inset, length - inset]^2: the two meshes then share no boundary
    edge, whose collinear overlap is a kink of the area in the vertices (one-sided
    derivatives differ), where finite differences do not measure the gradient."""
    sv, sf = S.jittered_grid(nv, length, 0.25, 1.0, seed)
    tv, tf = S.hex_voronoi(nt, nt, length, 0.1, seed + 1)
    if inset:
        tv = inset + tv * ((length - 2 * inset) / length)
    F = S.sinusoid_fields(sv[sf].mean(1) * (200.0 / length), n_fields, 1)
    return sv, sf, tv, tf, F


def build(sv, sf, tv, tf):
    import numba_celltree_amd as M
    from numba_celltree_amd.regrid import Regridder

    tree = M.CellTree2d(sv, sf, -1)
    return tree, Regridder(tree, tv, tf, -1)


def test_forward_matches_oracle_areas():
    import torch

    sv, sf, tv, tf, F = problem(41, 20, 10.0)
    tree, rg = build(sv, sf, tv, tf)
    y = rg(torch.from_numpy(F).cuda()).cpu().numpy()
    o = R.OracleCellTree2d(sv, sf, -1)
    i, j, area = o.intersect_faces(tv, tf, -1)
    topo = rg.topology
    assert np.array_equal(topo.tt.cpu().numpy(), i) and np.array_equal(topo.ss.cpu().numpy(), j)
    ccw = tf.astype(np.int64).copy()
    R.counter_clockwise(tv, ccw)
    tarea = np.array([R.polygon_area(tv[row[row >= 0]]) for row in ccw])
    y_ref = np.zeros((len(tf), F.shape[1]))
    np.add.at(y_ref, i, area[:, None] * F[j])
    y_ref /= tarea[:, None]
    assert np.allclose(y, y_ref, rtol=1e-12, atol=1e-12 * np.abs(F).max())
    # conservation: the regridded constant field is 1 wherever the target is covered
    ones = rg(torch.ones(len(sf), 1, dtype=torch.float64, device="cuda")).cpu().numpy()[:, 0]
    assert np.allclose(ones, 1.0, rtol=1e-12)


def _loss_grads(rg, F, sv, tv, G):
    import torch

    F = F.clone().requires_grad_(True)
    sv = sv.clone().requires_grad_(True)
    tv = tv.clone().requires_grad_(True)
    y = rg(F, sv, tv)
    loss = (y * G).sum()
    loss.backward()
    return F.grad, sv.grad, tv.grad


def test_gradients_match_finite_differences():
    import torch

    sv, sf, tv, tf, F = problem(33, 12, 4.0, inset=0.0731)
    tree, rg = build(sv, sf, tv, tf)
    dev = "cuda"
    F_t = torch.from_numpy(F).to(dev)
    sv_t = torch.from_numpy(sv).to(dev)
    tv_t = rg.topology.tgt_vertices.clone()
    G = torch.from_numpy(np.random.default_rng(0).normal(size=(len(tf), F.shape[1]))).to(dev)
    gF, gsv, gtv = _loss_grads(rg, F_t, sv_t, tv_t, G)

    def loss(Fx, svx, tvx):
        with torch.no_grad():
            return float((rg(Fx, svx, tvx) * G).sum())

    rng = np.random.default_rng(1)
    h = 1e-6
    for name, base, grad in (("F", F_t, gF), ("src", sv_t, gsv), ("tgt", tv_t, gtv)):
        for trial in range(3):
            d = torch.from_numpy(rng.normal(size=tuple(base.shape))).to(dev)
            args_p = [F_t, sv_t, tv_t]
            args_m = [F_t, sv_t, tv_t]
            k = {"F": 0, "src": 1, "tgt": 2}[name]
            args_p[k] = base + h * d
            args_m[k] = base - h * d
            fd = (loss(*args_p) - loss(*args_m)) / (2 * h)
            an = float((grad * d).sum())
            rel = abs(fd - an) / max(abs(an), abs(fd), 1e-300)
            print(f"[regrid fd] {name} trial {trial}: analytic {an:.15e} fd {fd:.15e} rel {rel:.2e}")
            assert rel < 1e-8, (name, an, fd, rel)


def test_vertex_gradients_match_oracle_adjoint():
    """The backward's pair adjoint (k_pair_backward and its corner sums) against the oracle's
    restatement (ref_pair_area_gradient, pinned by finite differences of the oracle's clip
    area): the source-vertex gradients of sum(y * G) equal the sum over pairs of
    (G_t . F_s / area_t) dA_ts / d(source vertices), gathered per vertex (the target vertices
    also carry the target-area term and are checked by the finite differences above)."""
    import torch

    sv, sf, tv, tf, F = problem(41, 20, 10.0)
    tree, rg = build(sv, sf, tv, tf)
    topo = rg.topology
    dev = "cuda"
    G = np.random.default_rng(2).normal(size=(len(tf), F.shape[1]))
    _, gsv, _ = _loss_grads(rg, torch.from_numpy(F).to(dev), torch.from_numpy(sv).to(dev),
                            topo.tgt_vertices.clone(), torch.from_numpy(G).to(dev))
    t = topo.tt.cpu().numpy()
    s_ = topo.ss.cpu().numpy()
    ccw = tf.astype(np.int64).copy()
    R.counter_clockwise(tv, ccw)
    tarea = np.array([R.polygon_area(tv[row[row >= 0]]) for row in ccw])
    gA = (G[t] * F[s_]).sum(1) / tarea[t]
    faces = tree.faces.astype(np.int64)
    _, gb, ops = R.pair_area_gradient(tv, sv, ccw, faces, t, s_, gA)
    assert (ops > 0).all()
    want = np.zeros_like(sv)
    rows = faces[s_]
    for slot in range(faces.shape[1]):
        ok = rows[:, slot] >= 0
        np.add.at(want, rows[ok, slot], gb[ok, slot])
    got = gsv.cpu().numpy()
    scale = np.abs(want).max()
    print(f"[regrid adjoint] {len(t)} pairs, max |grad| {scale:.3e}, max diff {np.abs(got - want).max():.3e}")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)


def test_torch_gradcheck_small():
    import torch

    sv, sf, tv, tf, F = problem(7, 3, 1.0, seed=3, n_fields=2, inset=0.0417)
    tree, rg = build(sv, sf, tv, tf)
    F_t = torch.from_numpy(F).cuda().requires_grad_(True)
    sv_t = torch.from_numpy(sv).cuda().requires_grad_(True)
    tv_t = rg.topology.tgt_vertices.clone().requires_grad_(True)
    # the vertex gradients are reduced in a fixed order: gradcheck's reentrancy check is exact
    assert torch.autograd.gradcheck(lambda a, b, c: rg(a, b, c), (F_t, sv_t, tv_t), eps=1e-6, atol=1e-9,
                                    rtol=1e-7)


def test_torch_gradcheck_2048_cells():
    """gradcheck at the 2,000-cell scale: 2,048 source triangles, 144 hexagon-Voronoi targets,
    F and both vertex sets.  A full numerical Jacobian would take ~10^4 evaluations, so this
    uses fast mode: vJu for random u, v from central differences against the adjoint."""
    import torch

    sv, sf, tv, tf, F = problem(33, 12, 1.0, seed=11, n_fields=4, inset=0.0173)
    assert len(sf) == 2048
    tree, rg = build(sv, sf, tv, tf)
    F_t = torch.from_numpy(F).cuda().requires_grad_(True)
    sv_t = torch.from_numpy(sv).cuda().requires_grad_(True)
    tv_t = rg.topology.tgt_vertices.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b, c: rg(a, b, c), (F_t, sv_t, tv_t), eps=1e-6, atol=1e-9,
                                    rtol=1e-7, fast_mode=True)


# ------------------------------------------------------------- torch restatement
def _torch_clip_area(subj, clip):
    """sutherland_hodgman.py:84-148 + geometry_utils.py:82-95 in torch ops (branches on
    values, arithmetic differentiable)."""
    import torch

    out = list(subj)
    r = clip[-1]
    for s in clip:
        U = s - r
        if float(U[0]) == 0 and float(U[1]) == 0:
            continue
        N = torch.stack([-U[1], U[0]])
        inp = out
        out = []
        a = inp[-1]
        a_in = float(U[0] * (a[1] - r[1])) > float(U[1] * (a[0] - r[0]))
        for b in inp:
            V = b - a
            if float(V[0]) == 0 and float(V[1]) == 0:
                continue
            b_in = float(U[0] * (b[1] - r[1])) > float(U[1] * (b[0] - r[0]))
            nv = (N * V).sum()
            if b_in:
                if not a_in and float(nv) != 0:
                    t = (N * (r - a)).sum() / nv
                    out.append(a + t * V)
                out.append(b)
            elif a_in:
                if float(nv) != 0:
                    t = (N * (r - a)).sum() / nv
                    out.append(a + t * V)
                else:
                    b_in = True
                    out.append(b)
            a = b
            a_in = b_in
        if len(out) < 3:
            return subj[0][0] * 0.0
        r = s
    area = out[0][0] * 0.0
    U = out[1] - out[0]
    for c in out[2:]:
        V = out[0] - c
        area = area + torch.abs(U[0] * V[1] - U[1] * V[0])
        U = V
    return 0.5 * area


def _fan(poly):
    area = poly[0][0] * 0.0
    U = poly[1] - poly[0]
    for c in poly[2:]:
        V = poly[0] - c
        area = area + abs(U[0] * V[1] - U[1] * V[0])
        U = V
    return 0.5 * area


def test_matches_torch_autograd_restatement():
    import torch

    # inset: no target edge collinear with a source boundary edge, where the clip's
    # arithmetic has degenerate (near-parallel) intersections and no unique derivative
    sv, sf, tv, tf, F = problem(9, 5, 2.0, seed=5, n_fields=3, inset=0.0613)
    tree, rg = build(sv, sf, tv, tf)
    topo = rg.topology
    G = np.random.default_rng(2).normal(size=(len(tf), F.shape[1]))
    gF, gsv, gtv = _loss_grads(rg, torch.from_numpy(F).cuda(), torch.from_numpy(sv).cuda(),
                               topo.tgt_vertices.clone(), torch.from_numpy(G).cuda())
    # reference on the CPU, same topology and the same CCW face rows
    Fc = torch.from_numpy(F).requires_grad_(True)
    svc = torch.from_numpy(sv.copy()).requires_grad_(True)
    tvc = topo.tgt_vertices.cpu().clone().requires_grad_(True)
    sfc = topo.src_faces.cpu().numpy()
    tfc = topo.tgt_faces.cpu().numpy()
    tt = topo.tt.cpu().numpy()
    ss = topo.ss.cpu().numpy()
    y = [torch.zeros(F.shape[1], dtype=torch.float64) for _ in range(len(tf))]
    for t, s in zip(tt, ss):
        subj = [tvc[v] for v in tfc[t] if v >= 0]
        clip = [svc[v] for v in sfc[s] if v >= 0]
        y[t] = y[t] + _torch_clip_area(subj, clip) * Fc[s]
    loss = 0.0
    for t in range(len(tf)):
        tarea = _fan([tvc[v] for v in tfc[t] if v >= 0])
        loss = loss + (y[t] / tarea * torch.from_numpy(G[t])).sum()
    loss.backward()
    for a, b, name in ((gF, Fc.grad, "F"), (gsv, svc.grad, "src"), (gtv, tvc.grad, "tgt")):
        a = a.cpu().numpy()
        b = b.numpy()
        scale = np.abs(b).max()
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12 * scale), name


# ------------------------------------------------------------------ barycentric adjoint
def test_barycentric_adjoint():
    """Triangle weights with faces held fixed: forward equals compute_barycentric_weights;
    gradients in points and vertices match central finite differences within 1e-8."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd.barycentric import barycentric_weights

    v, f, _ = S.triangle_grid_20k()
    tree = M.CellTree2d(v, f, -1)
    rng = np.random.default_rng(4)
    pts = torch.from_numpy(rng.uniform(0.2, 9.8, size=(20_000, 2))).cuda()
    face, w = barycentric_weights(tree, pts)
    ref_face, ref_w = tree.compute_barycentric_weights(pts)
    assert torch.equal(face, ref_face) and torch.equal(w, ref_w)
    # The weights are |cross| / area ratios: |.| has a kink where a point crosses an edge
    # line.  The finite-difference check keeps points at least 1e-3 (in weight) inside.
    inside = (face >= 0) & (w.min(dim=1).values > 1e-3)
    pts = pts[inside].contiguous()
    face = face[inside].contiguous()
    G = torch.from_numpy(rng.normal(size=(len(pts), 3))).cuda()
    p = pts.clone().requires_grad_(True)
    vv = tree.d_vertices.clone().requires_grad_(True)
    _, w = barycentric_weights(tree, p, vv)
    (w * G).sum().backward()
    from numba_celltree_amd.barycentric import BarycentricFunction

    tol = tree._default_tolerance()

    def loss(px, vx):
        with torch.no_grad():
            return float((BarycentricFunction.apply(px, vx, tree.d_elements, face, tol, False) * G).sum())

    # fourth-order central differences: truncation O(h^4), rounding ~ eps * |loss| / h
    h = 1e-6
    for name, base, grad, k in (("points", pts, p.grad, 0), ("vertices", tree.d_vertices, vv.grad, 1)):
        d = torch.from_numpy(rng.normal(size=tuple(base.shape))).cuda()

        def at(s):
            args = [pts, tree.d_vertices]
            args[k] = base + s * h * d
            return loss(*args)

        fd = (8.0 * (at(1) - at(-1)) - (at(2) - at(-2))) / (12.0 * h)
        an = float((grad * d).sum())
        rel = abs(fd - an) / max(abs(an), abs(fd))
        print(f"[bary fd] {name}: analytic {an:.15e} fd {fd:.15e} rel {rel:.2e}")
        assert rel < 1e-8, (name, an, fd, rel)
    small = pts[:50].clone().requires_grad_(True)
    vs = tree.d_vertices.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(
        lambda a, b: BarycentricFunction.apply(a, b, tree.d_elements, face[:50], tol, False), (small, vs),
        eps=1e-6, atol=1e-8, rtol=1e-6,
    )


def test_inputs_are_checked_or_converted():
    """float32 / host inputs are converted to float64 on the GPU by Regridder and
    barycentric_weights (same result as float64 device inputs); the autograd Functions
    called directly reject them, and wrong sizes raise before any kernel runs."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd.barycentric import BarycentricFunction, barycentric_weights
    from numba_celltree_amd.regrid import RegridFunction

    sv, sf, tv, tf, F = problem(21, 8, 4.0, n_fields=3)
    tree, rg = build(sv, sf, tv, tf)
    F64 = torch.from_numpy(F).cuda()
    want = rg(F64)
    F32 = F64.float()
    assert torch.equal(rg(F32), rg(F32.double()))
    assert torch.equal(rg(F), want)  # numpy in
    assert torch.equal(rg(F64.cpu()), want)  # host tensor in
    topo = rg.topology
    with pytest.raises(TypeError):
        RegridFunction.apply(F32, tree.d_vertices, topo.tgt_vertices, topo)
    with pytest.raises(ValueError):
        RegridFunction.apply(F64.cpu(), tree.d_vertices, topo.tgt_vertices, topo)
    with pytest.raises(ValueError):
        rg(F64[:-1])
    with pytest.raises(ValueError):
        rg(F64, src_vertices=tree.d_vertices[:-1])
    with pytest.raises(ValueError):
        rg(F64, tgt_vertices=topo.tgt_vertices[:, :1])

    v, f, _ = S.triangle_grid_20k()
    tri = M.CellTree2d(v, f, -1)
    pts = torch.from_numpy(np.random.default_rng(2).uniform(0.5, 9.5, (1000, 2))).cuda()
    face, w = barycentric_weights(tri, pts)
    face32, w32 = barycentric_weights(tri, pts.float())
    face_h, w_h = barycentric_weights(tri, pts.cpu().numpy())
    assert torch.equal(face_h, face) and torch.equal(w_h, w)
    assert torch.equal(w32, barycentric_weights(tri, pts.float().double())[1])
    with pytest.raises(ValueError):
        barycentric_weights(tri, pts, vertices=tri.d_vertices[:-1])
    with pytest.raises(TypeError):
        BarycentricFunction.apply(pts.float(), tri.d_vertices, tri.d_elements, face, 1e-12, False)


# ------------------------------------------------------------- operator assembly
def test_weights_match_oracle_and_dense_example():
    """Regridder.weights(): the CSR operator's (row, col) lists equal the oracle's
    intersect_faces pairs, its values are the oracle's areas over the target area (or over
    the overlap sum), and applying it (W.apply, torch sparse, scipy) equals the dense
    weight-matrix recipe of the reference's example (examples/spatial_indexing_2d_grids.py:
    268-298): np.dot(weight_matrix, z) / area_sum, NaN where nothing overlaps."""
    import torch

    from numba_celltree_amd.regrid import Regridder

    sv, sf, tv, tf, F = problem(41, 20, 10.0, n_fields=3)
    tv = tv * 1.15 - 0.5  # targets past the source mesh: partly covered and uncovered cells
    tree, rg = build(sv, sf, tv, tf)
    o = R.OracleCellTree2d(sv, sf, -1)
    i, j, area = o.intersect_faces(tv, tf, -1)
    ccw = tf.astype(np.int64).copy()
    R.counter_clockwise(tv, ccw)
    tarea = np.array([R.polygon_area(tv[row[row >= 0]]) for row in ccw])

    W = rg.weights()
    rows, cols, vals = (x.cpu().numpy() for x in W.coo())
    assert np.array_equal(rows, i) and np.array_equal(cols, j)
    np.testing.assert_allclose(vals, area / tarea[i], rtol=1e-12)
    assert W.shape == (len(tf), len(sf))

    dense = np.zeros((len(tf), len(sf)))
    dense[i, j] = area
    z = F[:, 0]
    area_sum = dense @ np.ones(len(sf))
    want = np.full(len(tf), np.nan)
    hit = area_sum > 0
    want[hit] = (dense @ z)[hit] / area_sum[hit]
    assert (~hit).any() and (area_sum[hit] < tarea[hit] * (1 - 1e-6)).any()

    mean = Regridder(tree, tv, tf, -1, normalize="overlap")
    got = mean(torch.from_numpy(z).cuda()).cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_allclose(got[hit], want[hit], rtol=1e-12)
    Wm = mean.weights()
    np.testing.assert_allclose(Wm.apply(z).cpu().numpy()[hit], want[hit], rtol=1e-12)
    np.testing.assert_allclose(Wm.to_scipy() @ z, np.nan_to_num(want, nan=0.0), rtol=1e-12, atol=1e-300)
    # conservative form: divided by the target's own area
    y = rg(torch.from_numpy(F).cuda()).cpu().numpy()
    np.testing.assert_allclose(y, (dense @ F) / tarea[:, None], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(W.apply(F).cpu().numpy(), y, rtol=1e-12, atol=1e-14)
    ts = (W.to_torch() @ torch.from_numpy(F).cuda()).cpu().numpy()
    np.testing.assert_allclose(ts, y, rtol=1e-12, atol=1e-14)
    raw = Regridder(tree, tv, tf, -1, normalize="none")
    np.testing.assert_allclose(raw(torch.from_numpy(F).cuda()).cpu().numpy(), dense @ F, rtol=1e-12, atol=1e-14)


def test_areas_cached_while_vertices_unchanged(monkeypatch):
    """With the same, unmodified vertex tensors the forward reuses the pair areas (the clip
    kernel does not run again); an in-place change of either vertex array recomputes them,
    and the result equals a regridder built on the changed mesh."""
    import torch

    from numba_celltree_amd import regrid as RG

    sv, sf, tv, tf, F = problem(33, 12, 4.0, n_fields=2, inset=0.0731)
    tree, rg = build(sv, sf, tv, tf)
    topo = rg.topology
    calls = []
    real = RG.RegridTopology.pair_areas

    def counted(self, a, b):
        calls.append(1)
        return real(self, a, b)

    monkeypatch.setattr(RG.RegridTopology, "pair_areas", counted)
    Fd = torch.from_numpy(F).cuda()
    svd = tree.d_vertices.clone()
    tvd = topo.tgt_vertices.clone()
    y0 = rg(Fd, svd, tvd)
    y1 = rg(Fd, svd, tvd)
    rg(Fd)
    rg(Fd)
    assert len(calls) == 2 and torch.equal(y0, y1)
    with torch.no_grad():
        svd[5] += 0.01  # in place: version counter moves
    y2 = rg(Fd, svd, tvd)
    assert len(calls) == 3 and not torch.equal(y2, y0)
    fresh = RG.Regridder(tree, tv, tf, -1).topology
    a = real(fresh, svd, tvd)  # not counted
    assert torch.equal(fresh.apply(a, fresh.target_areas(tvd), Fd), y2)
    topo.clear_cache()
    rg(Fd, svd, tvd)
    assert len(calls) == 4
    # with a vertex gradient requested the areas are always recomputed (gradcheck, for one,
    # perturbs its inputs through .data, which the version counter does not see)
    svg = svd.clone().requires_grad_(True)
    rg(Fd, svg, tvd)
    rg(Fd, svg, tvd)
    assert len(calls) == 6
