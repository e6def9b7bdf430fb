"""The regrid adjoint against the oracle, on meshes that exercise every path of the vertex
backward.

The vertex backward finishes the targets whose pairs share one workgroup block of 256 pairs
in LDS (k_pair_backward) and the others from per-pair rows (k_target_corners), so the meshes
here mix both: a fine target mesh (about 10 pairs per target, some targets across block
edges) and a coarse one over a fine source (hundreds of pairs per target, every target across
several blocks), partly outside the source (targets without pairs).  The gradients must not
depend on which inputs require them (F only, vertices only, all three: the same bits).
Against the oracle, per vertex:

* source vertices: sum over pairs of gA_k dA_k/d(source vertices), gA_k = G_t . F_s / area_t,
  with dA_k from the oracle's adjoint restatement (ref_pair_area_gradient, pinned by finite
  differences of the oracle's clip area);
* target vertices: the same with dA_k/d(target vertices), plus the target-area term
  -(G_t . y_t) / area_t times the area's derivative, which for a convex counter-clockwise
  polygon is ((y_{i+1} - y_{i-1}) / 2, (x_{i-1} - x_{i+1}) / 2) at vertex i;
* F: gF_s = sum over the source's pairs of (A_k / area_t) G_t.
"""

import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def _meshes(case):
    if case == "fine":
        sv, sf = S.jittered_grid(61, 10.0, 0.25, 1.0, 0)
        tv, tf = S.hex_voronoi(24, 24, 10.0, 0.1, 1)
    else:  # coarse targets over a fine source, shifted partly outside it
        sv, sf = S.jittered_grid(121, 10.0, 0.25, 1.0, 2)
        tv, tf = S.hex_voronoi(4, 4, 14.0, 0.1, 3)
        tv = tv - 5.0
    F = S.sinusoid_fields(sv[sf].mean(1) * 20.0, 16, 1)
    return sv, sf, tv, tf, F


def _grads(rg, F, sv, tv, G, which):
    import torch

    F = F.clone().requires_grad_("F" in which)
    sv = sv.clone().requires_grad_("v" in which)
    tv = tv.clone().requires_grad_("v" in which)
    y = rg(F, sv, tv)
    (y * G).sum().backward()
    return y.detach(), F.grad, sv.grad, tv.grad


@pytest.mark.parametrize("case", ["fine", "coarse"])
def test_backward_matches_oracle(case):
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd.regrid import Regridder

    sv, sf, tv, tf, F = _meshes(case)
    tree = M.CellTree2d(sv, sf, -1)
    rg = Regridder(tree, tv, tf, -1)
    topo = rg.topology
    counts = np.diff(topo.tgt_offsets.cpu().numpy())
    print(f"[regrid backward] {case}: {topo.n_pairs} pairs, {len(tf)} targets, pairs per target "
          f"{counts.min()}..{counts.max()}, {int((counts == 0).sum())} without pairs", file=sys.stderr, flush=True)
    if case == "coarse":
        assert counts.max() > 512 and (counts == 0).any()
    dev = "cuda"
    F_t = torch.from_numpy(F).to(dev)
    sv_t = torch.from_numpy(sv).to(dev)
    tv_t = topo.tgt_vertices.clone()
    G = np.random.default_rng(4).normal(size=(len(tf), F.shape[1]))
    G_t = torch.from_numpy(G).to(dev)

    y, gF, gsv, gtv = _grads(rg, F_t, sv_t, tv_t, G_t, "Fv")
    _, gF1, _, _ = _grads(rg, F_t, sv_t, tv_t, G_t, "F")
    _, _, gsv2, gtv2 = _grads(rg, F_t, sv_t, tv_t, G_t, "v")
    assert torch.equal(gF, gF1)
    assert torch.equal(gsv, gsv2)
    assert torch.equal(gtv, gtv2)

    t = topo.tt.cpu().numpy()
    s_ = topo.ss.cpu().numpy()
    tvh = tv_t.cpu().numpy()
    ccw = topo.tgt_faces.cpu().numpy().astype(np.int64)
    tarea = np.array([R.polygon_area(tvh[row[row >= 0]]) for row in ccw])
    faces = tree.faces.astype(np.int64)
    oi, oj, area = R.OracleCellTree2d(sv, sf, -1).intersect_faces(tvh, ccw, -1)
    assert np.array_equal(oi, t) and np.array_equal(oj, s_)
    gA = (G[t] * F[s_]).sum(1) / tarea[t]
    ga, gb, _ = R.pair_area_gradient(tvh, sv, ccw, faces, t, s_, gA)

    want_s = np.zeros_like(sv)
    for slot in range(faces.shape[1]):
        ok = faces[s_, slot] >= 0
        np.add.at(want_s, faces[s_[ok], slot], gb[ok, slot])
    want_t = np.zeros_like(tvh)
    for slot in range(ccw.shape[1]):
        ok = ccw[t, slot] >= 0
        np.add.at(want_t, ccw[t[ok], slot], ga[ok, slot])
    yh = y.cpu().numpy()
    garea = -(G * yh).sum(1) / tarea
    for ti, row in enumerate(ccw):
        idx = row[row >= 0]
        p = tvh[idx]
        nxt = np.roll(p, -1, 0)
        prv = np.roll(p, 1, 0)
        dA = 0.5 * np.column_stack([nxt[:, 1] - prv[:, 1], prv[:, 0] - nxt[:, 0]])
        np.add.at(want_t, idx, garea[ti] * dA)
    want_F = np.zeros_like(F)
    np.add.at(want_F, s_, (area / tarea[t])[:, None] * G[t])

    for name, got, want in (("src", gsv, want_s), ("tgt", gtv, want_t), ("F", gF, want_F)):
        got = got.cpu().numpy()
        scale = np.abs(want).max()
        print(f"[regrid backward] {case} {name}: max |grad| {scale:.3e}, max diff {np.abs(got - want).max():.3e}",
              file=sys.stderr, flush=True)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * scale)
