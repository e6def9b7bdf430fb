"""Batched persistent launch (numba_celltree_amd.batch) against per-tree calls and the oracle.

Bar: every tree's answers are bit-identical to that tree's own CellTree2d call (which the
parity tests pin to the oracle), for several launch shapes, through HIP-graph replays, and
for empty and one-query sets; a few trees are also checked against the oracle directly.
"""

import sys

import numpy as np
import pytest

from numba_celltree_amd import synthetic as S
from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[gpu batch] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def meshes():
    import numba_celltree_amd as M

    sets = S.batch_meshes(n_meshes=14, seed=3, cells=(200, 6000), n_points=(50, 20_000), n_segments=(10, 3000))
    # a triangle-only mesh (3 columns) among the mixed ones, an empty and a one-query set
    v, f, p = S.triangle_grid_20k(seed=2)
    seg = S.segments(500, 10.0, 0.1, 0.1, 50.0, 2)
    sets.append((v, f, p[:7000], seg))
    sets[2] = (sets[2][0], sets[2][1], sets[2][2][:0], sets[2][3][:1])
    sets[5] = (sets[5][0], sets[5][1], sets[5][2][:1], sets[5][3][:0])
    trees = [M.CellTree2d(v, f, -1) for v, f, _, _ in sets]
    log(f"{len(trees)} trees, depths {[t.depth for t in trees]}")
    return sets, trees


CONFIGS = [
    dict(),
    dict(block=64, wg_per_cu=2, tile=64, n_queues=1),
    dict(block=256, wg_per_cu=1, tile=1024, n_queues=3),
    dict(block=128, wg_per_cu=8, tile=192, n_queues=8, stack=128),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: ",".join(f"{k}={v}" for k, v in c.items()) or "default")
def test_batch_locate_points(meshes, cfg):
    import torch

    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    log(f"points {cfg}")
    got = B.locate_points(tb, [p for _, _, p, _ in sets], B.BatchConfig(**cfg))
    for m, (t, (_, _, p, _)) in enumerate(zip(trees, sets)):
        want = t.locate_points(torch.from_numpy(np.ascontiguousarray(p)).cuda()) if len(p) else None
        if want is None:
            assert got[m].numel() == 0
        else:
            assert torch.equal(got[m], want), m


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: ",".join(f"{k}={v}" for k, v in c.items()) or "default")
def test_batch_intersect_edges(meshes, cfg):
    import torch

    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    log(f"edges {cfg}")
    got = B.intersect_edges(tb, [s for _, _, _, s in sets], B.BatchConfig(**cfg))
    for m, (t, (_, _, _, s)) in enumerate(zip(trees, sets)):
        gi, gj, gxy = got[m]
        if len(s) == 0:
            assert gi.numel() == 0
            continue
        wi, wj, wxy = t.intersect_edges(torch.from_numpy(np.ascontiguousarray(s)).cuda())
        assert torch.equal(gi, wi) and torch.equal(gj, wj), m
        assert torch.equal(gxy.view(torch.int64), wxy.view(torch.int64)), m


@pytest.mark.parametrize("slots", [4, 0])
def test_batch_edges_hit_record_sizes(meshes, monkeypatch, slots):
    """A small hit record (4 slots: most long segments walk again) and none at all give the
    same rows as the default record."""
    import torch

    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    segs = [s for _, _, _, s in sets]
    want = B.intersect_edges(tb, segs)
    monkeypatch.setattr(B, "EDGE_HIT_SLOTS", slots)
    log(f"edges hit record slots={slots}")
    got = B.intersect_edges(tb, segs)
    for m in range(len(sets)):
        for a, b in zip(got[m], want[m]):
            assert torch.equal(a.view(torch.int64) if a.is_floating_point() else a,
                               b.view(torch.int64) if b.is_floating_point() else b), m


def test_batch_matches_oracle(meshes):
    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    pts = B.locate_points(tb, [p for _, _, p, _ in sets])
    edges = B.intersect_edges(tb, [s for _, _, _, s in sets])
    for m in (0, 3, len(sets) - 1):
        v, f, p, s = sets[m]
        o = R.OracleCellTree2d(v, f, -1)
        assert np.array_equal(pts[m].cpu().numpy(), o.locate_points(p))
        oi, oj, oxy = o.intersect_edges(s)
        gi, gj, gxy = (x.cpu().numpy() for x in edges[m])
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
        assert np.allclose(gxy, oxy, rtol=1e-12, atol=0.0, equal_nan=True)


def test_batch_graph_replays(meshes):
    """A captured step replayed several times, with eager work and new query values between
    replays, matches the eager batch every time."""
    import torch

    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    pts, pc = B.concat_queries([p for _, _, p, _ in sets], 2)
    segs, sc = B.concat_queries([s for _, _, _, s in sets], 4)
    segs = segs.reshape(-1, 2, 2)
    g = B.BatchGraph(tb, pts, pc, segs, sc)
    want_p = B.locate_points(tb, [p for _, _, p, _ in sets])
    want_e = B.intersect_edges(tb, [s for _, _, _, s in sets])
    rng = np.random.default_rng(0)
    for rep in range(4):
        log(f"graph replay {rep}")
        if rep == 2:
            # move every point to a new place; the graph reads the same buffer
            pts.copy_(pts[torch.from_numpy(rng.permutation(len(pts))).cuda()])
            want_p = None
        g.replay()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        torch.cuda.synchronize()
        res = g.result()
        if want_p is None:
            # new points: compare with per-tree calls on the moved points
            q = g.pplan.query_offsets
            for m, t in enumerate(trees):
                if q[m + 1] > q[m]:
                    assert torch.equal(res.points_of(m), t.locate_points(pts[q[m]: q[m + 1]].contiguous()))
        else:
            for m in range(len(trees)):
                assert torch.equal(res.points_of(m), want_p[m])
        ro = res.row_offsets()
        for m in range(len(trees)):
            gi, gj, gxy = res.edges_of(m, ro)
            wi, wj, wxy = want_e[m]
            assert torch.equal(gi, wi) and torch.equal(gj, wj)
            assert torch.equal(gxy.view(torch.int64), wxy.view(torch.int64))


def test_batch_graph_overflow_is_reported(meshes):
    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    segs, sc = B.concat_queries([s for _, _, _, s in sets], 4)
    segs = segs.reshape(-1, 2, 2)
    g = B.BatchGraph(tb, segments=segs, segment_counts=sc, capacity=10)
    g.replay()
    assert g.overflowed()
    with pytest.raises(RuntimeError):
        g.result()


def test_autotune_picks_a_candidate(meshes):
    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    pts, pc = B.concat_queries([p for _, _, p, _ in sets], 2)
    cands = [B.BatchConfig(block=64, tile=64), B.BatchConfig(block=256, tile=256)]
    best, timings = B.autotune(tb, "points", pts, pc, candidates=cands, reps=2)
    assert best in cands and len(timings) == 2 and all(ms > 0 for _, ms in timings)


def test_batch_rejects_bad_input(meshes):
    import torch

    from numba_celltree_amd import batch as B

    sets, trees = meshes
    tb = B.TreeBatch(trees)
    with pytest.raises(ValueError):
        tb.plan(np.ones(len(trees) + 1, dtype=np.int64), B.BatchConfig())
    with pytest.raises(ValueError):
        tb.plan(np.ones(len(trees), dtype=np.int64), B.BatchConfig(tile=100))
    with pytest.raises(ValueError):
        tb.plan(np.ones(len(trees), dtype=np.int64), B.BatchConfig(stack=16 if tb.max_depth > 16 else 8))
    plan = tb.plan(np.ones(len(trees), dtype=np.int64), B.BatchConfig())
    with pytest.raises(ValueError):  # wrong length
        plan.locate_points(torch.zeros((len(trees) + 1, 2), dtype=torch.float64, device="cuda"),
                           torch.zeros(len(trees), dtype=torch.int64, device="cuda"))
