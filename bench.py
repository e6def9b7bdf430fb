"""Benchmark: locate_points throughput at ~1M cells on one MI355X (plus secondary timings).

Workload (the headline metric): the mixed mesh of 999,941 cells (867x867 jittered vertex
grid on [0,1000]^2, a third of the quads split into triangles, int32 faces of width 4) and
1e7 float64 query points (95% uniform inside, 5% in a 50-wide frame outside).  One step =
one ``CellTree2d.locate_points`` call on the device-resident points.  Inputs are in HBM
before the timed region; the tree is built before it.

Prints ONE JSON line.  ``roofline`` uses the compulsory bytes of one launch,
N*(16 in + 8 out) + 32*n_nodes + 4*n_cells (bb_indices) + 4*k*n_cells (faces) +
16*n_vertices, over the average kernel time measured with HIP events on the launch stream.
``cpu_baseline`` times the CPU oracle (a C++/OpenMP restatement of the reference) on the
host's cores over the same points.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def time_calls(fn, steps: int):
    """Wall time of `steps` calls plus the per-call event time on the current stream."""
    import torch

    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(st)
        fn()
        b.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) for a, b in ev]
    return wall, float(np.mean(per))


def oracle_native():
    """The CPU oracle at the highest ISA level this host executes (same results at every
    level), and a description of the host for the baseline records."""
    from oracle import refcpu as R

    level = R.native_isa_level()
    R.use_isa_level(level)
    model, _ = R.host_cpu()
    return R, {"cores": int(R.lib().ref_num_threads()), "cpu": model, "isa": level}


def median_time(fn, runs: int = 5):
    """One warm-up call, then the median wall time of `runs` calls (seconds)."""
    fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), ts


def cpu_baseline(v, f, pts, runs: int = 5):
    R, host = oracle_native()
    tree = R.OracleCellTree2d(v, f, -1)
    t, ts = median_time(lambda: tree.locate_points(pts), runs)
    return {
        "value": len(pts) / t,
        "unit": "points/s",
        **host,
        "kind": "port",
        "sample": f"the whole workload ({len(pts):,} points, oracle tree of the same mesh): one warm-up, "
                  f"median of {runs} runs ({min(ts):.3f}-{max(ts):.3f} s)",
    }


def flop_counts():
    """The committed FP64 operation counts of the overlap workloads (oracle-counted)."""
    with open(os.path.join(ROOT, "tests", "golden", "flops_overlap.json")) as fh:
        return json.load(fh)


def fp64_roofline(kernel: str, flops: int, kernel_ms: float, note: str):
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    return {
        "bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": None, "kernel": kernel,
        "kernel_ms": round(kernel_ms, 4), "flops": int(flops), "flops_source": note,
    }


def extras(args):
    """Secondary workloads: a warm-up call each, then the mean over several calls
    (device-resident inputs)."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import _device as D
    from numba_celltree_amd import _native
    from numba_celltree_amd import synthetic as S
    from numba_celltree_amd.cast import cast_faces

    out = {}
    counts = flop_counts()
    v, f = S.mixed_mesh_1m()
    _, bx, sg = S.mixed_mesh_queries(10, args.boxes, args.segments)
    tree = M.CellTree2d(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda(), -1)
    d_bx = torch.from_numpy(bx).cuda()
    d_sg = torch.from_numpy(sg).cuda()
    for name, fn in (("locate_boxes_1e6_ms", lambda: tree.locate_boxes(d_bx)),
                     ("intersect_edges_1e6_ms", lambda: tree.intersect_edges(d_sg))):
        fn()
        wall, _ = time_calls(fn, 5)
        out[name] = 1e3 * wall / 5
    i, _ = tree.locate_boxes(d_bx)
    out["locate_boxes_pairs"] = int(i.numel())
    i, _, _ = tree.intersect_edges(d_sg)
    out["intersect_edges_rows"] = int(i.numel())

    # Regrid (16 fields) from the 999,941-cell mixed mesh to 250,000 hex-Voronoi cells:
    # the plain forward against forward + hand-written adjoint (dF, d src vertices,
    # d target vertices), topology fixed.
    from numba_celltree_amd.regrid import Regridder

    tv, tf = S.hex_voronoi(500, 500, 1000.0, 0.1, 7)
    F = torch.from_numpy(S.sinusoid_fields(v[np.maximum(f, 0)].mean(1), 16, 1)).cuda()
    d_tv = torch.from_numpy(tv).cuda()
    d_tf = torch.from_numpy(tf).cuda()
    for label in ("regrid_topology_cold_ms", "regrid_topology_ms"):
        t0 = time.perf_counter()
        rg = Regridder(tree, d_tv, d_tf, -1)
        torch.cuda.synchronize()
        out[label] = 1e3 * (time.perf_counter() - t0)
    topo = rg.topology
    out["regrid_pairs"] = topo.n_pairs
    sv_t = tree.d_vertices.clone()
    tv_t = topo.tgt_vertices.clone()
    G = torch.randn(topo.n_tgt, 16, dtype=torch.float64, device="cuda")

    def fwd():
        # the full forward: the pair areas are recomputed (the area cache is dropped first)
        topo.clear_cache()
        with torch.no_grad():
            return rg(F, sv_t, tv_t)

    def fwd_cached():
        # fixed vertices: the cached pair areas are reused, only the operator is applied
        with torch.no_grad():
            return rg(F, sv_t, tv_t)

    def fwd_bwd():
        # G is the output gradient (the gradient of sum(y * G)), handed to the backward
        # directly so the leg times the regrid forward and backward only, no loss kernels.
        Fg = F.detach().requires_grad_(True)
        svg = sv_t.detach().requires_grad_(True)
        tvg = tv_t.detach().requires_grad_(True)
        y = rg(Fg, svg, tvg)
        y.backward(G)
        return Fg.grad, svg.grad, tvg.grad

    for name, fn in (("regrid_fwd_16f_ms", fwd), ("regrid_fwd_bwd_16f_ms", fwd_bwd),
                     ("regrid_apply_cached_16f_ms", fwd_cached)):
        fn()
        wall, _ = time_calls(fn, 20)
        out[name] = 1e3 * wall / 20
    out["regrid_fwd_bwd_over_fwd"] = out["regrid_fwd_bwd_16f_ms"] / out["regrid_fwd_16f_ms"]
    t0 = time.perf_counter()
    W = rg.weights()
    torch.cuda.synchronize()
    out["regrid_weights_ms"] = 1e3 * (time.perf_counter() - t0)
    fn = lambda: W.apply(F)  # noqa: E731
    fn()
    wall, _ = time_calls(fn, 20)
    out["regrid_weights_apply_16f_ms"] = 1e3 * wall / 20
    # the pair-area kernel of the forward on its own (HIP events on the launch stream)
    pa = lambda: topo.pair_areas(sv_t, tv_t)  # noqa: E731
    pa()
    _, pa_ms = time_calls(pa, 20)
    # the vertex backward on its own (k_pair_backward, the corner sums and the vertex gathers;
    # HIP events on the launch stream), with the forward's outputs as saved by autograd
    with torch.no_grad():
        y_fb = rg(F, sv_t, tv_t)
    tarea = topo.target_areas(tv_t)
    vb = lambda: topo.vertex_backward(sv_t, tv_t, F, y_fb, G, tarea)  # noqa: E731
    vb()
    _, vb_ms = time_calls(vb, 20)
    out["regrid_vertex_backward_16f_ms"] = vb_ms
    rc = counts["regrid"]
    if rc["kept"] == topo.n_pairs:
        out["regrid_roofline"] = fp64_roofline(
            "k_pair_areas (regrid forward clip)", rc["clip_flops_kept"], pa_ms,
            "tests/golden/flops_overlap.json regrid.clip_flops_kept: Sutherland-Hodgman + polygon_area ops "
            "of every kept pair (oracle-counted)")
        # forward + backward: the forward's pair-area kernel plus the whole vertex backward;
        # counted are the clip ops (forward) and, per pair, the 16-field dot product with its
        # division and the overlap-area adjoint (backward); the backward's corner and vertex
        # sums are not counted (a few adds per pair)
        out["regrid_fwd_bwd_roofline"] = fp64_roofline(
            "k_pair_areas + vertex backward (k_pair_backward, k_target_corners, k_source_corners, "
            "k_vertex_gather)", rc["clip_flops_kept"] + rc["backward_pair_flops"] + rc["backward_ga_flops_16f"],
            pa_ms + vb_ms,
            "tests/golden/flops_overlap.json regrid: clip_flops_kept + backward_pair_flops (oracle "
            "ref_pair_area_gradient) + backward_ga_flops_16f")
    del tree, rg, topo, W, y_fb, tarea
    # compute_barycentric_weights with its adjoint: 1e6 points on the 20,000-triangle mesh
    # (points uniform over [-0.5, 10.5]^2; faces located once, then weights forward / backward).
    from numba_celltree_amd.barycentric import barycentric_weights

    v1, f1, _ = S.triangle_grid_20k()
    tri = M.CellTree2d(torch.from_numpy(v1).cuda(), torch.from_numpy(f1).cuda(), -1)
    bp = torch.from_numpy(np.random.default_rng(4).uniform(-0.5, 10.5, size=(1_000_000, 2))).cuda()
    Gb = torch.randn(1_000_000, 3, dtype=torch.float64, device="cuda")

    def bary_fwd():
        with torch.no_grad():
            return barycentric_weights(tri, bp)

    def bary_fwd_bwd():
        pg = bp.detach().requires_grad_(True)
        vg = tri.d_vertices.detach().requires_grad_(True)
        _, w = barycentric_weights(tri, pg, vg)
        w.backward(Gb)
        return pg.grad, vg.grad

    for name, fn in (("barycentric_1e6_fwd_ms", bary_fwd), ("barycentric_1e6_fwd_bwd_ms", bary_fwd_bwd),
                     ("compute_barycentric_weights_1e6_ms", lambda: tri.compute_barycentric_weights(bp))):
        fn()
        wall, _ = time_calls(fn, 10)
        out[name] = 1e3 * wall / 10
    del tri, bp, Gb
    sv, sf, tv, tf, _ = S.overlap_meshes()
    for label in ("build_999698_tris_cold_ms", "build_999698_tris_ms"):
        t0 = time.perf_counter()
        src = M.CellTree2d(torch.from_numpy(sv).cuda(), torch.from_numpy(sf).cuda(), -1)
        torch.cuda.synchronize()
        out[label] = 1e3 * (time.perf_counter() - t0)
    d_tv = torch.from_numpy(tv).cuda()
    d_tf = torch.from_numpy(tf).cuda()
    fn = lambda: src.intersect_faces(d_tv, d_tf, -1)  # noqa: E731
    fn()
    wall, _ = time_calls(fn, 5)
    out["intersect_faces_1Mx250k_ms"] = 1e3 * wall / 5
    out["intersect_faces_pairs"] = int(fn()[0].numel())
    # The pair stage (fused SAT + clip + area, k_face_pairs) on the same candidate pairs,
    # timed on its own with HIP events on the launch stream.
    d_faces = D.faces_to_dev(cast_faces(d_tf, -1), d_tv.shape[0])
    ii, jj = src._locate_boxes_dev(src._prepare_faces(d_tv, d_faces))
    n = int(ii.numel())
    area = D.empty(n, torch.float64)
    keep = D.empty(n, torch.int64)

    def pairs():
        _native.call("ct_face_pairs", D.ptr(d_tv), D.ptr(d_faces), int(d_faces.shape[1]), D.ptr(src.d_vertices),
                     D.ptr(src.d_elements), int(src.d_elements.shape[1]), D.ptr(ii), D.ptr(jj), n, 1, D.ptr(area),
                     D.ptr(keep), D.stream())

    pairs()
    _, fp_ms = time_calls(pairs, 20)
    fc = counts["intersect_faces"]
    out["intersect_faces_candidates"] = n
    if fc["candidates"] == n and fc["kept"] == int(keep.sum()):
        out["intersect_faces_roofline"] = fp64_roofline(
            "k_face_pairs (SAT + Sutherland-Hodgman + area)", fc["sat_flops"] + fc["clip_flops"], fp_ms,
            "tests/golden/flops_overlap.json intersect_faces.sat_flops + clip_flops: the ops the reference "
            "executes on these candidate pairs (oracle-counted)")
    if not args.no_cpu_baseline:
        # CPU oracle on the whole workload (all 250,000 target cells).
        R, host = oracle_native()
        osrc = R.OracleCellTree2d(sv, sf, -1)
        cpu_s, ts = median_time(lambda: osrc.intersect_faces(tv, tf, -1), 5)
        out["intersect_faces_cpu_baseline"] = {
            "value_ms": 1e3 * cpu_s, **host, "kind": "port",
            "sample": f"the whole workload ({len(tf):,} target cells x {len(sf):,} source cells): one warm-up, "
                      f"median of 5 runs ({min(ts):.3f}-{max(ts):.3f} s)",
        }
        out["intersect_faces_speedup_vs_cpu"] = 1e3 * cpu_s / out["intersect_faces_1Mx250k_ms"]
    return out


def batch_extras(args):
    """512 independent meshes, each with its own points and segments: a loop of plain
    per-mesh calls against one HIP-graph replay of the batched persistent kernels."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import batch as B
    from numba_celltree_amd import synthetic as S

    out = {}
    t0 = time.perf_counter()
    sets = S.batch_meshes(n_meshes=args.batch_meshes)
    log(f"batch workload generated in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    trees = [M.CellTree2d(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev), -1) for v, f, _, _ in sets]
    torch.cuda.synchronize()
    out["batch_build_trees_ms"] = 1e3 * (time.perf_counter() - t0)
    # The same 512 trees exported once (.npz in memory) and imported without a rebuild:
    # what a caller that builds its trees once pays per run instead of the build.
    import io

    blobs = []
    for t in trees:
        buf = io.BytesIO()
        t.save(buf)
        blobs.append(buf.getvalue())
    out["batch_saved_trees_mb"] = sum(len(b) for b in blobs) / 1e6
    t0 = time.perf_counter()
    loaded = [M.CellTree2d.load(io.BytesIO(b)) for b in blobs]
    torch.cuda.synchronize()
    out["batch_load_trees_ms"] = 1e3 * (time.perf_counter() - t0)
    out["batch_loaded_trees_identical"] = bool(all(
        np.array_equal(a.nodes, b.nodes) and np.array_equal(a.bb_indices, b.bb_indices) for a, b in zip(trees, loaded)))
    trees = loaded  # the timed legs below run on the imported trees
    del blobs
    d_pts = [torch.from_numpy(p).to(dev) for _, _, p, _ in sets]
    d_seg = [torch.from_numpy(s).to(dev) for _, _, _, s in sets]
    out["batch_meshes"] = len(sets)
    out["batch_cells"] = int(sum(len(f) for _, f, _, _ in sets))
    out["batch_points"] = int(sum(len(p) for _, _, p, _ in sets))
    out["batch_segments"] = int(sum(len(s) for _, _, _, s in sets))

    def plain_points():
        return [t.locate_points(p) for t, p in zip(trees, d_pts)]

    def plain_edges():
        return [t.intersect_edges(s) for t, s in zip(trees, d_seg)]

    def plain_both():
        return plain_points(), plain_edges()

    for name, fn in (("batch_plain_points_ms", plain_points), ("batch_plain_edges_ms", plain_edges),
                     ("batch_plain_ms", plain_both)):
        fn()
        wall, _ = time_calls(fn, 3)
        out[name] = 1e3 * wall / 3
    want_pts = plain_points()
    want_edges = plain_edges()

    tb = B.TreeBatch(trees)
    pts, pc = B.concat_queries(d_pts, 2)
    segs, sc = B.concat_queries(d_seg, 4)
    segs = segs.reshape(-1, 2, 2)
    t0 = time.perf_counter()
    cfg_p, tp = B.autotune(tb, "points", pts, pc, reps=2)
    cfg_e, te = B.autotune(tb, "edges", segs, sc, reps=2)
    out["batch_autotune_s"] = time.perf_counter() - t0
    out["batch_config_points"] = cfg_p.__dict__
    out["batch_config_edges"] = cfg_e.__dict__
    out["batch_autotune_points_ms"] = {"best": min(ms for _, ms in tp), "worst": max(ms for _, ms in tp)}
    out["batch_autotune_edges_count_ms"] = {"best": min(ms for _, ms in te), "worst": max(ms for _, ms in te)}
    graphs = {
        "batch_graph_points_ms": B.BatchGraph(tb, pts, pc, points_config=cfg_p),
        "batch_graph_edges_ms": B.BatchGraph(tb, segments=segs, segment_counts=sc, edges_config=cfg_e),
        "batch_graph_ms": B.BatchGraph(tb, pts, pc, segs, sc, points_config=cfg_p, edges_config=cfg_e),
    }
    for name, g in graphs.items():
        g.replay()
        wall, _ = time_calls(g.replay, 5)
        out[name] = 1e3 * wall / 5
    res = graphs["batch_graph_ms"].result()
    ro = res.row_offsets()
    ok = all(torch.equal(res.points_of(m), want_pts[m]) for m in range(len(trees)))
    for m in range(len(trees)):
        gi, gj, gxy = res.edges_of(m, ro)
        wi, wj, wxy = want_edges[m]
        ok = ok and torch.equal(gi, wi) and torch.equal(gj, wj) and torch.equal(gxy, wxy)
    out["batch_identical_to_plain"] = bool(ok)
    out["batch_speedup_points"] = out["batch_plain_points_ms"] / out["batch_graph_points_ms"]
    out["batch_speedup_edges"] = out["batch_plain_edges_ms"] / out["batch_graph_edges_ms"]
    out["batch_speedup"] = out["batch_plain_ms"] / out["batch_graph_ms"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=10_000_000)
    ap.add_argument("--boxes", type=int, default=1_000_000)
    ap.add_argument("--segments", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--batch-meshes", type=int, default=512)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--only", choices=["all", "locate", "extras", "batch"], default="all",
                    help="locate: the headline workload only; extras: the secondary workloads only; "
                         "batch: the many-mesh batch only")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run three 1 GiB device copies and exit (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU; --gpus {args.gpus} is not supported (queries shard by "
            "replicating the tree per GPU, which this single-GPU build does not launch)",
            file=sys.stderr,
        )
        return 2

    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import synthetic as S

    torch.cuda.set_device(0)
    if args.calibrate_copy:
        # A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE.
        nbytes = 1 << 30
        a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}), flush=True)
        return 0
    if args.only == "extras":
        print(json.dumps({"extras": extras(args)}), flush=True)
        return 0
    if args.only == "batch":
        print(json.dumps({"batch": batch_extras(args)}), flush=True)
        return 0
    log("generating workload")
    v, f = S.mixed_mesh_1m()
    pts, _, _ = S.mixed_mesh_queries(args.points, 10, 10)
    d_v = torch.from_numpy(v).cuda()
    d_f = torch.from_numpy(f).cuda()
    d_pts = torch.from_numpy(pts).cuda()
    torch.cuda.synchronize()
    build = {}
    for label in ("build_cold_ms", "build_ms"):  # the first call also loads the code objects
        t0 = time.perf_counter()
        tree = M.CellTree2d(d_v, d_f, -1)
        torch.cuda.synchronize()
        build[label] = 1e3 * (time.perf_counter() - t0)
    log(f"tree: {len(f)} cells, {tree.n_nodes} nodes, depth {tree.depth}, build {build['build_ms']:.1f} ms "
        f"(cold {build['build_cold_ms']:.1f} ms)")

    from numba_celltree_amd import _device as D

    tol = tree._default_tolerance()
    ordered = tree.ORDERED_MIN is not None and args.points >= tree.ORDERED_MIN

    def step(ev=None):
        # One locate_points call, split so the traversal kernel can be timed on its own:
        # [ev0] spatial order (3 sort kernels + scan) [ev1] traversal kernel [ev2].
        if ev:
            ev[0].record()
        order = D.query_order(0, d_pts, tree._bbox) if ordered else None
        if ev:
            ev[1].record()
        out = tree._locate_points_dev(d_pts, tol, order)
        if ev:
            ev[2].record()
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for ev in evs:
        step(ev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    order_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in evs]))
    kernel_ms = float(np.mean([b.elapsed_time(c) for _, b, c in evs]))
    ms_per_step = 1e3 * wall / args.steps
    value = args.steps * args.points / wall
    fn = step

    n_cells, k = f.shape
    bytes_per_launch = args.points * 24 + 32 * tree.n_nodes + 4 * n_cells + 4 * k * n_cells + 16 * len(v)
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": read_traffic("locate_points_1e7_traffic.txt"),
        "kernel": "k_locate_points (traversal)",
        "kernel_ms": round(kernel_ms, 4),
        "order_ms": round(order_ms, 4),
        "algorithmic_bytes": int(bytes_per_launch),
    }

    # spot parity on a sample, outside the timed region
    res = fn()[: 200_000].cpu().numpy()
    result = {
        "metric": "locate_points throughput at ~1M cells (points/s)",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded numpy generators; no dataset)",
        "config": {
            "workload": "locate_points on a 999,941-cell mixed tri/quad mesh, 1e7 float64 points",
            "cells": int(n_cells),
            "vertices": int(len(v)),
            "points": args.points,
            "tree_nodes": tree.n_nodes,
            "tree_depth": tree.depth,
            "n_buckets": 4,
            "cells_per_leaf": 2,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "build_ms": round(build["build_ms"], 2),
        "build_cold_ms": round(build["build_cold_ms"], 2),
        "inside_fraction": float((res >= 0).mean()),
    }
    if args.only == "locate":
        args.no_cpu_baseline = args.no_extras = True
    if not args.no_cpu_baseline:
        log("cpu baseline")
        from oracle import refcpu as R

        want = R.OracleCellTree2d(v, f, -1).locate_points(pts[:200_000])
        result["spot_parity_200k"] = bool(np.array_equal(res, want))
        result["cpu_baseline"] = cpu_baseline(v, f, pts)
        result["speedup_vs_cpu_baseline"] = value / result["cpu_baseline"]["value"]
    if not args.no_extras:
        log("extras")
        del tree
        torch.cuda.empty_cache()
        result["extras"] = extras(args)
        if not args.no_batch:
            log("batch")
            torch.cuda.empty_cache()
            result["batch"] = batch_extras(args)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
