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


def cpu_baseline(v, f, pts, min_seconds: float = 10.0, max_reps: int = 3):
    from oracle import refcpu as R

    tree = R.OracleCellTree2d(v, f, -1)
    reps, total, t_all = 0, 0, 0.0
    while reps < max_reps and t_all < min_seconds:
        t = time.perf_counter()
        tree.locate_points(pts)
        t_all += time.perf_counter() - t
        total += len(pts)
        reps += 1
    return {
        "value": total / t_all,
        "unit": "points/s",
        "cores": int(R.lib().ref_num_threads()),
        "kind": "port",
        "sample": f"{reps} x {len(pts):,} points of the same workload (oracle tree of the same mesh)",
    }


def extras(args):
    """Secondary workloads, timed once each after a warm-up (device-resident inputs)."""
    import torch

    import numba_celltree_amd as M
    from numba_celltree_amd import synthetic as S

    out = {}
    v, f = S.config2_mesh()
    _, bx, sg = S.config2_queries(10, args.boxes, args.segments)
    tree = M.CellTree2d(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda(), -1)
    d_bx = torch.from_numpy(bx).cuda()
    d_sg = torch.from_numpy(sg).cuda()
    for name, fn in (("locate_boxes_1e6_ms", lambda: tree.locate_boxes(d_bx)),
                     ("intersect_edges_1e6_ms", lambda: tree.intersect_edges(d_sg))):
        fn()
        wall, _ = time_calls(fn, 3)
        out[name] = 1e3 * wall / 3
    i, _ = tree.locate_boxes(d_bx)
    out["locate_boxes_pairs"] = int(i.numel())
    i, _, _ = tree.intersect_edges(d_sg)
    out["intersect_edges_rows"] = int(i.numel())
    del tree
    sv, sf, tv, tf, _ = S.config3()
    t0 = time.perf_counter()
    src = M.CellTree2d(torch.from_numpy(sv).cuda(), torch.from_numpy(sf).cuda(), -1)
    torch.cuda.synchronize()
    out["build_999698_tris_ms"] = 1e3 * (time.perf_counter() - t0)
    d_tv = torch.from_numpy(tv).cuda()
    d_tf = torch.from_numpy(tf).cuda()
    fn = lambda: src.intersect_faces(d_tv, d_tf, -1)  # noqa: E731
    fn()
    wall, _ = time_calls(fn, 3)
    out["intersect_faces_1Mx250k_ms"] = 1e3 * wall / 3
    out["intersect_faces_pairs"] = int(fn()[0].numel())
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
    log("generating workload")
    v, f = S.config2_mesh()
    pts, _, _ = S.config2_queries(args.points, 10, 10)
    d_v = torch.from_numpy(v).cuda()
    d_f = torch.from_numpy(f).cuda()
    d_pts = torch.from_numpy(pts).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = M.CellTree2d(d_v, d_f, -1)
    torch.cuda.synchronize()
    build_ms = 1e3 * (time.perf_counter() - t0)
    log(f"tree: {len(f)} cells, {tree.n_nodes} nodes, depth {tree.depth}, build {build_ms:.1f} ms")

    fn = lambda: tree.locate_points(d_pts)  # noqa: E731
    for _ in range(args.warmup):
        fn()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    wall, kernel_ms = time_calls(fn, args.steps)
    ms_per_step = 1e3 * wall / args.steps
    value = args.steps * args.points / wall

    n_cells, k = f.shape
    bytes_per_launch = args.points * 24 + 32 * tree.n_nodes + 4 * n_cells + 4 * k * n_cells + 16 * len(v)
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": read_traffic("locate_points_cfg2_traffic.txt"),
        "kernel_ms": round(kernel_ms, 4),
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
        "build_ms": round(build_ms, 2),
        "inside_fraction": float((res >= 0).mean()),
    }
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
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
