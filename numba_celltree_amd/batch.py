"""Many cell trees, one persistent launch (no reference counterpart).

The reference answers one mesh per call; a workload over many meshes becomes a loop of
``locate_points`` / ``intersect_edges`` calls (celltree.py:216-239, :294-303), each a few
small launches plus host work.  ``TreeBatch`` holds the device descriptors of many
``CellTree2d`` trees; its queries (one set per tree, concatenated) run in one persistent
grid over a tile queue (ct_batch_* in the C ABI).  Results equal the per-tree calls:
``locate_points`` gives each query its own tree's cell index; ``intersect_edges`` gives
rows (ii = concatenated segment index, jj = cell of that segment's tree, xy) in the
reference's per-segment order, tree after tree, with ``row_offsets`` marking each tree's
rows.

``BatchGraph`` captures a whole batch step (points, then segments: count, scan, write,
sort) in a HIP graph through ``torch.cuda.CUDAGraph``; a replay is one host call.  The
launch shape (workgroup size, workgroups per CU, tile size, stack class, queue count) is a
``BatchConfig``; ``autotune`` times candidates on the actual batch and keeps the fastest.
"""

import copy
import ctypes
import dataclasses
import heapq
import itertools
import json
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native
from numba_celltree_amd.celltree_base import CellTree2dBase

STACK_CLASSES = (16, 24, 32, 48, 64, 128, 256)
TILE_DTYPE = np.dtype([("begin", "<i8"), ("count", "<i4"), ("tree", "<i4")])


@dataclasses.dataclass(frozen=True)
class BatchConfig:
    """Launch shape of a batched persistent kernel."""

    block: int = 256        # workgroup size (64, 128, 256)
    wg_per_cu: int = 4      # persistent workgroups per compute unit
    tile: int = 256         # queries per tile (a multiple of 64, taken by one wavefront)
    stack: int = 0          # traversal stack class; 0 = the smallest class >= max depth
    n_queues: int = 8       # work queues (8 = one per XCD)

    def key(self) -> Tuple:
        return dataclasses.astuple(self)


def stack_class(depth: int) -> int:
    for s in STACK_CLASSES:
        if s >= depth:
            return s
    raise ValueError(f"tree depth {depth} exceeds {STACK_CLASSES[-1]}")


def _cu_count() -> int:
    return int(torch.cuda.get_device_properties(D.device()).multi_processor_count)


class TreeBatch:
    """Device descriptors of many face trees (``CellTree2d``, 3 to 8 vertices per face)."""

    def __init__(self, trees: Sequence, tolerances: Optional[Sequence[float]] = None):
        trees = list(trees)
        if not trees:
            raise ValueError("a batch needs at least one tree")
        for t in trees:
            nm = int(t.d_elements.shape[1])
            if not hasattr(t, "d_node_boxes") or not 3 <= nm <= 8:
                raise ValueError("batched queries take CellTree2d trees with 3..8 vertices per face")
            if int(t.depth) > _native.CT_MAX_DEPTH:
                raise ValueError(f"batched queries take trees of depth <= {_native.CT_MAX_DEPTH} (this one: "
                                 f"{int(t.depth)}); query deeper trees one by one")
        self.trees = trees
        self.n_trees = len(trees)
        raw = b"".join(ctypes.string_at(ctypes.addressof(t._ctree), ctypes.sizeof(t._ctree)) for t in trees)
        self.d_trees = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(D.device())
        if tolerances is None:
            tolerances = [t._default_tolerance() for t in trees]
        if len(tolerances) != self.n_trees:
            raise ValueError("one tolerance per tree")
        self.d_tols = torch.tensor([float(x) for x in tolerances], dtype=torch.float64, device=D.device())
        self.max_n_vert = max(int(t.d_elements.shape[1]) for t in trees)
        self.max_depth = max(int(t.depth) for t in trees)
        self.depths = np.array([int(t.depth) for t in trees])
        self.n_cells = np.array([int(t.d_elements.shape[0]) for t in trees])
        self.n_cu = _cu_count()

    # --------------------------------------------------------------- planning
    def plan(self, counts: Sequence[int], config: BatchConfig, weight: float = 1.0) -> "BatchPlan":
        return BatchPlan(self, np.asarray(counts, dtype=np.int64), config, weight)


class BatchPlan:
    """Tiles and queues of one query set (``counts[m]`` queries for tree m) under a config.

    Trees are assigned whole to queues, largest estimated work first, each to the least
    loaded queue (work ~ queries * (depth + 8)); a queue lists its trees' tiles in order.
    """

    def __init__(self, batch: TreeBatch, counts: np.ndarray, config: BatchConfig, weight: float):
        if counts.shape != (batch.n_trees,) or (counts < 0).any():
            raise ValueError("counts: one non-negative count per tree")
        if config.tile % 64 or config.tile <= 0:
            raise ValueError("tile must be a positive multiple of 64")
        stack = config.stack or stack_class(batch.max_depth)
        if stack not in STACK_CLASSES or stack < batch.max_depth:
            raise ValueError(f"stack {stack} cannot hold depth {batch.max_depth}")
        nq = int(config.n_queues)
        self.batch = batch
        self.config = config
        self.counts = counts
        self.n_queries = int(counts.sum())
        qoff = np.zeros(batch.n_trees + 1, dtype=np.int64)
        np.cumsum(counts, out=qoff[1:])
        self.query_offsets = qoff
        work = counts * (batch.depths + 8.0) * weight
        loads = [(0.0, r) for r in range(nq)]
        heapq.heapify(loads)
        members: List[List[int]] = [[] for _ in range(nq)]
        for m in np.argsort(-work, kind="stable"):
            if counts[m] == 0:
                continue
            load, r = heapq.heappop(loads)
            members[r].append(int(m))
            heapq.heappush(loads, (load + float(work[m]), r))
        T = int(config.tile)
        parts = []
        qstart = [0]
        for r in range(nq):
            in_queue = 0
            for m in members[r]:
                k = (int(counts[m]) + T - 1) // T
                tl = np.empty(k, dtype=TILE_DTYPE)
                tl["begin"] = np.arange(k, dtype=np.int64) * T
                tl["count"] = np.minimum(T, counts[m] - tl["begin"]).astype(np.int32)
                tl["tree"] = m
                parts.append(tl)
                in_queue += k
            qstart.append(qstart[-1] + in_queue)
        tiles = np.concatenate(parts) if parts else np.empty(0, dtype=TILE_DTYPE)
        if len(tiles) >= 2**31:
            raise ValueError("too many tiles")
        self.n_tiles = len(tiles)
        dev = D.device()
        self.d_tiles = torch.from_numpy(tiles.view(np.uint8).copy()).to(dev) if len(tiles) else \
            torch.zeros(16, dtype=torch.uint8, device=dev)
        self.d_qoff = torch.from_numpy(qoff).to(dev)
        self.d_qstart = torch.tensor(qstart, dtype=torch.int32, device=dev)
        self.d_counters = torch.zeros(nq, dtype=torch.int32, device=dev)
        s = _native.CtBatch()
        s.trees = batch.d_trees.data_ptr()
        s.tolerances = batch.d_tols.data_ptr()
        s.query_offsets = self.d_qoff.data_ptr()
        s.tiles = self.d_tiles.data_ptr()
        s.queue_starts = self.d_qstart.data_ptr()
        s.counters = self.d_counters.data_ptr()
        s.n_tiles = self.n_tiles
        s.n_trees = batch.n_trees
        s.n_queues = nq
        s.max_n_vert = batch.max_n_vert
        s.max_depth = batch.max_depth
        s.stack = stack
        s.n_workgroups = max(1, min(65536, batch.n_cu * int(config.wg_per_cu)))
        s.block = int(config.block)
        self._struct = s

    @property
    def ref(self):
        return ctypes.byref(self._struct)

    def relaunch(self, config: BatchConfig) -> "BatchPlan":
        """The same tiles and queues under another workgroup size, grid or stack class
        (``config.tile`` and ``config.n_queues`` must match this plan's)."""
        if (config.tile, config.n_queues) != (self.config.tile, self.config.n_queues):
            raise ValueError("relaunch keeps the tiles: tile and n_queues must match")
        stack = config.stack or stack_class(self.batch.max_depth)
        if stack not in STACK_CLASSES or stack < self.batch.max_depth:
            raise ValueError(f"stack {stack} cannot hold depth {self.batch.max_depth}")
        other = copy.copy(self)
        other.config = config
        other._struct = _native.CtBatch.from_buffer_copy(self._struct)
        other._struct.stack = stack
        other._struct.n_workgroups = max(1, min(65536, self.batch.n_cu * int(config.wg_per_cu)))
        other._struct.block = int(config.block)
        return other

    # ------------------------------------------------------------ kernels
    def locate_points(self, points: torch.Tensor, out: torch.Tensor) -> None:
        """Enqueue ct_batch_locate_points (points: (n_queries, 2) f64, out: (n_queries,) i64)."""
        _check(points, (self.n_queries, 2), torch.float64)
        _check(out, (self.n_queries,), torch.int64)
        _native.call("ct_batch_locate_points", self.ref, D.ptr(points), D.ptr(out), D.stream())

    def edges_count(self, segments: torch.Tensor, counts: torch.Tensor, hits: Optional[torch.Tensor] = None,
                    slots: int = 0) -> None:
        """First pass; with `hits` ((slots * n_queries,) int32) it also records hit cells."""
        _check(segments, (self.n_queries, 2, 2), torch.float64)
        _check(counts, (self.n_queries,), torch.int64)
        self._check_hits(hits, slots)
        _native.call("ct_batch_edges_count_rec", self.ref, D.ptr(segments), self.n_queries, D.ptr(counts),
                     D.ptr(hits), int(slots) if hits is not None else 0, D.stream())

    def edges_write(self, segments, offsets, capacity: int, ii, jj, xy, hits: Optional[torch.Tensor] = None,
                    slots: int = 0) -> None:
        _check(offsets, (self.n_queries + 1,), torch.int64)
        for t in (ii, jj):
            _check(t, (capacity,), torch.int64)
        _check(xy, (capacity, 2, 2), torch.float64)
        self._check_hits(hits, slots)
        _native.call("ct_batch_edges_write_rec", self.ref, D.ptr(segments), self.n_queries, D.ptr(offsets),
                     D.ptr(hits), int(slots) if hits is not None else 0, int(capacity), D.ptr(ii), D.ptr(jj),
                     D.ptr(xy), D.stream())

    def _check_hits(self, hits, slots: int) -> None:
        if hits is not None:
            if not 1 <= int(slots) <= 1024:
                raise ValueError("slots must be in [1, 1024]")
            _check(hits, (int(slots) * self.n_queries,), torch.int32)


def _check(t: torch.Tensor, shape, dtype) -> None:
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and tuple(t.shape) == tuple(shape)
            and t.is_contiguous() and t.data_ptr() % 32 == 0):
        raise ValueError(f"expected a contiguous 32-byte aligned {dtype} device tensor of shape {shape}")


def concat_queries(sets: Sequence, width: int) -> Tuple[torch.Tensor, np.ndarray]:
    """Concatenate per-tree query arrays (numpy or tensors) into one aligned device tensor."""
    parts = [D.to_dev(s, torch.float64).reshape(-1, width) for s in sets]
    counts = np.array([int(p.shape[0]) for p in parts], dtype=np.int64)
    cat = torch.cat(parts) if parts else D.empty((0, width), torch.float64)
    if cat.data_ptr() % 32:
        cat = cat.clone()
    return cat, counts


# Hit record of the batched segment passes (as CellTree2d.EDGE_HIT_SLOTS / EDGE_HIT_BUDGET):
# up to EDGE_HIT_SLOTS recorded cells per segment within EDGE_HIT_BUDGET int32 entries.  A
# batch is one large call, so the budget is 8 GiB (2.8% of a 288 GB device; slot-major, so
# only slots in use are written): 256 slots for 7.5M segments.  With 35 slots (1 GiB) the
# long segments, which carry most rows, still walked twice and the record did not pay.
EDGE_HIT_SLOTS = 256
EDGE_HIT_BUDGET = 1 << 31


class EdgeBuffers:
    """Device buffers of a batched segment pass with room for `capacity` rows, and the
    count pass's hit record (None below 4 slots per segment)."""

    def __init__(self, n_queries: int, capacity: int, record: bool = True):
        self.hits, self.slots = None, 0
        if record and n_queries:
            # Fewer slots (down to none) when the device cannot hold the full record.
            self.hits, self.slots = CellTree2dBase._hit_record(n_queries, EDGE_HIT_SLOTS, EDGE_HIT_BUDGET)
        self.capacity = int(capacity)
        self.counts = D.empty(n_queries, torch.int64)
        self.offsets = D.empty(n_queries + 1, torch.int64)
        self.scan_ws = D.empty(int(_native.lib().ct_scan_workspace_bytes(n_queries)), torch.uint8)
        self.ii = D.empty(self.capacity, torch.int64)
        self.jj = D.empty(self.capacity, torch.int64)
        self.xy = D.empty((self.capacity, 2, 2), torch.float64)
        self.jj_sorted = D.empty(self.capacity, torch.int64)
        self.xy_sorted = D.empty((self.capacity, 2, 2), torch.float64)


def run_edges(plan: BatchPlan, segments: torch.Tensor, buf: EdgeBuffers, sort_workgroups: int) -> None:
    """count -> scan -> write -> per-segment sort, all enqueued (no host synchronization)."""
    n = plan.n_queries
    st = D.stream()
    plan.edges_count(segments, buf.counts, buf.hits, buf.slots)
    _native.call("ct_exclusive_scan", D.ptr(buf.counts), n, D.ptr(buf.offsets), D.ptr(buf.scan_ws),
                 buf.scan_ws.numel(), st)
    plan.edges_write(segments, buf.offsets, buf.capacity, buf.ii, buf.jj, buf.xy, buf.hits, buf.slots)
    _native.call("ct_batch_sort_intersections", D.ptr(segments), n, D.ptr(buf.offsets), D.ptr(buf.ii),
                 D.ptr(buf.jj), D.ptr(buf.xy), buf.capacity, int(sort_workgroups), D.ptr(buf.jj_sorted),
                 D.ptr(buf.xy_sorted), st)


class BatchResult:
    """Concatenated batch results plus per-tree views."""

    def __init__(self, plan_points, plan_edges, cells, ii, jj, xy, offsets):
        self.plan_points = plan_points
        self.plan_edges = plan_edges
        self.cells = cells      # (n_points,) per-tree cell index or -1
        self.ii = ii            # (n_rows,) concatenated segment index
        self.jj = jj            # (n_rows,) cell of the segment's tree
        self.xy = xy            # (n_rows, 2, 2)
        self.offsets = offsets  # (n_segments + 1,) row offsets

    def points_of(self, m: int) -> torch.Tensor:
        q = self.plan_points.query_offsets
        return self.cells[q[m]: q[m + 1]]

    def row_offsets(self) -> np.ndarray:
        q = torch.from_numpy(self.plan_edges.query_offsets).to(self.offsets.device)
        return self.offsets[q].cpu().numpy()

    def edges_of(self, m: int, row_offsets: Optional[np.ndarray] = None):
        """(ii, jj, xy) of tree m with ii local to tree m's segments, as CellTree2d returns."""
        ro = self.row_offsets() if row_offsets is None else row_offsets
        a, b = int(ro[m]), int(ro[m + 1])
        return self.ii[a:b] - int(self.plan_edges.query_offsets[m]), self.jj[a:b], self.xy[a:b]


class BatchGraph:
    """A batch step (point location and/or segment intersection of every tree) captured in
    a HIP graph.  The segment rows are bounded by `capacity` (by default the exact count of
    a first eager run); ``replay()`` enqueues the graph, ``overflowed()`` (synchronizes)
    reports whether a replay produced more rows than the capacity -- then build a new graph."""

    def __init__(self, batch: TreeBatch, points: Optional[torch.Tensor] = None,
                 point_counts: Optional[Sequence[int]] = None, segments: Optional[torch.Tensor] = None,
                 segment_counts: Optional[Sequence[int]] = None, points_config: BatchConfig = BatchConfig(),
                 edges_config: BatchConfig = BatchConfig(), capacity: Optional[int] = None):
        self.batch = batch
        self.points = points
        self.segments = segments
        self.pplan = batch.plan(point_counts, points_config) if points is not None else None
        self.eplan = batch.plan(segment_counts, edges_config) if segments is not None else None
        self.cells = D.empty(self.pplan.n_queries, torch.int64) if self.pplan else None
        self.sort_wg = 4 * batch.n_cu
        self.buf = None
        if self.eplan is not None:
            if capacity is None:
                probe = EdgeBuffers(self.eplan.n_queries, 0, record=False)
                self.eplan.edges_count(segments, probe.counts)
                _native.call("ct_exclusive_scan", D.ptr(probe.counts), self.eplan.n_queries, D.ptr(probe.offsets),
                             D.ptr(probe.scan_ws), probe.scan_ws.numel(), D.stream())
                capacity = int(probe.offsets[-1].item())
                del probe
            self.buf = EdgeBuffers(self.eplan.n_queries, capacity)
        # warm up once eagerly on the capture stream, then capture
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._step()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._step()
        torch.cuda.synchronize()

    def _step(self) -> None:
        if self.pplan is not None:
            self.pplan.locate_points(self.points, self.cells)
        if self.eplan is not None:
            run_edges(self.eplan, self.segments, self.buf, self.sort_wg)

    def replay(self) -> None:
        self.graph.replay()

    def n_rows(self) -> int:
        return int(self.buf.offsets[-1].item()) if self.buf is not None else 0

    def overflowed(self) -> bool:
        return self.buf is not None and self.n_rows() > self.buf.capacity

    def result(self) -> BatchResult:
        if self.overflowed():
            raise RuntimeError("segment rows exceeded the graph's capacity; build a new BatchGraph")
        n = self.n_rows()
        b = self.buf
        return BatchResult(
            self.pplan, self.eplan, self.cells,
            b.ii[:n] if b else None, b.jj_sorted[:n] if b else None, b.xy_sorted[:n] if b else None,
            b.offsets if b else None,
        )


# -------------------------------------------------------------------- eager API
def locate_points(batch: TreeBatch, point_sets: Sequence, config: BatchConfig = BatchConfig()):
    """Per-tree ``locate_points`` of every query set in one launch; returns the list of
    per-tree results (device tensors)."""
    pts, counts = concat_queries(point_sets, 2)
    plan = batch.plan(counts, config)
    out = D.empty(plan.n_queries, torch.int64)
    plan.locate_points(pts, out)
    q = plan.query_offsets
    return [out[q[m]: q[m + 1]] for m in range(batch.n_trees)]


def intersect_edges(batch: TreeBatch, segment_sets: Sequence, config: BatchConfig = BatchConfig()):
    """Per-tree ``intersect_edges`` of every segment set in one launch sequence; returns the
    list of per-tree (ii, jj, xy) (device tensors; ii local to the tree's segments)."""
    segs, counts = concat_queries(segment_sets, 4)
    segs = segs.reshape(-1, 2, 2)
    plan = batch.plan(counts, config)
    n = plan.n_queries
    probe = EdgeBuffers(n, 0, record=False)
    plan.edges_count(segs, probe.counts)
    _native.call("ct_exclusive_scan", D.ptr(probe.counts), n, D.ptr(probe.offsets), D.ptr(probe.scan_ws),
                 probe.scan_ws.numel(), D.stream())
    total = int(probe.offsets[-1].item())
    buf = EdgeBuffers(n, total)
    run_edges(plan, segs, buf, 4 * batch.n_cu)
    res = BatchResult(None, plan, None, buf.ii, buf.jj_sorted, buf.xy_sorted, buf.offsets)
    ro = res.row_offsets()
    return [res.edges_of(m, ro) for m in range(batch.n_trees)]


# -------------------------------------------------------------------- autotuner
def default_candidates(batch: TreeBatch) -> List[BatchConfig]:
    base = stack_class(batch.max_depth)
    stacks = [base] + [s for s in STACK_CLASSES if s > base][:1]
    out = []
    for block, wg, tile, stack in itertools.product((64, 256), (2, 4, 8), (64, 256, 1024), stacks):
        out.append(BatchConfig(block=block, wg_per_cu=wg, tile=tile, stack=stack))
    return out


_TUNED: Dict[Tuple, BatchConfig] = {}


def autotune(batch: TreeBatch, kind: str, queries: torch.Tensor, counts: Sequence[int],
             candidates: Optional[Sequence[BatchConfig]] = None, reps: int = 3,
             cache_path: Optional[str] = None) -> Tuple[BatchConfig, List[Tuple[BatchConfig, float]]]:
    """Time every candidate config on this batch (kind "points" or "edges": the count pass,
    the dominant kernel) and return the fastest, with all (config, ms) timings.  Results are
    kept per (kind, tree count, query count, depth, vertex class), in memory and, with
    `cache_path`, in a JSON file."""
    if kind not in ("points", "edges"):
        raise ValueError("kind must be 'points' or 'edges'")
    key = (kind, batch.n_trees, int(np.sum(counts)), batch.max_depth, batch.max_n_vert)
    skey = json.dumps(key)
    if key in _TUNED:
        return _TUNED[key], []
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            saved = json.load(fh)
        if skey in saved:
            cfg = BatchConfig(**saved[skey])
            _TUNED[key] = cfg
            return cfg, []
    candidates = list(candidates or default_candidates(batch))
    n = int(np.sum(counts))
    out = D.empty(n, torch.int64)
    timings = []
    tiled: Dict[Tuple[int, int], BatchPlan] = {}  # plans differ only in tiles and queues
    for cfg in candidates:
        tk = (cfg.tile, cfg.n_queues)
        if tk not in tiled:
            tiled[tk] = batch.plan(counts, cfg)
        plan = tiled[tk].relaunch(cfg)

        def step():
            if kind == "points":
                plan.locate_points(queries, out)
            else:
                plan.edges_count(queries, out)

        step()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            step()
        ev1.record()
        ev1.synchronize()
        timings.append((cfg, ev0.elapsed_time(ev1) / reps))
    best = min(timings, key=lambda x: x[1])[0]
    _TUNED[key] = best
    if cache_path:
        saved = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                saved = json.load(fh)
        saved[skey] = dataclasses.asdict(best)
        with open(cache_path, "w") as fh:
            json.dump(saved, fh, indent=1)
    return best, timings
