"""Conservative area-weighted regridding with a hand-written adjoint (no reference
counterpart; the forward is the caller pattern of examples/spatial_indexing_2d_grids.py:268-323).

    y[t, :] = sum_{pairs (t, s)} A_ts(src_vertices, tgt_vertices) * F[s, :] / area_t(tgt_vertices)

A_ts is the Sutherland-Hodgman overlap area of target polygon t (subject) and source polygon
s (clipper), exactly as ``CellTree2d.intersect_faces`` computes it; area_t is the target's
polygon area.  The pair topology -- which pairs overlap, and the path each clip takes -- is
computed once and held fixed; gradients are exact derivatives of that fixed computation with
respect to F, the source vertices and the target vertices (libcelltree_hip.so:
ct_regrid_pair_areas / ct_regrid_apply / ct_regrid_apply_transpose /
ct_regrid_vertex_backward).
"""

import ctypes
import dataclasses
import weakref
from typing import Optional, Tuple

import numpy as np
import torch

from numba_celltree_amd import _device as D
from numba_celltree_amd import _native
from numba_celltree_amd.cast import cast_faces, cast_vertices


def _offsets(index: torch.Tensor, n: int) -> torch.Tensor:
    counts = torch.bincount(index, minlength=n).to(torch.int64) if index.numel() else torch.zeros(
        n, dtype=torch.int64, device=index.device)
    offsets, _ = D.exclusive_scan(counts)
    return offsets


def _vertex_corners(faces: torch.Tensor, n_vertices: int):
    """(offsets (n_vertices + 1), corners): the corners face * n_max + slot of every vertex,
    ascending (a stable sort of the corners by vertex)."""
    flat = faces.reshape(-1).to(torch.int64)
    ids = torch.arange(flat.numel(), dtype=torch.int64, device=flat.device)
    valid = flat >= 0
    v = flat[valid]
    order = torch.argsort(v, stable=True)
    return _offsets(v, n_vertices), ids[valid][order].contiguous()


class RegridTopology:
    """The fixed pair structure between a source CellTree2d and a target mesh."""

    def __init__(self, tree, tgt_vertices, tgt_faces, fill_value: int = -1):
        self.tree = tree
        d_tv = D.to_dev(cast_vertices(tgt_vertices), torch.float64)
        d_tf = D.faces_to_dev(cast_faces(tgt_faces, fill_value), d_tv.shape[0])
        if int(d_tf.shape[1]) > 8 or int(tree.d_elements.shape[1]) > 8:
            raise ValueError("regridding supports faces of at most 8 vertices")
        # intersect_faces reorders the target rows counter-clockwise in place (d_tf) and
        # returns the pairs grouped by target in ascending order.
        tt, ss, _ = tree._face_pairs(d_tv, d_tf, with_area=True)
        self.tgt_vertices = d_tv
        self.tgt_faces = d_tf
        self.src_faces = tree.d_elements
        self.n_tgt = int(d_tf.shape[0])
        self.n_src = int(tree.d_elements.shape[0])
        self.tt = tt
        self.ss = ss
        self.n_pairs = int(tt.numel())
        self.tgt_offsets = _offsets(tt, self.n_tgt)
        self.perm_by_src = torch.argsort(ss, stable=True)
        self.src_tgt = tt[self.perm_by_src].contiguous()
        self.src_offsets = _offsets(ss, self.n_src)
        self.n_tgt_vertices = int(d_tv.shape[0])
        self.n_src_vertices = int(tree.d_vertices.shape[0])
        self.tgt_vertex_offsets, self.tgt_vertex_corners = _vertex_corners(d_tf, self.n_tgt_vertices)
        self.src_vertex_offsets, self.src_vertex_corners = _vertex_corners(self.src_faces, self.n_src_vertices)
        s = _native.CtRegridTopology()
        for name in ("tgt_faces", "src_faces", "tt", "ss", "tgt_offsets", "src_offsets", "perm_by_src",
                     "tgt_vertex_offsets", "tgt_vertex_corners", "src_vertex_offsets", "src_vertex_corners"):
            setattr(s, name, getattr(self, name).data_ptr())
        s.n_pairs, s.n_tgt, s.n_src = self.n_pairs, self.n_tgt, self.n_src
        s.n_tgt_vertices, s.n_src_vertices = self.n_tgt_vertices, self.n_src_vertices
        s.n_max_t, s.n_max_s = int(d_tf.shape[1]), int(self.src_faces.shape[1])
        # the targets the vertex backward's pair pass leaves to its per-target pass
        rest = D.empty(max(self.n_tgt, 1), torch.int64)
        n_rest = ctypes.c_int64(0)
        _native.call("ct_regrid_rest_targets", ctypes.byref(s), D.ptr(rest), ctypes.byref(n_rest), D.stream())
        self.rest_targets = rest  # the first n_rest_targets entries
        s.rest_targets = rest.data_ptr()
        s.n_rest_targets = n_rest.value
        self._struct = s
        self._area_cache = None

    # ------------------------------------------------------------- area cache
    def areas(self, src_v: torch.Tensor, tgt_v: torch.Tensor, cache: bool = True) -> Tuple[torch.Tensor, torch.Tensor]:
        """(pair areas, target areas) for these vertex tensors.  With ``cache``, the last
        result is kept and reused while the same two tensor objects are passed unchanged
        (torch's in-place version counters unchanged), so repeated forwards with fixed
        vertices only apply the operator.  Writes that bypass the version counter (through
        ``tensor.data``) are not seen: ``clear_cache`` drops the result.  The regrid forward
        caches only when no vertex gradient is requested (gradcheck, for one, perturbs its
        inputs through ``.data``)."""
        c = self._area_cache
        key = (src_v._version, tgt_v._version)
        if cache and c is not None and c[0]() is src_v and c[1]() is tgt_v and c[2] == key:
            return c[3], c[4]
        area = self.pair_areas(src_v, tgt_v)
        tarea = self.target_areas(tgt_v)
        if cache:
            self._area_cache = (weakref.ref(src_v), weakref.ref(tgt_v), key, area, tarea)
        return area, tarea

    def clear_cache(self) -> None:
        self._area_cache = None

    def overlap_sums(self, area: torch.Tensor) -> torch.Tensor:
        """S_t = sum of the overlap areas of target t's pairs (0 for a target without pairs),
        summed in pair order (deterministic)."""
        ones_s = torch.ones((self.n_src, 1), dtype=torch.float64, device=area.device)
        ones_t = torch.ones(self.n_tgt, dtype=torch.float64, device=area.device)
        return self.apply(area, ones_t, ones_s)[:, 0].contiguous()

    # ------------------------------------------------------------------ kernels
    def pair_areas(self, src_v: torch.Tensor, tgt_v: torch.Tensor) -> torch.Tensor:
        area = D.empty(self.n_pairs, torch.float64)
        _native.call(
            "ct_regrid_pair_areas", D.ptr(tgt_v), D.ptr(self.tgt_faces), int(self.tgt_faces.shape[1]),
            D.ptr(src_v), D.ptr(self.src_faces), int(self.src_faces.shape[1]), D.ptr(self.tt), D.ptr(self.ss),
            self.n_pairs, D.ptr(area), D.stream(),
        )
        return area

    def target_areas(self, tgt_v: torch.Tensor) -> torch.Tensor:
        out = D.empty(self.n_tgt, torch.float64)
        _native.call("ct_polygon_areas", D.ptr(tgt_v), D.ptr(self.tgt_faces), int(self.tgt_faces.shape[1]),
                     self.n_tgt, D.ptr(out), D.stream())
        return out

    def apply(self, area, tarea, F: torch.Tensor) -> torch.Tensor:
        nf = int(F.shape[1])
        y = D.empty((self.n_tgt, nf), torch.float64)
        _native.call("ct_regrid_apply", D.ptr(self.tgt_offsets), D.ptr(self.ss), D.ptr(area), D.ptr(tarea),
                     D.ptr(F), nf, self.n_tgt, D.ptr(y), D.stream())
        return y

    def apply_transpose(self, area, tarea, gy: torch.Tensor) -> torch.Tensor:
        nf = int(gy.shape[1])
        gF = D.empty((self.n_src, nf), torch.float64)
        _native.call("ct_regrid_apply_transpose", D.ptr(self.src_offsets), D.ptr(self.perm_by_src),
                     D.ptr(self.src_tgt), D.ptr(area), D.ptr(tarea), D.ptr(gy), nf, self.n_src, D.ptr(gF),
                     D.stream())
        return gF

    def vertex_backward(self, src_v, tgt_v, F, y, gy, tarea):
        if int(tgt_v.shape[0]) != self.n_tgt_vertices or int(src_v.shape[0]) != self.n_src_vertices:
            raise ValueError("vertex arrays must keep the sizes of the topology's meshes")
        g_tv = torch.empty_like(tgt_v)
        g_sv = torch.empty_like(src_v)
        ref = ctypes.byref(self._struct)
        ws = D.empty(int(_native.lib().ct_regrid_backward_workspace_bytes(ref)), torch.uint8)
        _native.call(
            "ct_regrid_vertex_backward", ref, D.ptr(tgt_v), D.ptr(src_v), D.ptr(F), D.ptr(y), D.ptr(gy),
            D.ptr(tarea), int(F.shape[1]), D.ptr(ws), ws.numel(), D.ptr(g_tv), D.ptr(g_sv), D.stream(),
        )
        return g_sv, g_tv


def _require(name: str, t, shape) -> None:
    """Raise unless ``t`` is a float64 tensor on the current GPU with ``shape`` (None: any)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64, got {t.dtype}")
    if not t.is_cuda or t.device != D.device():
        raise ValueError(f"{name} must be on {D.device()}, got {t.device}")
    if t.ndim != len(shape) or any(want is not None and int(got) != want for got, want in zip(t.shape, shape)):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")


NORMALIZATIONS = ("target", "overlap", "none")


class RegridFunction(torch.autograd.Function):
    """y = regrid(F, src_vertices, tgt_vertices) under a fixed RegridTopology.

    F (n_src, n_fields), src_vertices (n_src_vertices, 2) and tgt_vertices
    (n_tgt_vertices, 2) must be float64 tensors on the current GPU (``Regridder`` converts
    other inputs).  ``normalize``: "target" divides each target's sum by its polygon area
    (conservative regridding; differentiable in F and both vertex sets), "overlap" by the sum
    of its overlap areas (the area-weighted mean of examples/spatial_indexing_2d_grids.py:
    285-298; NaN for a target that overlaps no source; differentiable in F only), "none"
    leaves the plain sum of area-weighted values."""

    @staticmethod
    def forward(ctx, F, src_vertices, tgt_vertices, topo: RegridTopology, normalize: str = "target"):
        _require("F", F, (topo.n_src, None))
        _require("src_vertices", src_vertices, (topo.n_src_vertices, 2))
        _require("tgt_vertices", tgt_vertices, (topo.n_tgt_vertices, 2))
        if normalize not in NORMALIZATIONS:
            raise ValueError(f"normalize must be one of {NORMALIZATIONS}")
        vertex_grad = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        area, tarea = topo.areas(src_vertices, tgt_vertices, cache=not vertex_grad)
        if normalize == "overlap":
            div = topo.overlap_sums(area)
        elif normalize == "none":
            div = torch.ones_like(tarea)
        else:
            div = tarea
        F = F.contiguous()
        y = topo.apply(area, div, F)
        ctx.topo = topo
        ctx.normalize = normalize
        ctx.save_for_backward(F, src_vertices.contiguous(), tgt_vertices.contiguous(), area, div, y)
        return y

    @staticmethod
    def backward(ctx, gy):
        F, src_v, tgt_v, area, div, y = ctx.saved_tensors
        topo = ctx.topo
        gy = gy.to(dtype=torch.float64).contiguous()
        gF = topo.apply_transpose(area, div, gy) if ctx.needs_input_grad[0] else None
        g_sv = g_tv = None
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            if ctx.normalize != "target":
                raise NotImplementedError("vertex gradients are implemented for normalize='target'")
            g_sv, g_tv = topo.vertex_backward(src_v, tgt_v, F, y, gy, div)
        return gF, g_sv, g_tv, None, None


@dataclasses.dataclass
class RegridWeights:
    """The regridding operator as a sparse (n_tgt, n_src) matrix in CSR form, rows grouped by
    target in intersect_faces order: y = W @ F.  ``values`` are the overlap areas divided by
    the normaliser (see RegridFunction); ``coo()`` gives the (target, source, value) lists the
    reference's intersect_* methods return (examples/spatial_indexing_2d_grids.py:268-323)."""

    crow_indices: torch.Tensor  # (n_tgt + 1,) int64
    col_indices: torch.Tensor   # (n_pairs,) int64 source cells
    values: torch.Tensor        # (n_pairs,) float64
    shape: Tuple[int, int]
    _topo: "RegridTopology" = dataclasses.field(repr=False, default=None)

    def coo(self):
        rows = torch.repeat_interleave(
            torch.arange(self.shape[0], device=self.values.device), torch.diff(self.crow_indices))
        return rows, self.col_indices, self.values

    def to_torch(self) -> torch.Tensor:
        return torch.sparse_csr_tensor(self.crow_indices, self.col_indices, self.values, size=self.shape)

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.values.cpu().numpy(), self.col_indices.cpu().numpy(),
                              self.crow_indices.cpu().numpy()), shape=self.shape)

    def apply(self, F) -> torch.Tensor:
        """W @ F for F (n_src,) or (n_src, n_fields): a deterministic segmented sum per row."""
        F = _as_f64(F)
        squeeze = F.ndim == 1
        F2 = (F[:, None] if squeeze else F).contiguous()
        _require("F", F2, (self.shape[1], None))
        ones = torch.ones(self.shape[0], dtype=torch.float64, device=F2.device)
        y = self._topo.apply(self.values, ones, F2)
        return y[:, 0] if squeeze else y


def _as_f64(x) -> torch.Tensor:
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x)
    return x.to(device=D.device(), dtype=torch.float64)


class Regridder:
    """Conservative regridding from the cells of ``tree`` (source) to a target mesh.

    >>> rg = Regridder(tree, tgt_vertices, tgt_faces)
    >>> y = rg(F)                                  # F: (n_src, n_fields) device float64
    >>> y = rg(F, src_vertices=v_s, tgt_vertices=v_t)   # differentiable in all three
    """

    def __init__(self, tree, tgt_vertices, tgt_faces, fill_value: int = -1, normalize: str = "target"):
        if normalize not in NORMALIZATIONS:
            raise ValueError(f"normalize must be one of {NORMALIZATIONS}")
        self.topology = RegridTopology(tree, tgt_vertices, tgt_faces, fill_value)
        self.normalize = normalize

    def weights(self, src_vertices: Optional[torch.Tensor] = None,
                tgt_vertices: Optional[torch.Tensor] = None) -> RegridWeights:
        """The operator of this regridder (at the given vertices) as a CSR matrix."""
        topo = self.topology
        src_vertices = topo.tree.d_vertices if src_vertices is None else _as_f64(src_vertices)
        tgt_vertices = topo.tgt_vertices if tgt_vertices is None else _as_f64(tgt_vertices)
        _require("src_vertices", src_vertices, (topo.n_src_vertices, 2))
        _require("tgt_vertices", tgt_vertices, (topo.n_tgt_vertices, 2))
        with torch.no_grad():
            area, tarea = topo.areas(src_vertices, tgt_vertices)
            if self.normalize == "target":
                values = area / tarea[topo.tt]
            elif self.normalize == "overlap":
                values = area / topo.overlap_sums(area)[topo.tt]
            else:
                values = area.clone()
        return RegridWeights(topo.tgt_offsets, topo.ss, values, (topo.n_tgt, topo.n_src), topo)

    def __call__(self, F, src_vertices: Optional[torch.Tensor] = None,
                 tgt_vertices: Optional[torch.Tensor] = None) -> torch.Tensor:
        topo = self.topology
        if src_vertices is None:
            src_vertices = topo.tree.d_vertices
        if tgt_vertices is None:
            tgt_vertices = topo.tgt_vertices
        # Host arrays, other float dtypes and other devices are converted (differentiably);
        # shapes are checked by RegridFunction before any kernel runs.
        F, src_vertices, tgt_vertices = (_as_f64(x) for x in (F, src_vertices, tgt_vertices))
        squeeze = F.ndim == 1
        if squeeze:
            F = F[:, None]
        y = RegridFunction.apply(F, src_vertices, tgt_vertices, topo, self.normalize)
        return y[:, 0] if squeeze else y
