"""Device plumbing: torch tensors hold every device array, the HIP library computes.

Inputs may be host arrays (numpy / lists) or device tensors.  ``Io`` remembers which kind
came in so results go back the same way: numpy in, numpy out (the reference's contract);
device tensor in, device tensor out (no PCIe round trip).
"""

import ctypes
import warnings

import numpy as np
import torch

from numba_celltree_amd import _native
from numba_celltree_amd.constants import NodeDType


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "numba_celltree_amd computes on a ROCm GPU only, and none is visible "
            "(torch.cuda.is_available() is False)"
        )
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def is_device_tensor(x) -> bool:
    return isinstance(x, torch.Tensor) and x.is_cuda


class Io:
    """Tracks whether a call received device tensors, and converts results back."""

    def __init__(self, *inputs):
        self.device_io = any(is_device_tensor(x) for x in inputs)

    def out(self, t: torch.Tensor):
        if self.device_io:
            return t
        return t.cpu().numpy()


def host_tensor(a: np.ndarray) -> torch.Tensor:
    """A CPU tensor sharing ``a``'s memory; read-only arrays (views of immutable bytes) are
    shared too, for copies that only read them."""
    a = np.ascontiguousarray(a)
    if a.flags.writeable:
        return torch.from_numpy(a)
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", message="The given NumPy array is not writable")
        return torch.from_numpy(a)


def to_dev(x, dtype: torch.dtype) -> torch.Tensor:
    """Contiguous, 32-byte aligned device tensor of ``dtype`` (copies host data).

    The kernels read points, boxes and segments with 16- and 32-byte vector loads; a view
    that starts mid-row is copied to a fresh (aligned) allocation.
    """
    dev = device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=dtype).contiguous()
    else:
        t = host_tensor(np.asarray(x)).to(device=dev, dtype=dtype).contiguous()
    if t.data_ptr() % 32:
        t = t.clone()
    return t


def faces_to_dev(faces, n_vertices: int, allow_fill: bool = True) -> torch.Tensor:
    """int32 device copy of a (n, n_max_vert) index array.

    Every index must address one of ``n_vertices`` vertices; -1 (the fill value) is allowed
    only where ``allow_fill`` (face rows, past their first 3 entries; edge rows have no
    fill).  The kernels read vertices without bounds checks, so this is checked here, before
    any launch."""
    if isinstance(faces, torch.Tensor):
        f = faces
    else:
        f = host_tensor(np.asarray(faces))
    if f.numel():
        # one host round trip for all three checks (a device tensor costs a synchronization
        # per value read back)
        lo, hi = torch.aminmax(f)
        first3 = allow_fill and f.ndim == 2 and f.shape[1] >= 3
        lo3 = f[:, :3].amin() if first3 else lo
        hi, lo, lo3 = (int(v) for v in torch.stack([hi, lo, lo3]).tolist())
        if hi >= min(int(n_vertices), 2**31):
            raise ValueError(f"element index {hi} is out of bounds for {int(n_vertices)} vertices")
        if lo < (-1 if allow_fill else 0):
            what = "[-1, n_vertices)" if allow_fill else "[0, n_vertices)"
            raise ValueError(f"element indices must lie in {what}, found {lo}")
        if first3 and lo3 < 0:
            # polygon_length (geometry_utils.py:72-79) takes the first 3 entries as vertices
            raise ValueError("every face needs at least 3 vertices (a fill value among its first 3 entries)")
    return f.to(device=device(), dtype=torch.int32).contiguous()


def index_to_dev(x, n: int, allow_missing: bool = False) -> torch.Tensor:
    """int64 device copy of an index array into n items (-1 allowed where ``allow_missing``)."""
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
    if t.numel():
        lo, hi = (int(v) for v in torch.stack(torch.aminmax(t)).tolist())  # one host round trip
        if hi >= int(n) or lo < (-1 if allow_missing else 0):
            raise ValueError(f"indices must lie in [{-1 if allow_missing else 0}, {int(n)}), found [{lo}, {hi}]")
    return to_dev(t, torch.int64)


def empty(shape, dtype) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device=device())


def exclusive_scan(counts: torch.Tensor, total: bool = True):
    """Offsets (n + 1, int64, device) of ``counts`` and the total (host int: synchronizes;
    None with ``total=False``)."""
    n = counts.numel()
    offsets = empty(n + 1, torch.int64)
    ws = empty(int(_native.lib().ct_scan_workspace_bytes(n)), torch.uint8)
    _native.call("ct_exclusive_scan", ptr(counts), n, ptr(offsets), ptr(ws), ws.numel(), stream())
    return offsets, (int(offsets[n].item()) if total else None)


def query_order(kind: int, queries: torch.Tensor, bbox, with_sorted: bool = False):
    """Processing order (int32, device) of n queries: kind 0 points, 1 boxes, 2 segments.
    With ``with_sorted``, returns (order, sorted) where sorted[i] = queries[order[i]]."""
    n = int(queries.shape[0])
    order = empty(n, torch.int32)
    sorted_q = empty((n, 2 if kind == 0 else 4), torch.float64) if with_sorted else None
    ws = empty(int(_native.lib().ct_order_workspace_bytes(n)), torch.uint8)
    bb = (ctypes.c_double * 4)(*[float(x) for x in bbox])
    _native.call("ct_query_order", kind, ptr(queries), n, bb, ptr(ws), ws.numel(), ptr(order), ptr(sorted_q),
                 stream())
    return (order, sorted_q) if with_sorted else order


# ------------------------------------------------------------------ node records
def nodes_to_host(raw: torch.Tensor, n_nodes: int) -> np.ndarray:
    """Device ``ct_node`` bytes -> the reference's packed NodeDType array."""
    buf = raw[: 32 * n_nodes].cpu().numpy()
    f = buf.view(np.float64).reshape(n_nodes, 4)
    i = buf.view(np.int32).reshape(n_nodes, 8)
    out = np.empty(n_nodes, dtype=NodeDType)
    out["child"] = i[:, 4]
    out["Lmax"] = f[:, 0]
    out["Rmin"] = f[:, 1]
    out["ptr"] = i[:, 5]
    out["size"] = i[:, 6]
    out["dim"] = i[:, 7] != 0
    return out


def nodes_to_device(nodes: np.ndarray) -> torch.Tensor:
    """NodeDType array -> device ``ct_node`` bytes."""
    n = len(nodes)
    buf = np.zeros(32 * n, dtype=np.uint8)
    f = buf.view(np.float64).reshape(n, 4)
    i = buf.view(np.int32).reshape(n, 8)
    f[:, 0] = nodes["Lmax"]
    f[:, 1] = nodes["Rmin"]
    i[:, 4] = nodes["child"]
    i[:, 5] = nodes["ptr"]
    i[:, 6] = nodes["size"]
    i[:, 7] = nodes["dim"].astype(np.int32)
    return torch.from_numpy(buf).to(device())
