"""Input normalisation with the reference's shape checks and errors (numba_celltree/cast.py:14-54).

Host (numpy, list) inputs are cast with numpy exactly as the reference does; device
tensors are checked for the same shapes and converted on the device.
"""

import numpy as np

from numba_celltree_amd.constants import FILL_VALUE, MAX_N_VERTEX, FloatDType, IntDType


def _is_tensor(x) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def cast_vertices(vertices, copy: bool = False):  # cast.py:14-21
    if _is_tensor(vertices):
        import torch

        v = vertices.to(torch.float64)
        if copy and v.data_ptr() == vertices.data_ptr():
            v = v.clone()
        if v.ndim != 2 or v.shape[1] != 2:
            raise ValueError("vertices must have shape (n_points, 2)")
        return v.contiguous()
    if isinstance(vertices, np.ndarray):
        vertices = vertices.astype(FloatDType, copy=copy)
    else:
        vertices = np.ascontiguousarray(vertices, dtype=FloatDType)
    if vertices.ndim != 2 or vertices.shape[1] != 2:
        raise ValueError("vertices must have shape (n_points, 2)")
    return np.ascontiguousarray(vertices)


def _check_faces_shape(ndim: int, n_max_vert: int) -> None:
    if ndim != 2:
        raise ValueError("faces must have shape (n_face, n_max_vert)")
    if n_max_vert > MAX_N_VERTEX:
        raise ValueError(
            f"faces contains up to {n_max_vert} vertices for a single face. "
            f"numba_celltree supports a maximum of {MAX_N_VERTEX} vertices. "
            f"Increase MAX_N_VERTEX in the source code, or alter the mesh."
        )


def cast_faces(faces, fill_value: int):  # cast.py:24-40
    if _is_tensor(faces):
        import torch

        f = faces.to(torch.int64).clone()
        _check_faces_shape(f.ndim, f.shape[1] if f.ndim == 2 else 0)
        if fill_value != FILL_VALUE:
            f[f == fill_value] = FILL_VALUE
        return f.contiguous()
    if isinstance(faces, np.ndarray):
        faces = faces.astype(IntDType, copy=True)
    else:
        faces = np.ascontiguousarray(faces, dtype=IntDType)
    _check_faces_shape(faces.ndim, faces.shape[1] if faces.ndim == 2 else 0)
    if fill_value != FILL_VALUE:
        faces[faces == fill_value] = FILL_VALUE
    return np.ascontiguousarray(faces)


def cast_bboxes(bbox_coords):  # cast.py:43-47
    if _is_tensor(bbox_coords):
        import torch

        b = bbox_coords.to(torch.float64).contiguous()
    else:
        b = np.ascontiguousarray(bbox_coords, dtype=FloatDType)
    if b.ndim != 2 or b.shape[1] != 4:
        raise ValueError("bbox_coords must have shape (n_box, 4)")
    return b


def cast_edges(edges):  # cast.py:50-54
    if _is_tensor(edges):
        import torch

        e = edges.to(torch.float64).contiguous()
    else:
        e = np.ascontiguousarray(edges, dtype=FloatDType)
    if e.ndim != 3 or e.shape[1] != 2 or e.shape[2] != 2:
        raise ValueError("edges must have shape (n_edge, 2, 2)")
    return e
