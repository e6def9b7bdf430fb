"""Types and constants of the cell tree (numba_celltree/constants.py:24-151).

The host-side record layouts are the reference's own, so that ``tree.nodes`` and
``tree.celltree_data`` are drop-in: ``NodeDType`` is the packed 41-byte record
(constants.py:92-106).  On the device a node is the 32-byte ``ct_node`` of
include/celltree_hip.h.
"""

from typing import NamedTuple

import numpy as np

IntDType = np.intp
FloatDType = np.float64
BoolArray = np.ndarray
IntArray = np.ndarray
FloatArray = np.ndarray

NDIM = 2
MAX_N_VERTEX = 32
FILL_VALUE = -1
MIN_TOLERANCE = 1e-15
TOLERANCE_FACTOR = 1e-12
FLOAT_MIN = np.finfo(FloatDType).min
FLOAT_MAX = np.finfo(FloatDType).max


class Point(NamedTuple):
    x: float
    y: float


class Vector(NamedTuple):
    x: float
    y: float


class Box(NamedTuple):
    xmin: float
    xmax: float
    ymin: float
    ymax: float


class CellTreeData(NamedTuple):
    elements: IntArray
    vertices: FloatArray
    nodes: np.ndarray
    bb_indices: IntArray
    bb_coords: FloatArray
    bbox: FloatArray
    cells_per_leaf: int


NodeDType = np.dtype(
    [
        ("child", IntDType),
        ("Lmax", FloatDType),
        ("Rmin", FloatDType),
        ("ptr", IntDType),
        ("size", IntDType),
        ("dim", bool),
    ]
)
