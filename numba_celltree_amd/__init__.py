"""numba_celltree_amd: the cell-tree hot path of numba_celltree, native to MI355X (gfx950).

Drop-in classes with the reference's API (numba_celltree/__init__.py:1-6):
``CellTree2d`` and ``EdgeCellTree2d``.  All computation runs in libcelltree_hip.so
(hand-written HIP kernels); torch provides device memory and streams.
"""

from numba_celltree_amd.celltree import CellTree2d
from numba_celltree_amd.edge_celltree import EdgeCellTree2d

__version__ = "0.1.0"

__all__ = ("CellTree2d", "EdgeCellTree2d")
