"""Freeze the unit-disk triangle mesh that the reference's "example material" test uses.

The reference test (tests/test_celltree.py:86-92, 423-525) triangulates points on
concentric rings of the unit disk with matplotlib's Delaunay (numba_celltree/demo.py:91-123,
``generate_disk(5, 5)``) and orders the triangles by the squared distance of their
centroid to (-1, -1).  The golden index arrays in that test refer to this exact
mesh, so it is generated once here and stored as data (``disk_5_5.npz``); the tests
never depend on matplotlib or on Delaunay tie-breaking.

Ring layout: ring r (r = 0..depth) holds max(1, partitions*r) points at radius
r/depth.  Ring angles start at pi/2 and advance by a constant step; they are formed
from a running sum over all points, minus the sum at the ring's first point, which
reproduces the reference's floating-point angles exactly.

Run:  python tests/golden/make_disk_mesh.py
"""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def disk_points(partitions: int, depth: int) -> np.ndarray:
    rings = depth + 1
    counts = partitions * np.arange(rings)
    counts[0] = 1
    step = np.repeat((2 * np.pi) / counts, counts)
    running = step.cumsum()
    first = np.repeat(np.concatenate([[0], counts.cumsum()[:-1]]), counts)
    angle = running - running[first] + 0.5 * np.pi
    radius = np.repeat(np.linspace(0.0, 1.0, rings), counts)
    return np.column_stack((np.cos(angle) * radius, np.sin(angle) * radius))


def main() -> None:
    import matplotlib.tri

    xy = disk_points(5, 5)
    triangles = matplotlib.tri.Triangulation(xy[:, 0], xy[:, 1]).triangles
    centroids = xy[triangles].mean(axis=1)
    rsquared = (centroids[:, 0] - -1.0) ** 2 + (centroids[:, 1] - -1.0) ** 2
    order = np.argsort(rsquared)
    np.savez(os.path.join(HERE, "disk_5_5.npz"), vertices=xy, faces=triangles[order].astype(np.int64))


if __name__ == "__main__":
    main()
