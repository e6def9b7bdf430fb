"""The HIP library against every golden vector of the reference's own test-suite."""

import sys
import types

import numpy as np
import pytest

from tests.golden import cases as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def impl():
    import numba_celltree_amd as M
    from numba_celltree_amd import algorithms as A

    return types.SimpleNamespace(
        CellTree2d=M.CellTree2d,
        EdgeCellTree2d=M.EdgeCellTree2d,
        area_of_intersection=A.area_of_intersection,
        box_area_of_intersection=A.box_area_of_intersection,
        polygons_intersect=A.polygons_intersect,
        barycentric_triangle_weights=A.barycentric_triangle_weights,
        barycentric_wachspress_weights=A.barycentric_wachspress_weights,
        counter_clockwise=A.counter_clockwise,
        build_face_bboxes=A.build_face_bboxes,
        build_edge_bboxes=A.build_edge_bboxes,
        sort_intersections_by_edge=A.sort_intersections_by_edge,
        bbox_tree=A.bbox_tree,
        bbox_distances=A.bbox_distances,
        default_tolerance=A.default_tolerance,
    )


@pytest.mark.parametrize("case", C.TREE_CASES + C.MODULE_CASES + C.UTIL_CASES, ids=lambda f: f.__name__)
def test_reference_case(impl, case):
    print(f"[gpu golden] {case.__name__}", file=sys.stderr, flush=True)
    case(impl)


def test_shape_errors(impl):  # test_celltree.py:134-164
    faces = [0, 1, 2, 1, 3, 2]
    nodes = [(1, 2, 3), (3, 4, 5), (4, 5, 6)]
    with pytest.raises(ValueError):
        impl.CellTree2d(nodes, C.faces2, -1)
    with pytest.raises(ValueError):
        impl.CellTree2d(C.nodes2, faces, -1)
    tree = impl.CellTree2d(C.nodes2, C.faces2, -1)
    with pytest.raises(ValueError):
        tree.locate_points(nodes)
    with pytest.raises(ValueError):
        tree.intersect_faces(nodes, C.faces2, -1)
    with pytest.raises(ValueError):
        tree.intersect_faces(faces, C.nodes2, -1)
    with pytest.raises(ValueError):
        tree.locate_boxes(np.array([0.0, 1.0, 2.0, 3.0]))
    with pytest.raises(ValueError):
        tree.intersect_boxes(np.array([0.0, 1.0, 2.0, 3.0]))
    with pytest.raises(ValueError):
        tree.intersect_edges(np.array([[0.0, 1.0], [2.0, 3.0]]))


def test_device_tensor_io(impl):
    """Device tensors in give device tensors out, equal to the host path."""
    import torch

    tree = impl.CellTree2d(C.nodes21, C.faces21, -1)
    pts = np.array([(4.2, 3.0), (7.7, 13.5), (3.4, 7.000000001), (7.0, 5.0)])
    host = tree.locate_points(pts)
    dev = tree.locate_points(torch.from_numpy(pts).cuda())
    assert isinstance(dev, torch.Tensor) and dev.is_cuda and dev.dtype == torch.int64
    assert np.array_equal(dev.cpu().numpy(), host)
