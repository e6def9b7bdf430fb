"""Golden cases transcribed from numba_celltree v0.4.2's own test-suite.

Each ``check_*`` function restates one reference test against an implementation
namespace ``impl`` that provides ``CellTree2d`` and ``EdgeCellTree2d`` with the
reference's API (and, for the algorithm cases, the module-level batch functions).
The same cases run against the CPU oracle (tests/test_oracle_golden.py) and the HIP
library (tests/test_gpu_golden.py).  Source of every case: tests/test_celltree.py,
tests/test_edgecelltree.py and tests/test_algorithms/*.py of the reference.
"""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")

# test_celltree.py:20-35
nodes2 = [[0.0, 0.0], [2.0, 0.0], [1.0, 2.0], [3.0, 2.0]]
faces2 = [[0, 1, 2], [1, 3, 2]]
fill_value = -1
nodes = np.array(nodes2, dtype=np.float64)
faces = np.array(faces2, dtype=np.intc)

# test_celltree.py:38-83
nodes21 = [
    (5, 1), (10, 1), (3, 3), (7, 3), (9, 4), (12, 4), (5, 5), (3, 7), (5, 7), (7, 7),
    (9, 7), (11, 7), (5, 9), (8, 9), (11, 9), (9, 11), (11, 11), (7, 13), (9, 13), (7, 15),
]
faces21 = [
    (0, 1, 3), (0, 2, 6), (0, 3, 6), (1, 3, 4), (1, 4, 5), (2, 6, 7), (6, 7, 8), (7, 8, 12),
    (6, 8, 9), (8, 9, 12), (9, 12, 13), (4, 5, 11), (4, 10, 11), (9, 10, 13), (10, 11, 14),
    (10, 13, 14), (13, 14, 15), (14, 15, 16), (15, 16, 18), (15, 17, 18), (17, 18, 19),
]

# test_celltree.py:200-215 (quad grid used by several cases)
GRID_NODES = np.array(
    [
        [0.0, 0.0], [0.0, 2.0], [2.0, 0.0], [2.0, 2.0], [4.0, 0.0], [4.0, 2.0],
        [6.0, 0.0], [6.0, 2.0], [0.0, 4.0], [2.0, 4.0], [4.0, 4.0], [6.0, 4.0],
    ]
)
MULTI_FACES = np.array([[0, 8, 9, 5, 2], [9, 11, 7, 5, -1], [4, 7, 6, -1, -1]], dtype=np.intc)


def load_disk():
    d = np.load(os.path.join(HERE, "disk_5_5.npz"))
    return d["vertices"].copy(), d["faces"].copy()


def load_xy_triangles():
    return np.loadtxt(os.path.join(DATA, "xy.txt"), dtype=float), np.loadtxt(
        os.path.join(DATA, "triangles.txt"), dtype=int
    )


def load_voronoi():
    return np.loadtxt(os.path.join(DATA, "voronoi_xy.txt"), dtype=float), np.loadtxt(
        os.path.join(DATA, "voronoi.txt"), dtype=int
    )


# ------------------------------------------------------------------ test_celltree.py
def check_init(impl):  # :95-104
    impl.CellTree2d(nodes, faces, fill_value)
    impl.CellTree2d(nodes, faces, fill_value, n_buckets=2, cells_per_leaf=1)
    impl.CellTree2d(nodes, faces, fill_value, n_buckets=4)
    impl.CellTree2d(nodes, faces, fill_value, cells_per_leaf=2)


def check_init_larger_mesh(impl):  # :107-112
    xy, tri = load_xy_triangles()
    impl.CellTree2d(xy, tri, fill_value, n_buckets=2)


def check_lists_and_types(impl):  # :115-124
    impl.CellTree2d(nodes2, faces2, fill_value)
    impl.CellTree2d(np.array(nodes2, dtype=np.float32), np.array(faces2, dtype=np.int32), fill_value)


def check_fill_value_conversion(impl):  # :127-131
    f = np.array([[0, 1, 2, -999], [1, 3, 2, -999]])
    tree = impl.CellTree2d(nodes, f, -999)
    assert tree.faces[0, -1] == -1
    assert tree.faces[1, -1] == -1


def check_triangle_lookup(impl):  # :167-178
    tree = impl.CellTree2d(nodes, faces, fill_value)
    point = np.array([[1.0, 1.0], [2.0, 1.0], [-1.0, 1.0]])
    assert np.array_equal(tree.locate_points(point), [0, 1, -1])


def check_triangle_lookup_tolerance(impl):  # :181-195
    tree = impl.CellTree2d(nodes, faces, fill_value)
    point = np.array([[-0.09, 0.0], [2.0, 1.0], [-1.0, 1.0]])
    assert np.array_equal(tree.locate_points(point, tolerance=1e-9), [-1, 1, -1])
    assert np.array_equal(tree.locate_points(point, tolerance=1e-1), [0, 1, -1])


def check_poly_lookup(impl):  # :198-256
    faces1 = np.array([[0, 2, 3, 1], [4, 6, 7, 5]], dtype=np.intc)
    faces_p = np.array([[0, 8, 9, 5, 2], [9, 11, 6, 2, 5]], dtype=np.intc)
    tree1 = impl.CellTree2d(GRID_NODES, faces1, fill_value, n_buckets=2, cells_per_leaf=1)
    point = np.array([[1.0, 1.0], [5.0, 1.0], [-1.0, 1.0]])
    assert np.array_equal(tree1.locate_points(point), [0, 1, -1])
    tree2 = impl.CellTree2d(GRID_NODES, faces_p, fill_value, n_buckets=2, cells_per_leaf=1)
    point = np.array([[1.0, 2.0], [5.0, 2.0], [-1.0, 2.0]])
    assert np.array_equal(tree2.locate_points(point), [0, 1, -1])


def check_multi_poly_lookup(impl):  # :259-313
    tree = impl.CellTree2d(GRID_NODES, MULTI_FACES, fill_value, n_buckets=2, cells_per_leaf=1)
    point = np.array([[1.0, 1.0], [5.0, 0.5], [5.0, 3.0], [-1.0, 1.0]])
    assert np.array_equal(tree.locate_points(point), [0, 2, 1, -1])
    point = np.array(
        [
            [-9e-9, 0.0], [2.0, -9e-9], [-9e-9, 1.0], [1.0, -9e-9],
            [-1.1e-8, 1.0], [1.0, -1.1e-8], [-1.1e-8, 0.0], [2.0, -1.5e-8],
        ]
    )
    assert np.array_equal(tree.locate_points(point), [-1] * 8)
    assert np.array_equal(tree.locate_points(point, tolerance=1e-8), [0, 0, 0, 0, -1, -1, -1, -1])


def check_multipoint(impl):  # :316-330
    tree = impl.CellTree2d(nodes21, faces21, fill_value)
    points = [(4.2, 3.0), (7.7, 13.5), (3.4, 7.000000001), (7.0, 5.0), (8.66, 10.99), (7.3, 0.74),
              (2.5, 5.5), (9.8, 12.3)]
    assert np.array_equal(tree.locate_points(points), (1, 20, 7, -1, -1, -1, -1, -1))


BOX_COORDS = np.array(
    [
        [1.0, 2.0, 1.0, 2.0], [4.0, 5.0, 0.0, 1.0], [4.0, 5.0, 2.0, 3.0], [-1.0, 0.0, 0.0, 4.0],
        [6.0, 8.0, 0.0, 4.0], [0.0, 6.0, -1.0, 0.0], [0.0, 6.0, 4.0, 5.0],
    ]
)


def check_box_lookup(impl):  # :333-374
    tree = impl.CellTree2d(GRID_NODES, MULTI_FACES, fill_value, n_buckets=2, cells_per_leaf=1)
    i, j = tree.locate_boxes(BOX_COORDS)
    assert np.array_equal(i, [0, 1, 2])
    assert np.array_equal(j, [0, 2, 1])
    i, j, _ = tree.intersect_boxes(BOX_COORDS)
    assert np.array_equal(i, [0, 1, 2])
    assert np.array_equal(j, [0, 2, 1])


EDGE_COORDS = np.array(
    [
        [[1.0, 1.0], [2.0, 2.0]], [[4.0, 3.0], [5.0, 4.0]], [[5.0, 0.0], [6.0, 1.0]],
        [[-2.0, -1.0], [0.0, 1.0]], [[-2.0, -1.0], [-2.0, -1.0]],
    ]
)


def check_edge_lookup(impl):  # :377-420
    tree = impl.CellTree2d(GRID_NODES, MULTI_FACES, fill_value, n_buckets=2, cells_per_leaf=1)
    i, j, xy = tree.intersect_edges(EDGE_COORDS)
    assert np.array_equal(i, [0, 1, 2])
    assert np.array_equal(j, [0, 1, 2])
    assert np.allclose(xy, EDGE_COORDS[:3])
    i, j, xy = tree.intersect_edges(EDGE_COORDS[:, ::-1])
    assert np.array_equal(i, [0, 1, 2])
    assert np.array_equal(j, [0, 1, 2])
    assert np.allclose(xy, EDGE_COORDS[:3][:, ::-1])


def check_example_material(impl):  # :423-525 (ordered goldens on the disk mesh)
    vertices, f = load_disk()
    vertices += 1.0
    vertices *= 5.0
    tree = impl.CellTree2d(vertices, f, -1)
    points = np.array([[-5.0, 1.0], [4.5, 2.5], [6.5, 4.5]])
    assert np.array_equal(tree.locate_points(points), [-1, 24, 63])

    box_coords = np.array([[4.0, 8.0, 4.0, 6.0], [0.0, 8.0, 8.0, 10.0], [10.0, 13.0, 2.0, 8.0]])
    expected_i = np.concatenate(
        [[0] * 22, [0] * 8 + [1] * 14, [1] * 9]
    )
    expected_j = np.concatenate(
        [
            [107, 109, 99, 106, 95, 93, 92, 80, 62, 73, 86, 97, 81, 79, 65, 63],
            [76, 57, 44, 59, 54, 58, 50, 45, 43, 31, 41, 36, 38, 27, 123, 112],
            [120, 117, 121, 108, 110, 115, 116, 101, 104, 103, 111, 88, 98],
            [90, 100, 72, 84, 85, 66, 75, 69],
        ]
    )
    i, j = tree.locate_boxes(box_coords)
    assert np.array_equal(i, expected_i)
    assert np.array_equal(j, expected_j)

    triangle_vertices = np.array([[5.0, 3.0], [7.0, 3.0], [7.0, 5.0], [0.0, 6.0], [4.0, 4.0], [6.0, 10.0]])
    triangles = np.array([[0, 1, 2], [3, 4, 5]])
    expected_i = np.concatenate([[0] * 10 + [1] * 12, [1] * 22, [1]])
    expected_j = np.concatenate(
        [
            [81, 79, 61, 65, 63, 76, 44, 59, 48, 40, 115, 116, 101],
            [104, 91, 103, 111, 88, 98, 83, 72, 84, 68, 74, 54, 58, 55],
            [67, 47, 35, 50, 66, 46, 53, 26, 34, 49, 56, 29, 37, 43],
            [36, 38, 27, 30],
        ]
    )
    i, j, _ = tree.intersect_faces(triangle_vertices, triangles, -1)
    assert np.array_equal(i, expected_i)
    assert np.array_equal(j, expected_j)

    edge_coords = np.array([[[0.0, 0.0], [10.0, 10.0]], [[0.0, 10.0], [10.0, 0.0]]])
    expected_i = np.concatenate([[0] * 18 + [1] * 5, [1] * 13])
    expected_j = np.concatenate(
        [
            [0, 2, 6, 11, 16, 21, 27, 36, 43, 62, 73, 80, 93],
            [96, 108, 112, 120, 124, 69, 75, 66, 72, 67, 55, 50, 58],
            [54, 57, 63, 59, 65, 61, 70, 64, 77, 71],
        ]
    )
    i, j, _ = tree.intersect_edges(edge_coords)
    assert np.array_equal(i, expected_i)
    assert np.array_equal(j, expected_j)


def check_barycentric_triangles(impl):  # :528-548
    tree = impl.CellTree2d(nodes, faces, fill_value)
    points = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 1.0]])
    fi, w = tree.compute_barycentric_weights(points)
    assert np.array_equal(fi, [0, 0, 1])
    assert np.allclose(w, [[1.0, 0.0, 0.0], [0.25, 0.25, 0.5], [0.5, 0.25, 0.25]])


def check_barycentric_triangle_quads(impl):  # :551-588
    n = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0], [4.0, 0.0], [4.0, 4.0]])
    f = np.array([[0, 1, 2, 3], [1, 4, 5, 2]])
    tree = impl.CellTree2d(n, f, -1)
    points = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 1.0]])
    fi, w = tree.compute_barycentric_weights(points)
    assert np.array_equal(fi, [0, 0, 0])
    assert np.allclose(w, [[1.0, 0.0, 0.0, 0.0], [0.25, 0.25, 0.25, 0.25], [0.0, 0.5, 0.5, 0.0]])


def check_node_bounds(impl):  # :591-603
    xy, f = load_voronoi()
    tree = impl.CellTree2d(xy, f, fill_value)
    nb = tree.node_bounds
    n_node = len(tree.celltree_data.nodes)
    xmin, xmax, ymin, ymax = tree.bbox
    assert nb.shape == (n_node, 4)
    assert (nb[:, 0] >= xmin).all()
    assert (nb[:, 1] <= xmax).all()
    assert (nb[:, 2] >= ymin).all()
    assert (nb[:, 3] <= ymax).all()


def check_node_validity(impl):  # :606-618 (pins DFS numbering: node 73 is a child of 71)
    xy, f = load_voronoi()
    tree = impl.CellTree2d(xy, f, fill_value)
    assert tree.validate_node_bounds().all()
    tree.nodes[71]["Lmax"] = -0.02319655
    validity = tree.validate_node_bounds()
    assert not validity[73]


def check_find_centroids(impl):  # :621-628
    xy, tri = load_xy_triangles()
    centroids = xy[tri].mean(axis=1)
    tree = impl.CellTree2d(xy, tri, fill_value)
    assert np.array_equal(np.arange(len(centroids)), tree.locate_points(centroids))


def check_to_dict_of_lists(impl):  # :631-639
    xy, tri = load_xy_triangles()
    tree = impl.CellTree2d(xy, tri, fill_value, n_buckets=4)
    d = tree.to_dict_of_lists()
    assert isinstance(d, dict)
    assert list(d.keys()) == list(range(len(tree.celltree_data.nodes)))
    assert max(len(v) for v in d.values()) == 2


def check_locate_point_on_edge(impl):  # :642-710
    n = np.array([[0.0, 0.0], [3.0, 0.0], [1.0, 1.0], [0.0, 2.0], [3.0, 2.0]])
    f = np.array([[0, 1, 2], [0, 2, 3], [2, 4, 3]])
    tree = impl.CellTree2d(n, f, -1, n_buckets=4)
    points = np.array([[0.0, 0.0]] + [[v, v] for v in (0.01, 0.05, 0.15, 0.25, 0.35, 0.45, 0.55, 0.65, 0.75)])
    assert (tree.locate_points(points) != -1).all()
    n = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0], [4.0, 0.0], [4.0, 4.0], [0.0, 4.0]])
    f = np.array([[0, 1, 2, 3], [1, 4, 5, 2], [3, 2, 5, 6]])
    tree = impl.CellTree2d(n, f, -1, n_buckets=4)
    points = np.array([[-1e-9, 1.0], [4.0 + 1e-9, 1.0], [2.0, 2.0], [3.0, -1e-9], [3.0, 4.0 + 1e-9]])
    np.testing.assert_array_equal(tree.locate_points(points, tolerance=0.9e-9), [-1, -1, 2, -1, -1])
    assert (tree.locate_points(points, tolerance=1.1e-9) != -1).all()


# -------------------------------------------------------------- test_edgecelltree.py
E_VERTICES = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [2.0, 1.0]], dtype=float)
E_EDGES = np.array([[0, 1], [1, 2], [2, 3]], dtype=np.int32)


def check_edge_tree_init(impl):  # test_edgecelltree.py:12-39
    tree = impl.EdgeCellTree2d(E_VERTICES, E_EDGES)
    assert tree.vertices.shape == (4, 2)
    assert tree.edges.shape == (3, 2)
    assert tree.n_buckets == 4
    assert tree.cells_per_leaf == 2
    assert tree.nodes.shape == (3,)
    assert tree.bb_indices.shape == (3,)
    assert tree.bb_coords.shape == (3, 4)
    assert tree.bbox.shape == (4,)
    np.testing.assert_array_equal(tree.bb_indices, [0, 1, 2])
    expected = np.array([[0.0, 1.0, 0.0, 0.0], [1.0, 2.0, 0.0, 0.0], [2.0, 2.0, 0.0, 1.0]])
    np.testing.assert_allclose(tree.bb_coords, expected, atol=1e-9)
    np.testing.assert_allclose(tree.bbox, [0.0, 2.0, 0.0, 1.0], atol=1e-9)


def check_edge_tree_locate_points(impl):  # test_edgecelltree.py:42-69
    tree = impl.EdgeCellTree2d(E_VERTICES, E_EDGES)
    pts = np.array([[0.5, 0.0], [1.5, 0.0], [2.0, 0.5]])
    np.testing.assert_array_equal(tree.locate_points(pts), [0, 1, 2])
    pts = np.array([[0.5, 0.5], [1.5, 0.5], [2.0, 0.5]])
    np.testing.assert_array_equal(tree.locate_points(pts), [-1, -1, 2])
    v = np.array([[171805.657000002, 563516.366], [171889.594000001, 563437.333000001]])
    tree = impl.EdgeCellTree2d(v, np.array([[0, 1]]))
    p = np.array([[171882.49385095935, 563444.0183244612]])
    np.testing.assert_array_equal(tree.locate_points(p, tolerance=1e-8), [0])
    np.testing.assert_array_equal(tree.locate_points(p, tolerance=1e-9), [0])


def check_edge_tree_intersect_edges(impl):  # test_edgecelltree.py:72-92
    tree = impl.EdgeCellTree2d(E_VERTICES, E_EDGES)
    edge_coords = np.array(
        [
            [[1.0, -1.0], [1.0, 1.0]], [[3.0, 1.0], [-1.0, -1.0]], [[0.0, -1.0], [0.0, 1.0]],
            [[-2.0, -1.0], [-3.0, -1.0]],
        ]
    )
    e, te, xy = tree.intersect_edges(edge_coords)
    np.testing.assert_array_equal(e, [0, 1, 1])
    np.testing.assert_array_equal(te, [0, 2, 1])
    np.testing.assert_allclose(xy, [[1.0, 0.0], [2.0, 0.5], [1.0, 0.0]], atol=1e-9)


TREE_CASES = [
    check_init, check_init_larger_mesh, check_lists_and_types, check_fill_value_conversion,
    check_triangle_lookup, check_triangle_lookup_tolerance, check_poly_lookup,
    check_multi_poly_lookup, check_multipoint, check_box_lookup, check_edge_lookup,
    check_example_material, check_barycentric_triangles, check_barycentric_triangle_quads,
    check_node_bounds, check_node_validity, check_find_centroids, check_to_dict_of_lists,
    check_locate_point_on_edge, check_edge_tree_init, check_edge_tree_locate_points,
    check_edge_tree_intersect_edges,
]

# ---------------------------------------------------- test_sutherland_hodgman.py data
SH_A = np.array(
    [
        [[0.98599114, 0.16203056], [0.64839124, 0.6552714], [0.44528724, 0.88567472]],
        [[0.96182162, 0.3642742], [0.03478739, 0.54268026], [0.57582971, 0.41541277]],
        [[0.32556365, 0.03800701], [0.74000686, 0.04684465], [0.89527188, 0.55061165]],
        [[0.2988294, 0.96608896], [0.01212383, 0.00144037], [0.75113002, 0.54797261]],
        [[0.06522962, 0.43735202], [0.791499, 0.5229509], [0.40651803, 0.94317979]],
        [[0.06544202, 0.16735701], [0.67916353, 0.95843272], [0.33545733, 0.86368003]],
        [[0.43129575, 0.27998206], [0.49468229, 0.75438255], [0.01542992, 0.80696797]],
        [[0.29449023, 0.32433138], [0.46157048, 0.22492393], [0.82442969, 0.75853821]],
        [[0.66113797, 0.88485505], [0.70164374, 0.24393423], [0.89565423, 0.89407158]],
        [[0.92226655, 0.82771688], [0.42243438, 0.17562404], [0.82885357, 0.17541439]],
    ]
)
SH_B = np.array(
    [
        [[0.8141854, 0.06821897], [0.37086004, 0.49067617], [0.79810508, 0.07873283]],
        [[0.74948185, 0.8942076], [0.59654411, 0.87755533], [0.3023107, 0.68256513]],
        [[0.46670989, 0.31716127], [0.68408985, 0.75792215], [0.41437824, 0.79509823]],
        [[0.60715923, 0.67648133], [0.40045464, 0.79676831], [0.06332723, 0.69679141]],
        [[0.24057248, 0.16433727], [0.58871277, 0.05499277], [0.59144784, 0.24476056]],
        [[0.23183198, 0.41619006], [0.66566902, 0.30110111], [0.60418791, 0.60702136]],
        [[0.09393344, 0.87976118], [0.994083, 0.00532686], [0.95176396, 0.79836557]],
        [[0.89063751, 0.5880825], [0.03881315, 0.82436939], [0.61391092, 0.45027842]],
        [[0.63168954, 0.75135847], [0.8726944, 0.06387274], [0.89585471, 0.92837592]],
        [[0.94379596, 0.64164962], [0.95787609, 0.65627618], [0.6212529, 0.89153053]],
    ]
)
# shapely areas (test_sutherland_hodgman.py:77-90)
SH_EXPECTED = np.array(
    [0.0, 0.0, 0.0, 0.0262324, 0.0, 0.00038042, 0.03629781, 0.01677156, 0.05417924, 0.00108787]
)


def check_area_of_intersection(impl):  # test_sutherland_hodgman.py:190-200
    va = SH_A.reshape(-1, 2)
    vb = SH_B.reshape(-1, 2)
    fa = np.arange(len(va)).reshape(-1, 3)
    fb = np.arange(len(vb)).reshape(-1, 3)
    idx = np.arange(len(fa))
    actual = impl.area_of_intersection(va, vb, fa, fb, idx, idx)
    assert np.allclose(actual, SH_EXPECTED)


def check_box_area_of_intersection(impl):  # test_sutherland_hodgman.py:203-234
    box_coords = np.array([[0.0, 1.0, 0.0, 1.0], [1.0, 2.0, 1.0, 2.0]])
    v = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [-2.0, 0.0], [-2.0, 2.0]])
    f = np.array([[0, 1, 2], [0, 3, 4]])
    actual = impl.box_area_of_intersection(box_coords, v, f, np.array([0, 0, 1, 1]), np.array([0, 1, 0, 1]))
    assert np.allclose(actual, [0.5, 0.0, 0.5, 0.0])


def check_polygons_intersect(impl):  # test_separating_axis.py:139-180
    va = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    fa = np.array([[0, 1, 2, 3]])
    vb = np.array([[0.25, 0.25], [0.75, 0.75], [0.6, 0.6], [2.0, 1.0], [3.0, 0.0], [3.0, 1.0]])
    fb = np.array([[0, 1, 2, -1], [3, 4, 5, -1]])
    actual = impl.polygons_intersect(va, vb, fa, fb, np.array([0, 0]), np.array([0, 1]))
    assert np.array_equal(actual, [True, False])


def check_barycentric_module_weights(impl):  # test_barycentric.py:282-326, 352-394
    points = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.5, 0.0], [1.0, 0.5], [0.5, 0.5], [2.0, 2.0]])
    v = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    f = np.array([[0, 1, 2]])
    fi = np.array([0, 0, 0, 0, 0, 0, -1])
    expected = np.array(
        [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5],
         [0.0, 0.0, 0.0]]
    )
    for fn in (impl.barycentric_triangle_weights, impl.barycentric_wachspress_weights):
        assert np.allclose(fn(points, fi, f, v, 1e-9), expected)
    v = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    f = np.array([[0, 1, 2, 3]])
    fi = np.array([0, 0, 0, 0, 0, 0, 0, -1])
    points = np.array(
        [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0], [0.25, 0.0], [0.25, 1.0], [0.5, 0.5], [2.0, 2.0]]
    )
    expected = np.array(
        [[1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0],
         [0.75, 0.25, 0.0, 0.0], [0.0, 0.0, 0.25, 0.75], [0.25, 0.25, 0.25, 0.25], [0.0, 0.0, 0.0, 0.0]]
    )
    assert np.allclose(impl.barycentric_wachspress_weights(points, fi, f, v, 1e-9), expected)


def check_counter_clockwise(impl):  # test_geometry_utils.py:319-348
    v = np.array([[0.0, 0.0], [0.5, 0.0], [1.0, 0.0], [1.0, 0.5], [1.0, 1.0], [0.0, 1.0]])
    ccw = np.array([[0, 2, 4, 5, -1, -1], [0, 1, 2, 3, 4, 5]], dtype=np.int64)
    cw = np.array([[5, 4, 2, 0, -1, -1], [5, 4, 3, 2, 1, 0]], dtype=np.int64)
    expected = ccw.copy()
    impl.counter_clockwise(v, ccw)
    assert np.array_equal(expected, ccw)
    impl.counter_clockwise(v, cw)
    assert np.array_equal(expected, cw)


def check_build_bboxes(impl):  # test_geometry_utils.py:196-234
    f = np.array([[0, 1, 2, -1], [0, 1, 2, 3]])
    v = np.array([[0.0, 5.0], [5.0, 0.0], [5.0, 5.0], [0.0, 5.0]])
    assert np.array_equal(impl.build_face_bboxes(f, v), [[0.0, 5.0, 0.0, 5.0], [0.0, 5.0, 0.0, 5.0]])
    v = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [2.0, 1.0]], dtype=float)
    e = np.array([[0, 1], [1, 2], [2, 3]], dtype=np.int32)
    np.testing.assert_allclose(
        impl.build_edge_bboxes(e, v, 1e-9),
        [[0.0, 1.0, 0.0, 0.0], [1.0, 2.0, 0.0, 0.0], [2.0, 2.0, 0.0, 1.0]],
        atol=1e-9,
    )


def check_sort_intersections_by_edge(impl):  # test_geometry_utils.py:476-505
    edge_coords = np.array([[[0.0, 0.0], [1.0, 1.0]], [[2.0, 2.0], [3.0, 2.0]]])
    _ii = np.array([0, 0, 1, 1])
    _jj = np.array([0, 1, 2, 3])
    _xy = np.array(
        [
            [[0.5, 0.5], [1.0, 1.0]], [[0.0, 0.0], [0.5, 0.5]], [[2.0, 2.0], [2.5, 2.0]],
            [[2.5, 3.0], [3.0, 2.0]],
        ]
    )
    ii, jj, xy = impl.sort_intersections_by_edge(_ii, _jj, _xy, edge_coords)
    np.testing.assert_array_equal(ii, _ii)
    np.testing.assert_array_equal(jj, [1, 0, 2, 3])
    np.testing.assert_array_equal(
        xy,
        [[[0.0, 0.0], [0.5, 0.5]], [[0.5, 0.5], [1.0, 1.0]], [[2.0, 2.0], [2.5, 2.0]], [[2.5, 3.0], [3.0, 2.0]]],
    )


MODULE_CASES = [
    check_area_of_intersection, check_box_area_of_intersection, check_polygons_intersect,
    check_barycentric_module_weights, check_counter_clockwise, check_build_bboxes,
    check_sort_intersections_by_edge,
]


# -------------------------------------------------------------- test_celltree_utils.py
UTIL_BBOX_COORDS = np.array(  # test_celltree_utils.py:6-16
    [
        [1.0, 2.0, 1.0, 2.0], [4.0, 5.0, 0.0, 1.0], [4.0, 5.0, 2.0, 3.0], [-1.0, 0.0, 0.0, 4.0],
        [6.0, 8.0, 0.0, 4.0], [0.0, 6.0, -1.0, 0.0], [0.0, 6.0, 4.0, 5.0],
    ]
)
MIN_TOLERANCE = 1e-15  # constants.py
TOLERANCE_FACTOR = 1e-12


def check_default_tolerance(impl):  # test_celltree_utils.py:19-30
    bb_diagonal = np.array([2.4, 0.5])
    np.testing.assert_allclose(
        impl.default_tolerance(bb_diagonal), 2.4 * TOLERANCE_FACTOR, rtol=0, atol=MIN_TOLERANCE / 1e5
    )
    np.testing.assert_allclose(
        impl.default_tolerance(bb_diagonal / 1e4), MIN_TOLERANCE, rtol=0, atol=MIN_TOLERANCE / 1e5
    )


def check_bbox_tree(impl):  # test_celltree_utils.py:33-36
    np.testing.assert_allclose(impl.bbox_tree(UTIL_BBOX_COORDS), [-1.0, 8.0, -1.0, 5.0], rtol=0, atol=1e-5)


def check_bbox_distances(impl):  # test_celltree_utils.py:39-52
    expected = np.array(
        [
            [1.0, 1.0, 1.41421356], [1.0, 1.0, 1.41421356], [1.0, 1.0, 1.41421356], [1.0, 4.0, 4.12310563],
            [2.0, 4.0, 4.47213595], [6.0, 1.0, 6.08276253], [6.0, 1.0, 6.08276253],
        ]
    )
    np.testing.assert_allclose(impl.bbox_distances(UTIL_BBOX_COORDS), expected, rtol=0, atol=1e-5)


def check_bbox_distances_pythagoras(impl):
    """Mathematics, not a reference vector: the third column is the Euclidean diagonal
    (correctly rounded sqrt of dx^2 + dy^2 agrees with hypot to 1 ulp), on boxes spanning
    many magnitudes, including zero-width ones."""
    rng = np.random.default_rng(11)
    lo = rng.uniform(-1e3, 1e3, (500, 2))
    ext = 10.0 ** rng.uniform(-6, 4, (500, 2))
    ext[:5, 0] = 0.0
    bb = np.column_stack([lo[:, 0], lo[:, 0] + ext[:, 0], lo[:, 1], lo[:, 1] + ext[:, 1]])
    d = np.asarray(impl.bbox_distances(bb))
    np.testing.assert_array_equal(d[:, 0], bb[:, 1] - bb[:, 0])
    np.testing.assert_array_equal(d[:, 1], bb[:, 3] - bb[:, 2])
    np.testing.assert_allclose(d[:, 2], np.hypot(d[:, 0], d[:, 1]), rtol=4e-16, atol=0)
    assert (d[:, 2] >= np.maximum(d[:, 0], d[:, 1])).all()


def check_default_tolerance_of_tree(impl):
    """The default point tolerance is 1e-12 x the LARGEST cell diagonal: a unit square next to
    a 1000 x 1 strip gives tol = 1e-12 * sqrt(1000^2 + 1) ~ 1.0000005e-9, so a point 5e-10
    outside the square's left edge is on it (a tolerance from the smallest diagonal, 1.4e-12,
    would miss it) and a point 2e-9 outside is not (a diagonal without the sqrt, 1e6, would
    give 1e-6 and find it)."""
    v = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1001.0, 0.0], [1001.0, 1.0]])
    f = np.array([[0, 1, 2, 3], [1, 4, 5, 2]])
    tree = impl.CellTree2d(v, f, -1)
    pts = np.array([[-5e-10, 0.5], [-2e-9, 0.5], [0.5, -5e-10], [0.5, 1.0 + 2e-9]])
    np.testing.assert_array_equal(tree.locate_points(pts), [0, -1, 0, -1])


UTIL_CASES = [
    check_default_tolerance, check_bbox_tree, check_bbox_distances, check_bbox_distances_pythagoras,
    check_default_tolerance_of_tree,
]
