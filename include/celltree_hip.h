/*
 * celltree_hip.h -- C ABI of libcelltree_hip.so, the MI355X (gfx950) cell-tree library.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no HIP or torch types.
 * Conventions:
 *   - every array pointer is a DEVICE pointer unless its comment says "host";
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream); every call is
 *     stream-ordered and asynchronous unless its comment says it synchronizes;
 *   - every function returns CT_OK (0) or a negative CT_ERR_* code and never throws;
 *     ct_last_error() returns the message of the calling thread's last failure;
 *   - result index arrays are int64 (numpy's intp, the reference's IntDType);
 *   - element arrays (faces, edges) are int32 on the device, padded with -1.
 *
 * Each declaration names the numba_celltree v0.4.2 function it replaces (file:line,
 * relative to the numba_celltree source tree).
 */
#ifndef CELLTREE_HIP_H
#define CELLTREE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_OK 0
#define CT_ERR_ARG (-1)        /* invalid argument (shape, size, null pointer) */
#define CT_ERR_HIP (-2)        /* a HIP runtime call failed */
#define CT_ERR_DEGENERATE (-3) /* tree build: a centroid falls in no bucket (the reference raises IndexError) */
#define CT_ERR_DEPTH (-4)      /* tree deeper than CT_MAX_DEPTH without ct_tree.parents */
#define CT_ERR_WORKSPACE (-5)  /* workspace smaller than the matching *_workspace_bytes() */

#define CT_MAX_N_VERTEX 32     /* constants.py:134 */
#define CT_MAX_DEPTH 254       /* deepest tree the stack traversals take; deeper trees are walked
                                  through parent links (ct_tree.parents, see ct_node_parents) */

/* One tree node, 32 bytes, the device form of NodeDType (constants.py:92-106).
 * child == -1 marks a leaf; the right child is child + 1. */
typedef struct ct_node {
  double Lmax;
  double Rmin;
  int32_t child;
  int32_t ptr;
  int32_t size;
  int32_t dim;
} ct_node;

/* A built tree: the device form of CellTreeData (constants.py:82-89). */
typedef struct ct_tree {
  const double* vertices;    /* (n_vertices, 2) */
  const int32_t* elements;   /* (n_elements, n_max_vert), faces (CCW) or edges (n_max_vert 2) */
  const double* bb_coords;   /* (n_elements, 4): xmin, xmax, ymin, ymax */
  const ct_node* nodes;      /* (n_nodes) */
  const int32_t* bb_indices; /* (n_elements) */
  int64_t n_vertices;
  int64_t n_elements;
  int64_t n_nodes;
  int32_t n_max_vert;
  int32_t depth;             /* longest root-to-leaf path, in edges */
  double bbox[4];            /* host values: tree bounding box */
  const double* leaf_coords; /* optional (may be NULL): (n_elements, n_max_vert, 2), row i =
                                the vertices of cell bb_indices[i], unused slots repeating
                                its last vertex; see ct_pack_leaf_coords */
  const double* node_boxes;  /* optional (may be NULL): (n_nodes + 1, 4), row j + 1 = the
                                tight box of node j's cells; see ct_node_boxes */
  const int32_t* start_grid; /* optional (may be NULL): (start_grid_n, start_grid_n) start
                                nodes of point walks over bbox; see ct_start_grid */
  int32_t start_grid_n;
  int32_t reserved;          /* 0 */
  const double* leaf_bb;     /* optional (may be NULL): (n_elements, 4), row k = the bbox of
                                cell bb_indices[k]; see ct_pack_leaf_bboxes */
  const int32_t* parents;    /* (n_nodes) parent of every node (-1 for the root), required when
                                depth > CT_MAX_DEPTH, else unused (may be NULL); see
                                ct_node_parents */
} ct_tree;

const char* ct_last_error(void);
int ct_version(void);

/* ---------------------------------------------------------------- geometry helpers */

/* geometry_utils.py:541-561 counter_clockwise: reorders clockwise rows in place. */
int ct_counter_clockwise(const double* vertices, int32_t* faces, int64_t n_faces,
                         int32_t n_max_vert, void* stream);

/* geometry_utils.py:443-456 build_face_bboxes -> out (n_faces, 4). */
int ct_build_face_bboxes(const int32_t* faces, int64_t n_faces, int32_t n_max_vert,
                         const double* vertices, double* out, void* stream);

/* geometry_utils.py:473-487 build_edge_bboxes (padded by tol) -> out (n_edges, 4). */
int ct_build_edge_bboxes(const int32_t* edges, int64_t n_edges, const double* vertices,
                         double tol, double* out, void* stream);

/* celltree_base.py:20-48 bbox_tree + bbox_distances.  distances: (n, 3) dx, dy, dxy;
 * out5: device array of 5 doubles = xmin, xmax, ymin, ymax, max(dxy) (the argument of
 * default_tolerance, celltree_base.py:51-52). */
int ct_bbox_stats(const double* bb_coords, int64_t n, double* distances, double* out5,
                  void* stream);

/* edge_celltree.py:60-64: out4 = xmin, xmax, ymin, ymax of the vertices. */
int ct_vertex_extent(const double* vertices, int64_t n, double* out4, void* stream);

/* ------------------------------------------------------------------------- build */

/* creation.py:216-230 pessimistic_n_nodes: capacity of the nodes_out array of ct_build. */
int64_t ct_max_nodes(int64_t n_elements);

/* Device workspace, in bytes, that ct_build needs for n_elements and n_buckets. */
size_t ct_build_workspace_bytes(int64_t n_elements, int32_t n_buckets);

/* creation.py:384-413 initialize (+ build, :233-381): level-synchronous bucket-split
 * build on the device.  Writes nodes_out[0 .. *n_nodes) in the reference's DFS numbering
 * and bb_indices_out (n_elements).  n_nodes and depth are HOST outputs; the call
 * synchronizes the stream once per tree level. */
int ct_build(const double* bb_coords, int64_t n_elements, int32_t n_buckets,
             int32_t cells_per_leaf, void* workspace, size_t workspace_bytes,
             ct_node* nodes_out, int32_t* bb_indices_out, int64_t* n_nodes, int32_t* depth,
             void* stream);

/* Host-side check of an imported tree (HOST arrays; no device work), the arrays a reference
 * tree holds as attributes (celltree.py:80-97) in this library's layouts: nodes (n_nodes
 * ct_node records), bb_indices (n_elements), elements (n_elements, n_max_vert).  Returns
 * CT_OK and the tree depth in *depth when every child link points forwards inside the node
 * array, every node but the root has exactly one parent, every node's cell range lies in
 * [0, n_elements), bb_indices is a permutation and every element index addresses one of
 * n_vertices vertices (allow_fill: -1 padding past a face's first three entries); else
 * CT_ERR_ARG with the first violation in ct_last_error().  The traversal kernels rely on
 * exactly these properties. */
int ct_check_tree(const ct_node* nodes, int64_t n_nodes, const int32_t* bb_indices, const int32_t* elements,
                  int64_t n_elements, int32_t n_max_vert, int64_t n_vertices, int32_t allow_fill, int32_t* depth);

/* query.py:568-622 collect_node_bounds -> out (n_nodes, 4); bbox is a HOST array of 4;
 * parent_scratch: device int32 array of n_nodes. */
int ct_collect_node_bounds(const ct_node* nodes, int64_t n_nodes, const double* bbox,
                           int32_t* parent_scratch, double* out, void* stream);

/* query.py:625-664 validate_node_bounds -> validity (n_nodes) of 0/1 bytes. */
int ct_validate_node_bounds(const ct_tree* tree, const double* node_bounds, uint8_t* validity,
                            void* stream);

/* ----------------------------------------------------------------------- queries */

/* query.py:110-117 locate_points: first containing face in DFS order, else -1. */
int ct_locate_points(const ct_tree* tree, const double* points, int64_t n, double tol,
                     int64_t* out, void* stream);

/* Layout helper, no reference counterpart: fills `out` (n_elements * n_max_vert * 2 doubles)
 * with the vertices of every cell in bb_indices order, slots past the cell's polygon
 * (polygon_length's rule) repeating its last vertex, so traversals read a leaf's cells
 * contiguously instead of through bb_indices -> faces -> vertices, and test every cell as
 * a fixed n_max_vert-edge loop (a repeated vertex adds a zero-length edge, which changes no
 * point-in-polygon answer).  Requires 3 <= n_max_vert <= 8. */
int ct_pack_leaf_coords(const ct_tree* tree, double* out, void* stream);

/* Layout helper, no reference counterpart: out (n_nodes) gets the parent of every node (-1
 * for the root).  Trees deeper than CT_MAX_DEPTH need it (ct_tree.parents): their queries
 * keep no traversal stack (the reference's grows without bound, utils.py:34-97) and find a
 * deferred sibling again by climbing these links.  Results are the same as the stack
 * traversal's. */
int ct_node_parents(const ct_tree* tree, int32_t* out, void* stream);

/* Layout helper, no reference counterpart: out (n_elements, 4) gets, at row k, the bbox
 * (xmin, xmax, ymin, ymax) of cell bb_indices[k], so box and segment queries of a tree with
 * leaf_bb set read a leaf's cell boxes contiguously instead of through bb_indices; their
 * results are the same with or without it. */
int ct_pack_leaf_bboxes(const ct_tree* tree, double* out, void* stream);

/* Layout helper, no reference counterpart: out (n_nodes + 1, 4) gets, at row j + 1, the
 * tight box (xmin, xmax, ymin, ymax) of node j's cells bb_indices[ptr .. ptr + size); any
 * NaN cell coordinate makes the box NaN.  Segment-face queries (ct_locate_edges_*, kind 1)
 * of a tree with node_boxes set skip subtrees the segment cannot reach; their results are
 * the same with or without it.  scratch: device int32 array of n_nodes.  Enqueues
 * 1 + depth kernels. */
int ct_node_boxes(const ct_tree* tree, int32_t* scratch, double* out, void* stream);

/* Layout helper, no reference counterpart: out (n, n) int32, row-major (grid row = y), gets
 * for every cell of an n x n grid over the tree bbox (cell (gx, gy) spans
 * [xmin + gx * w, xmin + (gx + 1) * w] with w = (xmax - xmin) / n, the same in y, and the
 * border cells reach to infinity outwards) the deepest node that every point walk through
 * that cell reaches with no deferred sibling: from the root, while the cell's span along the
 * node's split axis is on one side of the split for all its points (v <= Lmax and v < Rmin
 * throughout: left; v > Lmax and v >= Rmin throughout: right).  Point queries
 * (ct_locate_points*, ct_compute_barycentric_weights, ct_locate_points_on_edge) of a tree
 * with start_grid set begin each walk at its cell's node; their results are the same with
 * or without it.  Requires 1 <= n <= 4096. */
int ct_start_grid(const ct_tree* tree, int32_t n, int32_t* out, void* stream);

size_t ct_locate_workspace_bytes(int64_t n);

/* query.py:110-117 locate_points, same results as ct_locate_points: the queries are first
 * ordered with ct_query_order (kind 0), so the lanes of a wavefront walk nearly the same
 * root-to-leaf path; every result is written back to its query's own slot. */
int ct_locate_points_ordered(const ct_tree* tree, const double* points, int64_t n, double tol,
                             void* workspace, size_t workspace_bytes, int64_t* out,
                             void* stream);

/* query.py:110-117 locate_points with a caller-supplied processing order (from
 * ct_query_order kind 0, or any permutation of 0..n-1); same results as ct_locate_points.
 * `sorted` (may be NULL): the points in that order, sorted[i] = points[order[i]], as
 * ct_query_order writes them; lane i then reads its point contiguously. */
int ct_locate_points_in_order(const ct_tree* tree, const double* points, int64_t n,
                              const int32_t* order, const double* sorted, double tol, int64_t* out,
                              void* stream);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same results
 * as ct_locate_points_ordered (query.py:110-117 locate_points), with the ordering workspace
 * taken from a static device array inside the library instead of an argument: no hipMalloc,
 * hipFree or other allocation happens in the call.  n is at most CT_NOALLOC_MAX_POINTS (larger
 * n returns CT_ERR_ARG).  Each device keeps its own array (about 36 MiB) for the life of the
 * process, resolved on the first call on that device (make that first call outside graph
 * capture), and shared by every call on it: calls that overlap on one device race and must
 * be serialised by the caller (one stream, or events).  `stream` must belong to the current
 * device. */
#define CT_NOALLOC_MAX_POINTS (1 << 21)
int ct_locate_points_ordered_noalloc(const ct_tree* tree, const double* points, int64_t n, double tol,
                                     int64_t* out, void* stream);

/* query.py:167-174 locate_points_on_edge (EdgeCellTree2d). */
int ct_locate_points_on_edge(const ct_tree* tree, const double* points, int64_t n, double tol,
                             int64_t* out, void* stream);

/* Scheduling helper, no reference counterpart: a processing order for n queries (kind 0
 * points (n, 2), 1 boxes (n, 4), 2 segments (n, 2, 2)) from a bucket sort on a 20-bit key:
 * points by the Morton key of a 1024 x 1024 grid over `bbox` (HOST array of 4); boxes and
 * segments by a work class (log2 of their bbox area relative to `bbox`) above the Morton key
 * of their centre.  The order is grouped by the top 8 key bits and sorted by the full key
 * within chunks of 32768 queries of one group (a group larger than that is a sequence of
 * such sorted chunks).  Passing the order to a count / write pass gives each wavefront
 * neighbouring queries of similar cost; results are unchanged by it.  `sorted` (may be NULL)
 * receives the queries in that order (n rows of 2 or 4 doubles). */
size_t ct_order_workspace_bytes(int64_t n);
int ct_query_order(int32_t kind, const double* queries, int64_t n, const double* bbox,
                   void* workspace, size_t workspace_bytes, int32_t* order, double* sorted,
                   void* stream);

/* query.py:278-289 locate_boxes, first pass: counts[i] = number of tree cells whose bbox
 * strictly overlaps box i.  `order` (from ct_query_order, or NULL) only schedules; `sorted`
 * (NULL, or the boxes in that order from ct_query_order) lets lanes read them contiguously. */
int ct_locate_boxes_count(const ct_tree* tree, const double* boxes, int64_t n,
                          const int32_t* order, const double* sorted, int64_t* counts, void* stream);

/* query.py:278-289 locate_boxes, second pass: writes pairs at offsets[i] (from
 * ct_exclusive_scan), in box order then per-box DFS order (the reference's order). */
int ct_locate_boxes_write(const ct_tree* tree, const double* boxes, int64_t n,
                          const int32_t* order, const double* sorted, const int64_t* offsets,
                          int64_t* ii, int64_t* jj, void* stream);

/* The same two passes with a hit record (as ct_locate_edges_count_rec / _write_rec): the
 * count pass keeps the first hit_slots hit cells of every box in `hits` (device int32,
 * hit_slots * n, in [1, 1024]); the write pass copies them for boxes with at most hit_slots
 * hits instead of walking the tree again.  hits == NULL: ct_locate_boxes_count / _write. */
int ct_locate_boxes_count_rec(const ct_tree* tree, const double* boxes, int64_t n,
                              const int32_t* order, const double* sorted, int64_t* counts,
                              int32_t* hits, int32_t hit_slots, void* stream);
int ct_locate_boxes_write_rec(const ct_tree* tree, const double* boxes, int64_t n,
                              const int32_t* order, const double* sorted, const int64_t* offsets,
                              const int32_t* hits, int32_t hit_slots, int64_t* ii, int64_t* jj,
                              void* stream);

/* query.py:500-544 locate_edge_faces (kind 1) / locate_edge_edges (kind 0), first pass;
 * `order` and `sorted` as for ct_locate_boxes_count. */
int ct_locate_edges_count(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int32_t* order, const double* sorted, int64_t* counts, void* stream);

/* Second pass: ii, jj and xy (n_found, 2, 2) in segment order then DFS order; offsets
 * (n + 1) is the exclusive scan of the counts. */
int ct_locate_edges_write(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int32_t* order, const double* sorted, const int64_t* offsets,
                          int64_t* ii, int64_t* jj, double* xy, void* stream);

/* The same two passes with a hit record (query.py:500-544; the reference's query-twice
 * scheme walks the tree in both passes): the count pass also keeps the first hit_slots hit
 * cells of every segment in `hits` (device int32, hit_slots * n, in [1, 1024]), and the write
 * pass rebuilds the rows of segments with at most hit_slots hits from those cells (same
 * predicates, same DFS order, identical rows) instead of walking the tree again.  Both
 * passes take the same order / sorted.  hits == NULL: ct_locate_edges_count / _write. */
int ct_locate_edges_count_rec(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                              const int32_t* order, const double* sorted, int64_t* counts,
                              int32_t* hits, int32_t hit_slots, void* stream);
int ct_locate_edges_write_rec(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                              const int32_t* order, const double* sorted, const int64_t* offsets,
                              const int32_t* hits, int32_t hit_slots, int64_t* ii, int64_t* jj,
                              double* xy, void* stream);

/* ------------------------------------------------------------------ work splitting
 * No reference counterpart: the largest box and segment queries (query.py:177-244, 357-455)
 * walk hundreds of leaves in one lane and set a launch's tail.  A query is cut into
 * subqueries by subtree: its walk from the root is expanded into at most max_parts viable
 * subtrees, in the walk's visit order (the subtree with the most cells expanded first, while
 * it holds more than leaf_cells cells), and each subquery walks one subtree.  The subqueries
 * of a query, one after the other, emit exactly the query's rows in the query's order.
 * Queries expected to meet at most min_cells cells (a box: (w / c + 1) (h / c + 1), a segment
 * 1.5 |b - a| / c + 1, c = sqrt(tree bbox area / n_elements)) stay whole: one subquery at
 * the root.  Trees of depth <= CT_MAX_DEPTH. */
typedef struct ct_subquery {
  int32_t query;  /* query index */
  int32_t node;   /* subtree root (0: the whole walk, tree-bbox test included) */
} ct_subquery;

/* kind 1 boxes (n, 4), 2 segment-face, 3 segment-edge (segments (n, 2, 2)); `order` (may be
 * NULL) lists the queries in processing order (ct_query_order).  Count call (subs NULL):
 * parts[i] = number of subqueries of query order[i].  Write call: subs[part_offsets[i] + k]
 * (part_offsets: ct_exclusive_scan of parts), subqueries in processing order.
 * frontier_buf (may be NULL; device int32, n * max_parts): the count call keeps the frontiers
 * there and the write call copies them instead of computing them again; both calls must get
 * the same buffer.  1 <= max_parts <= 16; trees of fewer than 2**28 nodes. */
int ct_split_queries(const ct_tree* tree, int32_t kind, const double* queries, int64_t n, const int32_t* order,
                     int32_t min_cells, int32_t leaf_cells, int32_t max_parts, int64_t* parts,
                     const int64_t* part_offsets, ct_subquery* subs, int32_t* frontier_buf, void* stream);

/* After a count pass over the subqueries: counts[j] becomes the exclusive prefix of the
 * counts of its query's earlier subqueries, and per_query[q] the query's total (the input of
 * ct_exclusive_scan, whose output is query_offsets). */
int ct_subquery_offsets(const int32_t* order, int64_t n, const int64_t* part_offsets, int64_t* counts,
                        int64_t* per_query, void* stream);

/* query.py:278-289 locate_boxes over subqueries: count pass (counts per subquery, hit record
 * hits[slot * n_sub + j]) and write pass (rows of subquery j at query_offsets[q] + prefix[j],
 * copied from the record when they fit). */
int ct_locate_boxes_count_sub(const ct_tree* tree, const double* boxes, const ct_subquery* subs, int64_t n_sub,
                              int64_t* counts, int32_t* hits, int32_t hit_slots, void* stream);
int ct_locate_boxes_write_sub(const ct_tree* tree, const double* boxes, const ct_subquery* subs, int64_t n_sub,
                              const int64_t* query_offsets, const int64_t* prefix, const int32_t* hits,
                              int32_t hit_slots, int64_t* ii, int64_t* jj, void* stream);

/* query.py:500-544 locate_edge_faces (kind 1) / locate_edge_edges (kind 0) over subqueries.
 * Count pass: counts[j] per subquery and, with hits, the record of its first hits;
 * hit_xy (optional, with hits) keeps the clipped segment of each of the first xy_slots
 * (<= hit_slots) hits, hit_xy[4 * (slot * n_sub + j) + 0..3] = (x0, y0, x1, y1).
 * Then sub_offsets = ct_exclusive_scan(counts) (staging rows in processing order: subquery j
 * at sub_offsets[j], rows in DFS order) and ct_subquery_spans (per query: its row count, to
 * scan into the output offsets, and its first staging row).  Write pass: the staging rows
 * (query, cell, clipped segment, sort key t = (x0 - a) . (b - a)), copied from the record for
 * subqueries whose hits all have a clipped segment in it.  ct_sort_staged_rows then writes the
 * output rows: per query at offsets[q], ordered by t as sort_intersections_by_edge
 * (geometry_utils.py:564-574). */
int ct_locate_edges_count_sub(const ct_tree* tree, const double* segments, const ct_subquery* subs, int64_t n_sub,
                              int32_t kind, int64_t* counts, int32_t* hits, int32_t hit_slots, double* hit_xy,
                              int32_t xy_slots, void* stream);
int ct_subquery_spans(const int32_t* order, int64_t n, const int64_t* part_offsets, const int64_t* sub_offsets,
                      int64_t* per_query, int64_t* query_start, void* stream);
int ct_locate_edges_write_sub(const ct_tree* tree, const double* segments, const ct_subquery* subs, int64_t n_sub,
                              int32_t kind, const int64_t* sub_offsets, const int32_t* hits, int32_t hit_slots,
                              const double* hit_xy, int32_t xy_slots, int32_t* rows_query, int32_t* rows_cell,
                              double* rows_xy, double* rows_t, void* stream);
int ct_sort_staged_rows(const int64_t* offsets, const int64_t* query_start, const int32_t* rows_query,
                        const int32_t* rows_cell, const double* rows_xy, const double* rows_t, int64_t n_rows,
                        int64_t* ii, int64_t* jj, double* xy, void* stream);

/* geometry_utils.py:564-574 sort_intersections_by_edge on grouped output (ii ascending, the
 * rows of segment s in [offsets[s], offsets[s+1])): writes jj and xy (n_rows, 2, 2) stably
 * sorted within each segment by t = (xy[:,0] - a) . (b - a), NaN last, into jj_out / xy_out
 * (ii is unchanged by the sort). */
int ct_sort_intersections_by_edge(const double* segments, const int64_t* offsets,
                                  const int64_t* ii, const int64_t* jj, const double* xy,
                                  int64_t n_rows, int64_t* jj_out, double* xy_out, void* stream);

/* ------------------------------------------------------------- scan / compaction */

size_t ct_scan_workspace_bytes(int64_t n);

/* out[0] = 0, out[i+1] = out[i] + in[i]: exclusive offsets plus the total in out[n]. */
int ct_exclusive_scan(const int64_t* in, int64_t n, int64_t* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ pair kernels */

/* separating_axis.py:58-75 polygons_intersect -> out (n) of 0/1 bytes. */
int ct_polygons_intersect(const double* vertices_a, const double* vertices_b,
                          const int32_t* faces_a, int32_t n_max_a, const int32_t* faces_b,
                          int32_t n_max_b, const int64_t* indices_a, const int64_t* indices_b,
                          int64_t n, uint8_t* out, void* stream);

/* sutherland_hodgman.py:151-168 area_of_intersection (subject a, clipper b) -> out (n). */
int ct_area_of_intersection(const double* vertices_a, const double* vertices_b,
                            const int32_t* faces_a, int32_t n_max_a, const int32_t* faces_b,
                            int32_t n_max_b, const int64_t* indices_a, const int64_t* indices_b,
                            int64_t n, double* out, void* stream);

/* sutherland_hodgman.py:171-187 box_area_of_intersection -> out (n). */
int ct_box_area_of_intersection(const double* bbox_coords, const double* vertices,
                                const int32_t* faces, int32_t n_max_vert,
                                const int64_t* indices_bbox, const int64_t* indices_face,
                                int64_t n, double* out, void* stream);

/* celltree.py:186-269 locate_faces + intersect_faces, fused pair stage: for every bbox
 * candidate pair k, keep[k] = SAT(a, b) (separating_axis.py:58-75) and, for intersect
 * (with_area != 0), also area[k] = SH clip area and keep[k] &= area[k] > 0
 * (celltree.py:266-269).  keep is int64 0/1 (it feeds ct_exclusive_scan). */
int ct_face_pairs(const double* vertices_a, const int32_t* faces_a, int32_t n_max_a,
                  const double* vertices_b, const int32_t* faces_b, int32_t n_max_b,
                  const int64_t* ii, const int64_t* jj, int64_t n, int32_t with_area,
                  double* area, int64_t* keep, void* stream);

/* celltree.py:150-184 intersect_boxes pair stage: area[k] = box clip area, keep[k] = area > 0. */
int ct_box_pairs(const double* bbox_coords, const double* vertices, const int32_t* faces,
                 int32_t n_max_vert, const int64_t* ii, const int64_t* jj, int64_t n,
                 double* area, int64_t* keep, void* stream);

/* Stable compaction of (ii, jj, area): rows with keep[k] go to position offsets[k]
 * (offsets from ct_exclusive_scan over keep).  area/area_out may be NULL. */
int ct_compact_pairs(const int64_t* ii, const int64_t* jj, const double* area,
                     const int64_t* keep, const int64_t* offsets, int64_t n, int64_t* ii_out,
                     int64_t* jj_out, double* area_out, void* stream);

/* ----------------------------------------------------------------- barycentric */

/* barycentric_triangle.py:46-64 -> out (n, 3). */
int ct_barycentric_triangle_weights(const double* points, const int64_t* face_indices,
                                    int64_t n, const int32_t* faces, int32_t n_max_vert,
                                    const double* vertices, double tol, double* out,
                                    void* stream);

/* barycentric_wachspress.py:88-107 -> out (n, n_max_vert). */
int ct_barycentric_wachspress_weights(const double* points, const int64_t* face_indices,
                                      int64_t n, const int32_t* faces, int32_t n_max_vert,
                                      const double* vertices, double tol, double* out,
                                      void* stream);

/* celltree.py:306-351 compute_barycentric_weights, fused: locate each point (query.py:62-107)
 * then its weights, Wachspress when n_max_vert > 3 else triangle.  face_out (n),
 * weights (n, n_max_vert). */
int ct_compute_barycentric_weights(const ct_tree* tree, const double* points, int64_t n,
                                   double tol, int64_t* face_out, double* weights,
                                   void* stream);

/* ----------------------------------------------------- regridding (no reference counterpart)
 * Conservative area-weighted regridding y[t,:] = sum_k A_k F[s_k,:] / area_t over a fixed pair
 * list (tt[k] target, ss[k] source; grouped by target, ascending), the caller pattern of
 * examples/spatial_indexing_2d_grids.py:268-323, and its adjoint under fixed topology.  A_k is
 * the Sutherland-Hodgman area (subject = target face, clipper = source face, as in
 * intersect_faces); faces are counter-clockwise rows of at most 8 vertices. */

/* A_k for every pair. */
int ct_regrid_pair_areas(const double* tgt_vertices, const int32_t* tgt_faces, int32_t n_max_t,
                         const double* src_vertices, const int32_t* src_faces, int32_t n_max_s,
                         const int64_t* tt, const int64_t* ss, int64_t n_pairs, double* area,
                         void* stream);

/* Fan area of every face (geometry_utils.py:82-95). */
int ct_polygon_areas(const double* vertices, const int32_t* faces, int32_t n_max_vert, int64_t n,
                     double* out, void* stream);

/* y (n_tgt, n_fields) = segmented sum over each target's pairs [tgt_offsets[t], tgt_offsets[t+1]). */
int ct_regrid_apply(const int64_t* tgt_offsets, const int64_t* ss, const double* area,
                    const double* tgt_area, const double* F, int32_t n_fields, int64_t n_tgt,
                    double* y, void* stream);

/* gF (n_src, n_fields) = transpose, over perm (pairs sorted by source) and src_offsets;
 * src_tgt[m] = tt[perm[m]], the pairs' targets in that order. */
int ct_regrid_apply_transpose(const int64_t* src_offsets, const int64_t* perm, const int64_t* src_tgt,
                              const double* area, const double* tgt_area, const double* gy,
                              int32_t n_fields, int64_t n_src, double* gF, void* stream);

/* The fixed pair structure of a regrid (device arrays).  Pairs are grouped by target in
 * ascending order (intersect_faces order); corner lists map every vertex to the face corners
 * (face * n_max + slot, ascending) that use it. */
typedef struct ct_regrid_topology {
  const int32_t* tgt_faces;          /* (n_tgt, n_max_t), counter-clockwise, -1 padded */
  const int32_t* src_faces;          /* (n_src, n_max_s), counter-clockwise, -1 padded */
  const int64_t* tt;                 /* (n_pairs) target of each pair */
  const int64_t* ss;                 /* (n_pairs) source of each pair */
  const int64_t* tgt_offsets;        /* (n_tgt + 1): pairs of target t */
  const int64_t* src_offsets;        /* (n_src + 1): perm_by_src rows of source s */
  const int64_t* perm_by_src;        /* (n_pairs) pairs ordered by source (stable) */
  const int64_t* tgt_vertex_offsets; /* (n_tgt_vertices + 1) */
  const int64_t* tgt_vertex_corners; /* corners of the target faces, grouped by vertex */
  const int64_t* src_vertex_offsets; /* (n_src_vertices + 1) */
  const int64_t* src_vertex_corners; /* corners of the source faces, grouped by vertex */
  int64_t n_pairs;
  int64_t n_tgt;
  int64_t n_src;
  int64_t n_tgt_vertices;
  int64_t n_src_vertices;
  int32_t n_max_t;                   /* 3 .. 8 */
  int32_t n_max_s;                   /* 3 .. 8 */
  const int64_t* rest_targets;       /* optional (may be NULL): the targets whose corner sums the
                                        vertex backward's pair pass does not finish (see
                                        ct_regrid_rest_targets); NULL: every target is checked */
  int64_t n_rest_targets;
} ct_regrid_topology;

/* Fills rest_targets (device, capacity n_tgt) with the targets the vertex backward's pair pass
 * leaves to its per-target pass -- those without pairs and those whose pairs cross a block of
 * the pair pass -- and *n_out (HOST) with their count.  Depends on the topology only, so it
 * is called once per topology.  Synchronizes the stream. */
int ct_regrid_rest_targets(const ct_regrid_topology* topo, int64_t* rest_targets, int64_t* n_out, void* stream);

size_t ct_regrid_backward_workspace_bytes(const ct_regrid_topology* topo);

/* Writes dL/d(tgt_vertices) and dL/d(src_vertices) for dL/dy = gy: through every pair's
 * overlap area (the derivative of the clipped polygon's area along the input edges it lies
 * on) and through every target's area.  No atomics: per-pair corner gradients are reduced
 * per face, then gathered per vertex, in a fixed order (run-to-run identical). */
int ct_regrid_vertex_backward(const ct_regrid_topology* topo, const double* tgt_vertices,
                              const double* src_vertices, const double* F, const double* y,
                              const double* gy, const double* tgt_area, int32_t n_fields,
                              void* workspace, size_t workspace_bytes, double* g_tgt_vertices,
                              double* g_src_vertices, void* stream);

/* Adjoint of barycentric_triangle.py:28-64 with each point's face held fixed: for dL/dw
 * (n, 3) writes dL/dpoints (n, 2) (NULL to skip) and adds dL/dvertices (NULL to skip; the
 * caller zeroes it).  Points with face -1 get zero gradient. */
int ct_barycentric_triangle_backward(const double* points, const int64_t* face_indices, int64_t n,
                                     const int32_t* faces, int32_t n_max_vert,
                                     const double* vertices, const double* g_weights,
                                     double* g_points, double* g_vertices, void* stream);

/* Adjoint of barycentric_wachspress.py:38-85 with each point's face held fixed, and the
 * forward's branch held fixed (interior weights, or the linear interpolation along the first
 * edge within `tol`): for dL/dw (n, n_max_vert) writes dL/dpoints (n, 2) (NULL to skip) and
 * adds dL/dvertices (NULL to skip; the caller zeroes it).  3 <= n_max_vert <= 16.  Points with
 * face -1 get zero gradient. */
int ct_barycentric_wachspress_backward(const double* points, const int64_t* face_indices, int64_t n,
                                       const int32_t* faces, int32_t n_max_vert, const double* vertices, double tol,
                                       const double* g_weights, double* g_points, double* g_vertices, void* stream);

/* ------------------------------------------------------------------ batched trees */
/* No reference counterpart: many trees and their query sets in one persistent launch,
 * replacing a loop of per-mesh calls (celltree.py:216-239 locate_points, celltree.py:294-303
 * intersect_edges).  Results equal the per-tree calls'.  Every call clears `counters` with
 * a fill kernel and allocates nothing, so a sequence of batch calls may be captured in a
 * HIP graph; calls sharing one `counters` array must not overlap. */

/* A tile: up to 64 * k consecutive queries of one tree, taken by one wavefront. */
typedef struct ct_batch_tile {
  int64_t begin;  /* first query, relative to the tree's first query */
  int32_t count;  /* number of queries, >= 1 */
  int32_t tree;   /* index into ct_batch.trees */
} ct_batch_tile;

/* The batch description; the struct itself is a HOST value, its arrays are on the device. */
typedef struct ct_batch {
  const ct_tree* trees;         /* (n_trees) tree descriptors */
  const double* tolerances;     /* (n_trees) point-location tolerance per tree */
  const int64_t* query_offsets; /* (n_trees + 1): tree m owns queries [off[m], off[m + 1]) */
  const ct_batch_tile* tiles;   /* (n_tiles), grouped by queue */
  const int32_t* queue_starts;  /* (n_queues + 1): queue r owns tiles [s[r], s[r + 1]) */
  int32_t* counters;            /* (n_queues) scratch */
  int64_t n_tiles;
  int32_t n_trees;
  int32_t n_queues;             /* 1 .. 64; workgroup b starts on queue b % n_queues */
  int32_t max_n_vert;           /* largest n_max_vert over the trees, 3 .. 8 (faces) */
  int32_t max_depth;            /* largest tree depth */
  int32_t stack;                /* traversal stack class, >= max_depth: 16, 24, 32, 48, 64, 128, 256 */
  int32_t n_workgroups;         /* persistent grid size */
  int32_t block;                /* workgroup size: 64, 128 or 256 */
  int32_t reserved;
} ct_batch;

/* locate_points for every tree's queries: out[q] = the first containing cell of q's tree
 * (that tree's cell index), else -1.  points: (n_queries, 2) concatenated. */
int ct_batch_locate_points(const ct_batch* batch, const double* points, int64_t* out, void* stream);

/* intersect_edges (segment-face) for every tree's segments, first pass: counts[q]. */
int ct_batch_edges_count(const ct_batch* batch, const double* segments, int64_t* counts, void* stream);

/* Second pass: rows of query q at offsets[q] (ct_exclusive_scan of the counts), written only
 * when offsets[q + 1] <= capacity; ii = concatenated query index, jj = cell of q's tree,
 * xy (rows, 2, 2).  Rows are in DFS order per query; ct_sort_intersections_by_edge then gives
 * the reference's order. */
int ct_batch_edges_write(const ct_batch* batch, const double* segments, const int64_t* offsets,
                         int64_t capacity, int64_t* ii, int64_t* jj, double* xy, void* stream);

/* The two passes with a hit record (as ct_locate_edges_count_rec / _write_rec): `hits` is a
 * device int32 array of hit_slots * n_queries (slot-major over the concatenated query index),
 * hit_slots in [1, 1024].  hits == NULL: ct_batch_edges_count / _write. */
int ct_batch_edges_count_rec(const ct_batch* batch, const double* segments, int64_t n_queries,
                             int64_t* counts, int32_t* hits, int32_t hit_slots, void* stream);
int ct_batch_edges_write_rec(const ct_batch* batch, const double* segments, int64_t n_queries,
                             const int64_t* offsets, const int32_t* hits, int32_t hit_slots,
                             int64_t capacity, int64_t* ii, int64_t* jj, double* xy, void* stream);

/* geometry_utils.py:564-574 sort_intersections_by_edge over the rows of ct_batch_edges_write
 * (ii = concatenated segment index), like ct_sort_intersections_by_edge but with the row
 * count read on the device from offsets[n_segments]: rows sorted only if it is <= capacity.
 * A fixed grid of n_workgroups, so the call can be captured before the count is known. */
int ct_batch_sort_intersections(const double* segments, int64_t n_segments, const int64_t* offsets,
                                const int64_t* ii, const int64_t* jj, const double* xy, int64_t capacity,
                                int32_t n_workgroups, int64_t* jj_out, double* xy_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CELLTREE_HIP_H */
