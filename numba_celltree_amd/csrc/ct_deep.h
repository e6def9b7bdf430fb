// ct_deep.h -- queries of trees deeper than CT_MAX_DEPTH (ct_deep.hip): stackless walks
// through ct_tree.parents.  Each returns CT_ERR_DEPTH when the tree has no parent links.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/celltree_hip.h"

namespace ct {

// CT_OK, or CT_ERR_DEPTH (with the message set) when tree->parents is null.
int deep_check(const ct_tree* tree);

// query.py:62-117 locate_points (on_edge: :120-174 locate_points_on_edge).
int deep_locate_points(const ct_tree* tree, const double* pts, const int32_t* order, int64_t n, double tol,
                       bool on_edge, int64_t* out, hipStream_t st);

// query.py:177-289 locate_boxes, count pass (write = false) or write pass.
int deep_boxes(const ct_tree* tree, const double* boxes, const int32_t* order, int64_t n, bool write,
               int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, hipStream_t st);

// query.py:357-544 locate_edge_faces (faces) / locate_edge_edges, count or write pass.
int deep_edges(const ct_tree* tree, const double* segs, const int32_t* order, int64_t n, bool faces, bool write,
               int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy, hipStream_t st);

}  // namespace ct
