// ct_deep.hip -- traversals of trees deeper than CT_MAX_DEPTH, without a stack.
//
// The stack traversals (ct_traverse.h; query.py:62-117, 177-244, 357-455) enter the near
// child of a node whose two children are both viable and defer the far one; at most one
// deferred node per level is pending, so they need a stack of `depth` entries, and their
// stack classes end at 256.  The reference's stacks grow without bound (utils.py:34-97), so
// deeper trees (strongly graded meshes) must still be walked.  Here no stack is kept: when a
// subtree is finished, the walk climbs parent links (ct_tree.parents) from the subtree's
// root to the nearest ancestor that had both children viable and whose near child is the
// subtree just left, and continues in that ancestor's far child.  That is exactly the node
// the stack would pop (the far child of the deepest such ancestor whose far child is not yet
// visited), because an ancestor's split test, re-evaluated for the same query, gives the
// same answer it gave on the way down.  The visit order -- and so every first-hit answer and
// every row order -- is the stack traversal's.  The optional layouts (packed leaves, start
// grid, node boxes, hit records) are not used here; results are the same without them.
#include "ct_deep.h"
#include "ct_internal.h"
#include "ct_primitives.h"
#include "ct_traverse.h"

namespace ct {
namespace {

// A split test: 0 = no child viable, 1 = only the left, 2 = only the right, 3 = both; for 3,
// *near is the child entered first and *far the deferred one.
struct PointSplit {  // query.py:93-105 (points and points on edges)
  P2 p;
  __device__ int operator()(const NodeLd& nd, int* near, int* far) const {
    const double v = nd.dim ? p.y : p.x;
    const bool left = v <= nd.Lmax;
    const bool right = v >= nd.Rmin;
    if (left && right) {
      const bool right_first = (nd.Lmax - v) < (v - nd.Rmin);
      *near = right_first ? nd.child + 1 : nd.child;
      *far = right_first ? nd.child : nd.child + 1;
      return 3;
    }
    return (left ? 1 : 0) | (right ? 2 : 0);
  }
};

struct BoxSplit {  // query.py:228-242: the left child is pushed first, the right visited first
  Box box;
  __device__ int operator()(const NodeLd& nd, int* near, int* far) const {
    const double bmin = nd.dim ? box.ymin : box.xmin;
    const double bmax = nd.dim ? box.ymax : box.xmax;
    const bool left = bmin <= nd.Lmax;
    const bool right = bmax >= nd.Rmin;
    *near = nd.child + 1;
    *far = nd.child;
    return (left ? 1 : 0) | (right ? 2 : 0);
  }
};

struct SegmentSplit {  // query.py:407-453 (the slab test of edge_query, without pruning)
  P2 a, b;
  double vx, vy;
  __device__ int operator()(const NodeLd& nd, int* near, int* far) const {
    const double dx = nd.dim ? vy : vx;
    const double an = nd.dim ? a.y : a.x;
    const double bn = nd.dim ? b.y : b.x;
    double dl, dr;
    if (dx > 0.0) {
      dl = nd.Lmax - an;
      dr = nd.Rmin - bn;
    } else {
      dl = nd.Lmax - bn;
      dr = nd.Rmin - an;
    }
    bool left = dl >= 0.0;
    bool right = dr <= 0.0;
    if (dx > 0.0) {
      if (left) left = (dl / dx) >= 0.0;
      if (right) right = (dr / dx) <= 1.0;
    } else if (dx < 0.0) {
      if (left) left = (1.0 - (dl / dx)) >= 0.0;
      if (right) right = (1.0 - (dr / dx)) <= 1.0;
    }
    *near = nd.child + 1;
    *far = nd.child;
    return (left ? 1 : 0) | (right ? 2 : 0);
  }
};

// The next node after the subtree rooted at `node` is finished, or -1 when the walk is over.
template <class Split>
__device__ __forceinline__ int climb(const ct_node* __restrict__ nodes, const int32_t* __restrict__ parents,
                                     int node, const Split& split) {
  while (node != 0) {
    const int p = parents[node];
    const NodeLd nd = ld_node(nodes, p);
    int near = -1, far = -1;
    if (split(nd, &near, &far) == 3 && node == near) return far;
    node = p;
  }
  return -1;
}

// Visits the leaves the stack traversal visits, in its order; leaf(nd) returns true to stop.
template <class Split, class Leaf>
__device__ __forceinline__ void deep_walk(const TreeView& t, const int32_t* __restrict__ parents, const Split& split,
                                          Leaf&& leaf) {
  int node = 0;
  while (node >= 0) {
    const NodeLd nd = ld_node(t.nodes, node);
    if (nd.child == -1) {
      if (leaf(nd)) return;
      node = climb(t.nodes, parents, node, split);
      continue;
    }
    int near = -1, far = -1;
    const int s = split(nd, &near, &far);
    if (s == 3) node = near;
    else if (s == 1) node = nd.child;
    else if (s == 2) node = nd.child + 1;
    else node = climb(t.nodes, parents, node, split);
  }
}

// query.py:62-107 / :120-164: first hit in DFS order, else -1.
template <int MAXV, bool ON_EDGE>
__global__ void __launch_bounds__(kBlock) k_locate_deep(TreeView t, const int32_t* __restrict__ parents,
                                                        const double* __restrict__ pts,
                                                        const int32_t* __restrict__ order, int64_t n, double tol,
                                                        int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  const P2 p = ld_point(pts, q);
  int found = -1;
  deep_walk(t, parents, PointSplit{p}, [&](const NodeLd& nd) {
    int f;
    if (leaf_hit<MAXV, 16, ON_EDGE, kLeafFaces>(t, p, tol, nd, &f)) {
      found = f;
      return true;
    }
    return false;
  });
  out[q] = found;
}

// query.py:177-244 locate_box: count pass (counts) or write pass (rows at offsets[q]).
template <bool WRITE>
__global__ void __launch_bounds__(kBlock) k_boxes_deep(TreeView t, const int32_t* __restrict__ parents,
                                                       const double* __restrict__ boxes,
                                                       const int32_t* __restrict__ order, int64_t n,
                                                       int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                       int64_t* __restrict__ ii, int64_t* __restrict__ jj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  const Box box = ld_box(boxes, q);
  int64_t cnt = 0;
  int64_t o = WRITE ? offsets[q] : 0;
  if (boxes_intersect(box, t.bbox)) {
    deep_walk(t, parents, BoxSplit{box}, [&](const NodeLd& nd) {
      for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
        const int c = t.bb_indices[k];
        if (boxes_intersect(box, leaf_box(t, k, c))) {
          if (WRITE) {
            ii[o] = q;
            jj[o] = c;
            ++o;
          } else {
            ++cnt;
          }
        }
      }
      return false;
    });
  }
  if (!WRITE) counts[q] = cnt;
}

// query.py:357-455 locate_edge: Cohen-Sutherland against the tree bbox, then the slab walk;
// count pass or write pass (rows in DFS order at offsets[q]).
template <bool WRITE, int MAXV, bool FACES>
__global__ void __launch_bounds__(kBlock) k_edges_deep(TreeView t, const int32_t* __restrict__ parents,
                                                       const double* __restrict__ segs,
                                                       const int32_t* __restrict__ order, int64_t n,
                                                       int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                       int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                                       double* __restrict__ xy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  const double4 s4 = reinterpret_cast<const double4*>(segs)[q];
  const P2 a = {s4.x, s4.y}, b = {s4.z, s4.w};
  int64_t cnt = 0;
  const int64_t o = WRITE ? offsets[q] : 0;
  P2 c0, d0;
  if (cohen_sutherland(a, b, t.bbox, &c0, &d0)) {
    deep_walk(t, parents, SegmentSplit{a, b, b.x - a.x, b.y - a.y}, [&](const NodeLd& nd) {
      for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
        const int c = t.bb_indices[k];
        P2 cc, dd;
        bool hit;
        if constexpr (FACES) hit = edge_face_box<MAXV>(t, c, leaf_box(t, k, c), a, b, &cc, &dd);
        else hit = edge_edge(t, c, a, b, &cc, &dd);
        if (hit) {
          if (WRITE) {
            ii[o + cnt] = q;
            jj[o + cnt] = c;
            reinterpret_cast<double4*>(xy)[o + cnt] = make_double4(cc.x, cc.y, dd.x, dd.y);
          }
          ++cnt;
        }
      }
      return false;
    });
  }
  if (!WRITE) counts[q] = cnt;
}

template <bool ON_EDGE>
int launch_locate(const TreeView& t, const int32_t* parents, const double* pts, const int32_t* order, int64_t n,
                  double tol, int64_t* out, hipStream_t st) {
  const unsigned g = grid_for(n);
  const int nm = t.nmax;
  if (ON_EDGE) k_locate_deep<4, true><<<g, kBlock, 0, st>>>(t, parents, pts, order, n, tol, out);
  else if (nm <= 3) k_locate_deep<3, false><<<g, kBlock, 0, st>>>(t, parents, pts, order, n, tol, out);
  else if (nm <= 4) k_locate_deep<4, false><<<g, kBlock, 0, st>>>(t, parents, pts, order, n, tol, out);
  else if (nm <= 8) k_locate_deep<8, false><<<g, kBlock, 0, st>>>(t, parents, pts, order, n, tol, out);
  else if (nm <= 16) k_locate_deep<16, false><<<g, kBlock, 0, st>>>(t, parents, pts, order, n, tol, out);
  else k_locate_deep<32, false><<<g, kBlock, 0, st>>>(t, parents, pts, order, n, tol, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

template <bool WRITE>
int launch_edges(const TreeView& t, const int32_t* parents, const double* segs, const int32_t* order, int64_t n,
                 bool faces, int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy,
                 hipStream_t st) {
  const unsigned g = grid_for(n);
  const int nm = t.nmax;
  // the vertex-count classes of the stack kernels (ct_query.hip launch_edges)
  if (!faces) k_edges_deep<WRITE, 4, false><<<g, kBlock, 0, st>>>(t, parents, segs, order, n, counts, offsets, ii, jj, xy);
  else if (nm <= 4) k_edges_deep<WRITE, 4, true><<<g, kBlock, 0, st>>>(t, parents, segs, order, n, counts, offsets, ii, jj, xy);
  else if (nm <= 8) k_edges_deep<WRITE, 8, true><<<g, kBlock, 0, st>>>(t, parents, segs, order, n, counts, offsets, ii, jj, xy);
  else if (nm <= 16) k_edges_deep<WRITE, 16, true><<<g, kBlock, 0, st>>>(t, parents, segs, order, n, counts, offsets, ii, jj, xy);
  else k_edges_deep<WRITE, 32, true><<<g, kBlock, 0, st>>>(t, parents, segs, order, n, counts, offsets, ii, jj, xy);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace

int deep_check(const ct_tree* tree) {
  if (tree->parents == nullptr) {
    set_error("tree depth %d exceeds %d: its queries need ct_tree.parents (ct_node_parents)", tree->depth,
              CT_MAX_DEPTH);
    return CT_ERR_DEPTH;
  }
  return CT_OK;
}

int deep_locate_points(const ct_tree* tree, const double* pts, const int32_t* order, int64_t n, double tol,
                       bool on_edge, int64_t* out, hipStream_t st) {
  int rc = deep_check(tree);
  if (rc || n == 0) return rc;
  const TreeView t = view(*tree);
  return on_edge ? launch_locate<true>(t, tree->parents, pts, order, n, tol, out, st)
                 : launch_locate<false>(t, tree->parents, pts, order, n, tol, out, st);
}

int deep_boxes(const ct_tree* tree, const double* boxes, const int32_t* order, int64_t n, bool write,
               int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, hipStream_t st) {
  int rc = deep_check(tree);
  if (rc || n == 0) return rc;
  const TreeView t = view(*tree);
  const unsigned g = grid_for(n);
  if (write) k_boxes_deep<true><<<g, kBlock, 0, st>>>(t, tree->parents, boxes, order, n, counts, offsets, ii, jj);
  else k_boxes_deep<false><<<g, kBlock, 0, st>>>(t, tree->parents, boxes, order, n, counts, offsets, ii, jj);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int deep_edges(const ct_tree* tree, const double* segs, const int32_t* order, int64_t n, bool faces, bool write,
               int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy, hipStream_t st) {
  int rc = deep_check(tree);
  if (rc || n == 0) return rc;
  const TreeView t = view(*tree);
  return write ? launch_edges<true>(t, tree->parents, segs, order, n, faces, counts, offsets, ii, jj, xy, st)
               : launch_edges<false>(t, tree->parents, segs, order, n, faces, counts, offsets, ii, jj, xy, st);
}

}  // namespace ct
