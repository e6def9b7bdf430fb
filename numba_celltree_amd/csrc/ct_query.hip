// ct_query.hip -- tree traversals on the device.
//
// One lane per query, depth-first with a private stack sized from the tree depth
// (template STACK >= depth + 1, so it can never overflow: every pop pushes at most
// two children, one level deeper).  Visit order, push order and leaf predicates are the
// reference's (query.py:62-455), so first-hit answers and the order of variable-length
// results match it exactly.  Variable-length queries run twice: a count pass, a device
// exclusive scan (ct_exclusive_scan), and a write pass at the scanned offsets -- no
// atomic append, output order = query order then DFS order.
#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {
namespace {

struct TreeView {
  const double* vertices;
  const int32_t* elements;
  const double* bb;
  const ct_node* nodes;
  const int32_t* bb_indices;
  int nmax;
  int n_elements;
  Box bbox;
};

TreeView view(const ct_tree* t) {
  return TreeView{t->vertices, t->elements, t->bb_coords, t->nodes, t->bb_indices, t->n_max_vert,
                  (int)t->n_elements, Box{t->bbox[0], t->bbox[1], t->bbox[2], t->bbox[3]}};
}

struct NodeLd {
  double Lmax, Rmin;
  int child, ptr, size, dim;
};

// A node is 32 bytes, 32-byte aligned: two 16-byte loads.
__device__ __forceinline__ NodeLd ld_node(const ct_node* __restrict__ nodes, int i) {
  const double2 a = reinterpret_cast<const double2*>(nodes)[2 * i];
  const int4 b = reinterpret_cast<const int4*>(nodes)[2 * i + 1];
  return NodeLd{a.x, a.y, b.x, b.y, b.z, b.w};
}

__device__ __forceinline__ P2 ld_point(const double* __restrict__ pts, int64_t i) {
  const double2 d = reinterpret_cast<const double2*>(pts)[i];
  return {d.x, d.y};
}

// ----------------------------------------------------------------- points
// query.py:62-107 locate_point (faces) / :120-164 locate_point_on_edge (edges).
template <int MAXV, int STACK, bool ON_EDGE>
__device__ __forceinline__ int locate_one(const TreeView& t, P2 p, double tol) {
  int stack[STACK];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int ni = stack[--sp];
    const NodeLd nd = ld_node(t.nodes, ni);
    if (nd.child == -1) {
      for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
        const int c = t.bb_indices[k];
        if constexpr (ON_EDGE) {
          const int2 e = reinterpret_cast<const int2*>(t.elements)[c];
          if (point_on_edge(p, ld_vertex(t.vertices, e.x), ld_vertex(t.vertices, e.y), tol)) return c;
        } else {
          P2 poly[MAXV];
          const int len = load_poly<MAXV>(t.elements, t.nmax, c, t.vertices, poly);
          if (point_in_poly_or_edge<MAXV>(p, poly, len, tol)) return c;
        }
      }
      continue;
    }
    const double v = nd.dim ? p.y : p.x;
    const bool left = v <= nd.Lmax;
    const bool right = v >= nd.Rmin;
    const int lc = nd.child, rc = lc + 1;
    if (left && right) {
      // The side with the larger margin is popped first (query.py:93-101).
      if ((nd.Lmax - v) < (v - nd.Rmin)) {
        stack[sp++] = lc;
        stack[sp++] = rc;
      } else {
        stack[sp++] = rc;
        stack[sp++] = lc;
      }
    } else if (left) {
      stack[sp++] = lc;
    } else if (right) {
      stack[sp++] = rc;
    }
  }
  return -1;
}

template <int MAXV, int STACK, bool ON_EDGE>
__global__ void __launch_bounds__(kBlock) k_locate_points(TreeView t, const double* __restrict__ pts,
                                                          int64_t n, double tol,
                                                          int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = locate_one<MAXV, STACK, ON_EDGE>(t, ld_point(pts, i), tol);
}

// ----------------------------------------------------------------- boxes
// query.py:177-244 locate_box: strict overlap with the tree bbox first; left pushed before
// right, so the right child is popped first.
template <bool WRITE, int STACK>
__global__ void __launch_bounds__(kBlock) k_boxes(TreeView t, const double* __restrict__ boxes,
                                                  int64_t n, int64_t* __restrict__ counts,
                                                  const int64_t* __restrict__ offsets,
                                                  int64_t* __restrict__ ii, int64_t* __restrict__ jj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Box box = ld_box(boxes, i);
  int64_t cnt = 0;
  int64_t o = WRITE ? offsets[i] : 0;
  if (boxes_intersect(box, t.bbox)) {
    int stack[STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int ni = stack[--sp];
      const NodeLd nd = ld_node(t.nodes, ni);
      if (nd.child == -1) {
        for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
          const int c = t.bb_indices[k];
          if (boxes_intersect(box, ld_box(t.bb, c))) {
            if (WRITE) {
              ii[o] = i;
              jj[o] = c;
              ++o;
            } else {
              ++cnt;
            }
          }
        }
        continue;
      }
      const double bmin = nd.dim ? box.ymin : box.xmin;
      const double bmax = nd.dim ? box.ymax : box.xmax;
      const bool left = bmin <= nd.Lmax;
      const bool right = bmax >= nd.Rmin;
      if (left) stack[sp++] = nd.child;
      if (right) stack[sp++] = nd.child + 1;
    }
  }
  if (!WRITE) counts[i] = cnt;
}

// ----------------------------------------------------------------- segments
// query.py:309-328 compute_edge_face_intersect: Cohen-Sutherland on the cell bbox, then
// Cyrus-Beck on the cell with tol = max(1e-15, 1e-12 * max(w, h)).
template <int MAXV>
__device__ __forceinline__ bool edge_face(const TreeView& t, int c, P2 a, P2 b, P2* cc, P2* dd) {
  const Box box = ld_box(t.bb, c);
  if (!cohen_sutherland(a, b, box, cc, dd)) return false;
  P2 poly[MAXV];
  const int len = load_poly<MAXV>(t.elements, t.nmax, c, t.vertices, poly);
  const double tol = pmax(kMinTolerance, kToleranceFactor * pmax(box.xmax - box.xmin, box.ymax - box.ymin));
  return cyrus_beck<MAXV>(a, b, poly, len, tol, cc, dd);
}

// query.py:292-306 compute_edge_edge_intersect: (c, c).
__device__ __forceinline__ bool edge_edge(const TreeView& t, int c, P2 a, P2 b, P2* cc, P2* dd) {
  const int2 e = reinterpret_cast<const int2*>(t.elements)[c];
  P2 x;
  const bool r = lines_intersect(a, b, ld_vertex(t.vertices, e.x), ld_vertex(t.vertices, e.y), &x);
  *cc = x;
  *dd = x;
  return r;
}

// query.py:357-455 locate_edge: Cohen-Sutherland against the tree bbox, then slab pruning
// per node (the t-tests of :428-441 always hold once the slab test passed, but are kept).
template <bool WRITE, int MAXV, int STACK, bool FACES>
__global__ void __launch_bounds__(kBlock) k_edges(TreeView t, const double* __restrict__ segs,
                                                  int64_t n, int64_t* __restrict__ counts,
                                                  const int64_t* __restrict__ offsets,
                                                  int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                                  double* __restrict__ xy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 s4 = reinterpret_cast<const double4*>(segs)[i];
  const P2 a = {s4.x, s4.y}, b = {s4.z, s4.w};
  int64_t cnt = 0;
  int64_t o = WRITE ? offsets[i] : 0;
  P2 c0, d0;
  if (cohen_sutherland(a, b, t.bbox, &c0, &d0)) {
    const double vx = b.x - a.x, vy = b.y - a.y;
    int stack[STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int ni = stack[--sp];
      const NodeLd nd = ld_node(t.nodes, ni);
      if (nd.child == -1) {
        for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
          const int c = t.bb_indices[k];
          P2 cc, dd;
          const bool hit = FACES ? edge_face<MAXV>(t, c, a, b, &cc, &dd) : edge_edge(t, c, a, b, &cc, &dd);
          if (hit) {
            if (WRITE) {
              ii[o] = i;
              jj[o] = c;
              reinterpret_cast<double4*>(xy)[o] = make_double4(cc.x, cc.y, dd.x, dd.y);
              ++o;
            } else {
              ++cnt;
            }
          }
        }
        continue;
      }
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
      if (left) stack[sp++] = nd.child;
      if (right) stack[sp++] = nd.child + 1;
    }
  }
  if (!WRITE) counts[i] = cnt;
}

// geometry_utils.py:564-574 sort_intersections_by_edge on grouped rows: a stable insertion
// sort by t = (c - a) . (b - a) within each segment's row range; NaN sorts last (np.lexsort).
__device__ __forceinline__ bool t_less(double x, double y) {
  return x < y || (!isnan(x) && isnan(y));
}

__global__ void __launch_bounds__(kBlock) k_sort_by_edge(const double* __restrict__ segs, int64_t n,
                                                         const int64_t* __restrict__ offsets,
                                                         int64_t* __restrict__ jj,
                                                         double* __restrict__ xy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t lo = offsets[i], hi = offsets[i + 1];
  if (hi - lo < 2) return;
  const double4 s4 = reinterpret_cast<const double4*>(segs)[i];
  const double abx = s4.z - s4.x, aby = s4.w - s4.y;
  double4* rows = reinterpret_cast<double4*>(xy);
  auto tval = [&](const double4& r) { return (r.x - s4.x) * abx + (r.y - s4.y) * aby; };
  for (int64_t k = lo + 1; k < hi; ++k) {
    const double4 r = rows[k];
    const int64_t j = jj[k];
    const double t = tval(r);
    int64_t m = k;
    while (m > lo) {
      const double4 q = rows[m - 1];
      if (!t_less(t, tval(q))) break;
      rows[m] = q;
      jj[m] = jj[m - 1];
      --m;
    }
    rows[m] = r;
    jj[m] = j;
  }
}

// ----------------------------------------------------------------- barycentric
// barycentric_triangle.py:28-43 compute_weights.
__device__ __forceinline__ void tri_weights(P2 a, P2 b, P2 c, P2 p, double* w) {
  const double abx = b.x - a.x, aby = b.y - a.y;
  const double acx = c.x - a.x, acy = c.y - a.y;
  const double apx = p.x - a.x, apy = p.y - a.y;
  const double Aa = fabs(cross(abx, aby, apx, apy));
  const double Ac = fabs(cross(acx, acy, apx, apy));
  const double A = fabs(cross(abx, aby, acx, acy));
  const double inv_denom = 1.0 / A;
  const double ww = inv_denom * Aa;
  const double v = inv_denom * Ac;
  w[0] = 1.0 - v - ww;
  w[1] = v;
  w[2] = ww;
}

// barycentric_wachspress.py:26-85 compute_weights (w has nw columns, zeroed by the caller).
template <int MAXV>
__device__ __forceinline__ void wachspress(const P2 (&poly)[MAXV], int n, P2 p, double* w, int nw,
                                           double tol) {
  auto edge_case = [&](P2 a, double ux, double uy, int i, int j) {  // :26-35
    for (int k = 0; k < nw; ++k) w[k] = 0.0;
    const double vx = p.x - a.x, vy = p.y - a.y;
    const double f = sqrt(dot(vx, vy, vx, vy)) / sqrt(dot(ux, uy, ux, uy));
    w[i] = 1.0 - f;
    w[j] = f;
  };
  P2 a = poly[0];
#pragma unroll
  for (int k = 0; k < MAXV; ++k)
    if (k == n - 1) a = poly[k];
  P2 b = poly[0];
  double ux = b.x - a.x, uy = b.y - a.y;
  double Ai = fabs(cross(ux, uy, p.x - a.x, p.y - a.y));
  if (within_perp(Ai, ux, uy, tol)) {
    edge_case(a, ux, uy, n - 1, 0);
    return;
  }
  double w_sum = 0.0;
  double wl[MAXV];
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    if (i < n) {
      const int inext = (i + 1 == n) ? 0 : i + 1;
      P2 c = poly[0];
#pragma unroll
      for (int k = 0; k < MAXV; ++k)
        if (k == inext) c = poly[k];
      const double Ci = fabs(cross(ux, uy, c.x - a.x, c.y - a.y));
      ux = c.x - b.x;
      uy = c.y - b.y;
      const double Aj = fabs(cross(ux, uy, p.x - b.x, p.y - b.y));
      if (within_perp(Aj, ux, uy, tol)) {
        edge_case(b, ux, uy, i, inext);
        return;
      }
      const double wi = 2 * Ci / (Ai * Aj);
      wl[i] = wi;
      w_sum += wi;
      a = b;
      b = c;
      Ai = Aj;
    }
  }
#pragma unroll
  for (int i = 0; i < MAXV; ++i)
    if (i < n) w[i] = wl[i] / w_sum;
}

template <int MAXV>
__device__ __forceinline__ void weights_for(const double* __restrict__ vertices,
                                            const int32_t* __restrict__ faces, int nmax, int64_t f,
                                            P2 p, double tol, bool use_wachspress, double* w) {
  P2 poly[MAXV];
  const int len = load_poly<MAXV>(faces, nmax, f, vertices, poly);
  if (use_wachspress) {
    wachspress<MAXV>(poly, len, p, w, nmax, tol);
  } else {
    tri_weights(poly[0], poly[1], poly[2], p, w);
  }
}

template <int MAXV>
__global__ void __launch_bounds__(kBlock) k_bary(const double* __restrict__ pts,
                                                 const int64_t* __restrict__ face_indices, int64_t n,
                                                 const int32_t* __restrict__ faces, int nmax,
                                                 const double* __restrict__ vertices, double tol,
                                                 int wachs, int nw, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double w[CT_MAX_N_VERTEX];
  for (int k = 0; k < nw; ++k) w[k] = 0.0;
  const int64_t f = face_indices[i];
  if (f != -1) weights_for<MAXV>(vertices, faces, nmax, f, ld_point(pts, i), tol, wachs != 0, w);
  for (int k = 0; k < nw; ++k) out[i * nw + k] = w[k];
}

template <int MAXV, int STACK>
__global__ void __launch_bounds__(kBlock) k_locate_bary(TreeView t, const double* __restrict__ pts,
                                                        int64_t n, double tol,
                                                        int64_t* __restrict__ face_out,
                                                        double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P2 p = ld_point(pts, i);
  const int f = locate_one<MAXV, STACK, false>(t, p, tol);
  face_out[i] = f;
  const int nw = t.nmax;
  double w[MAXV];
#pragma unroll
  for (int k = 0; k < MAXV; ++k) w[k] = 0.0;
  if (f != -1) weights_for<MAXV>(t.vertices, t.elements, t.nmax, f, p, tol, nw > 3, w);
#pragma unroll
  for (int k = 0; k < MAXV; ++k)
    if (k < nw) out[i * nw + k] = w[k];
}

// ----------------------------------------------------------------- node bounds
// parent[] must be filled with -1 first.  Child indices are bounds-checked: the host
// array these nodes come from may have been edited by the caller.
__global__ void k_parents(const ct_node* __restrict__ nodes, int64_t n_nodes, int32_t* __restrict__ parent) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n_nodes) return;
  const int c = nodes[x].child;
  if (c >= 0 && c + 1 < n_nodes) {
    parent[c] = (int32_t)x;
    parent[c + 1] = (int32_t)x;
  }
}

// query.py:568-622 collect_node_bounds: a child copies its parent's bounds and replaces
// slot dim*2 + side (left: Lmax into the max slot, right: Rmin into the min slot).  Walking
// up from a node, the first ancestor edge that writes a slot decides it; the root bbox
// fills the rest.
__global__ void k_node_bounds(const ct_node* __restrict__ nodes, int64_t n_nodes,
                              const int32_t* __restrict__ parent, double b0, double b1, double b2,
                              double b3, double* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n_nodes) return;
  double v[4] = {b0, b1, b2, b3};
  bool set[4] = {false, false, false, false};
  int64_t y = x;
  for (int guard = 0; guard < 4096 && y > 0; ++guard) {
    const int p = parent[y];
    if (p < 0 || p >= n_nodes) break;
    const ct_node pn = nodes[p];
    const int side = (y == pn.child) ? 1 : 0;  // 1 = left, 0 = right
    const int slot = (pn.dim ? 2 : 0) + side;
    if (!set[slot]) {
      v[slot] = side ? pn.Lmax : pn.Rmin;
      set[slot] = true;
    }
    y = p;
  }
  reinterpret_cast<double4*>(out)[x] = make_double4(v[0], v[1], v[2], v[3]);
}

// query.py:625-664 validate_node_bounds.
__global__ void k_validate(TreeView t, int64_t n_nodes, const double* __restrict__ nb,
                           uint8_t* __restrict__ valid) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n_nodes) return;
  const ct_node nd = t.nodes[x];
  const Box me = ld_box(nb, x);
  bool ok = true;
  if (nd.child == -1) {
    if (nd.ptr < 0 || nd.size < 0 || (int64_t)nd.ptr + nd.size > t.n_elements) {
      ok = false;
    } else {
      for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
        const int c = t.bb_indices[k];
        ok = ok && c >= 0 && c < t.n_elements && box_contained(ld_box(t.bb, c), me);
      }
    }
  } else if (nd.child < 0 || (int64_t)nd.child + 1 >= n_nodes) {
    ok = false;
  } else {
    ok = box_contained(ld_box(nb, nd.child), me) && box_contained(ld_box(nb, nd.child + 1), me);
  }
  valid[x] = ok ? 1 : 0;
}

// ----------------------------------------------------------------- dispatch
template <int MAXV, bool ON_EDGE>
int launch_locate_v(const TreeView& t, int depth, const double* pts, int64_t n, double tol,
                    int64_t* out, hipStream_t st) {
  const unsigned g = grid_for(n);
  if (depth < 32) k_locate_points<MAXV, 32, ON_EDGE><<<g, kBlock, 0, st>>>(t, pts, n, tol, out);
  else if (depth < 64) k_locate_points<MAXV, 64, ON_EDGE><<<g, kBlock, 0, st>>>(t, pts, n, tol, out);
  else if (depth < 128) k_locate_points<MAXV, 128, ON_EDGE><<<g, kBlock, 0, st>>>(t, pts, n, tol, out);
  else k_locate_points<MAXV, 256, ON_EDGE><<<g, kBlock, 0, st>>>(t, pts, n, tol, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int check_tree(const ct_tree* tree) {
  CT_REQUIRE(tree != nullptr && tree->nodes && tree->bb_indices && tree->elements && tree->vertices,
             "incomplete tree");
  if (tree->depth > CT_MAX_DEPTH) {
    set_error("tree depth %d exceeds the supported %d", tree->depth, CT_MAX_DEPTH);
    return CT_ERR_DEPTH;
  }
  return CT_OK;
}

template <bool WRITE>
int launch_boxes(const ct_tree* tree, const double* boxes, int64_t n, int64_t* counts,
                 const int64_t* offsets, int64_t* ii, int64_t* jj, hipStream_t st) {
  const TreeView t = view(tree);
  const unsigned g = grid_for(n);
  const int d = tree->depth;
  if (d < 32) k_boxes<WRITE, 32><<<g, kBlock, 0, st>>>(t, boxes, n, counts, offsets, ii, jj);
  else if (d < 64) k_boxes<WRITE, 64><<<g, kBlock, 0, st>>>(t, boxes, n, counts, offsets, ii, jj);
  else if (d < 128) k_boxes<WRITE, 128><<<g, kBlock, 0, st>>>(t, boxes, n, counts, offsets, ii, jj);
  else k_boxes<WRITE, 256><<<g, kBlock, 0, st>>>(t, boxes, n, counts, offsets, ii, jj);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

template <bool WRITE, int MAXV, bool FACES>
int launch_edges_v(const TreeView& t, int d, const double* segs, int64_t n, int64_t* counts,
                   const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy, hipStream_t st) {
  const unsigned g = grid_for(n);
  if (d < 32) k_edges<WRITE, MAXV, 32, FACES><<<g, kBlock, 0, st>>>(t, segs, n, counts, offsets, ii, jj, xy);
  else if (d < 64) k_edges<WRITE, MAXV, 64, FACES><<<g, kBlock, 0, st>>>(t, segs, n, counts, offsets, ii, jj, xy);
  else if (d < 128) k_edges<WRITE, MAXV, 128, FACES><<<g, kBlock, 0, st>>>(t, segs, n, counts, offsets, ii, jj, xy);
  else k_edges<WRITE, MAXV, 256, FACES><<<g, kBlock, 0, st>>>(t, segs, n, counts, offsets, ii, jj, xy);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

template <bool WRITE>
int launch_edges(const ct_tree* tree, const double* segs, int64_t n, int kind, int64_t* counts,
                 const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy, hipStream_t st) {
  const TreeView t = view(tree);
  const int d = tree->depth, nm = tree->n_max_vert;
  if (kind == 0) return launch_edges_v<WRITE, 4, false>(t, d, segs, n, counts, offsets, ii, jj, xy, st);
  if (nm <= 4) return launch_edges_v<WRITE, 4, true>(t, d, segs, n, counts, offsets, ii, jj, xy, st);
  if (nm <= 8) return launch_edges_v<WRITE, 8, true>(t, d, segs, n, counts, offsets, ii, jj, xy, st);
  if (nm <= 16) return launch_edges_v<WRITE, 16, true>(t, d, segs, n, counts, offsets, ii, jj, xy, st);
  return launch_edges_v<WRITE, 32, true>(t, d, segs, n, counts, offsets, ii, jj, xy, st);
}

template <int MAXV>
int launch_locate_bary_v(const TreeView& t, int d, const double* pts, int64_t n, double tol,
                         int64_t* fo, double* w, hipStream_t st) {
  const unsigned g = grid_for(n);
  if (d < 32) k_locate_bary<MAXV, 32><<<g, kBlock, 0, st>>>(t, pts, n, tol, fo, w);
  else if (d < 64) k_locate_bary<MAXV, 64><<<g, kBlock, 0, st>>>(t, pts, n, tol, fo, w);
  else if (d < 128) k_locate_bary<MAXV, 128><<<g, kBlock, 0, st>>>(t, pts, n, tol, fo, w);
  else k_locate_bary<MAXV, 256><<<g, kBlock, 0, st>>>(t, pts, n, tol, fo, w);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_bary(const double* points, const int64_t* fi, int64_t n, const int32_t* faces, int nmax,
                const double* vertices, double tol, int wachs, double* out, hipStream_t st) {
  CT_REQUIRE(n >= 0 && nmax >= 3 && nmax <= CT_MAX_N_VERTEX, "bad sizes");
  if (n == 0) return CT_OK;
  const int nw = wachs ? nmax : 3;
  const unsigned g = grid_for(n);
  if (nmax <= 4) k_bary<4><<<g, kBlock, 0, st>>>(points, fi, n, faces, nmax, vertices, tol, wachs, nw, out);
  else if (nmax <= 8) k_bary<8><<<g, kBlock, 0, st>>>(points, fi, n, faces, nmax, vertices, tol, wachs, nw, out);
  else if (nmax <= 16) k_bary<16><<<g, kBlock, 0, st>>>(points, fi, n, faces, nmax, vertices, tol, wachs, nw, out);
  else k_bary<32><<<g, kBlock, 0, st>>>(points, fi, n, faces, nmax, vertices, tol, wachs, nw, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_locate_points(const ct_tree* tree, const double* points, int64_t n, double tol, int64_t* out,
                     void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  const TreeView t = view(tree);
  hipStream_t st = as_stream(stream);
  const int nm = tree->n_max_vert, d = tree->depth;
  if (nm <= 3) return launch_locate_v<3, false>(t, d, points, n, tol, out, st);
  if (nm <= 4) return launch_locate_v<4, false>(t, d, points, n, tol, out, st);
  if (nm <= 8) return launch_locate_v<8, false>(t, d, points, n, tol, out, st);
  if (nm <= 16) return launch_locate_v<16, false>(t, d, points, n, tol, out, st);
  return launch_locate_v<32, false>(t, d, points, n, tol, out, st);
}

int ct_locate_points_on_edge(const ct_tree* tree, const double* points, int64_t n, double tol,
                             int64_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  CT_REQUIRE(tree->n_max_vert == 2, "edge tree expected (n_max_vert == 2)");
  if (n == 0) return CT_OK;
  return launch_locate_v<4, true>(view(tree), tree->depth, points, n, tol, out, as_stream(stream));
}

int ct_locate_boxes_count(const ct_tree* tree, const double* boxes, int64_t n, int64_t* counts,
                          void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  return launch_boxes<false>(tree, boxes, n, counts, nullptr, nullptr, nullptr, as_stream(stream));
}

int ct_locate_boxes_write(const ct_tree* tree, const double* boxes, int64_t n, const int64_t* offsets,
                          int64_t* ii, int64_t* jj, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  return launch_boxes<true>(tree, boxes, n, nullptr, offsets, ii, jj, as_stream(stream));
}

int ct_locate_edges_count(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          int64_t* counts, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 || tree->n_max_vert == 2, "edge-edge queries need an edge tree");
  if (n == 0) return CT_OK;
  return launch_edges<false>(tree, segments, n, kind, counts, nullptr, nullptr, nullptr, nullptr,
                             as_stream(stream));
}

int ct_locate_edges_write(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy,
                          void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 || tree->n_max_vert == 2, "edge-edge queries need an edge tree");
  if (n == 0) return CT_OK;
  return launch_edges<true>(tree, segments, n, kind, nullptr, offsets, ii, jj, xy, as_stream(stream));
}

int ct_sort_intersections_by_edge(const double* segments, int64_t n, const int64_t* offsets,
                                  int64_t* jj, double* xy, void* stream) {
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  k_sort_by_edge<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(segments, n, offsets, jj, xy);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_barycentric_triangle_weights(const double* points, const int64_t* face_indices, int64_t n,
                                    const int32_t* faces, int32_t n_max_vert, const double* vertices,
                                    double tol, double* out, void* stream) {
  return launch_bary(points, face_indices, n, faces, n_max_vert, vertices, tol, 0, out,
                     as_stream(stream));
}

int ct_barycentric_wachspress_weights(const double* points, const int64_t* face_indices, int64_t n,
                                      const int32_t* faces, int32_t n_max_vert,
                                      const double* vertices, double tol, double* out,
                                      void* stream) {
  return launch_bary(points, face_indices, n, faces, n_max_vert, vertices, tol, 1, out,
                     as_stream(stream));
}

int ct_compute_barycentric_weights(const ct_tree* tree, const double* points, int64_t n, double tol,
                                   int64_t* face_out, double* weights, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && tree->n_max_vert >= 3, "bad arguments");
  if (n == 0) return CT_OK;
  const TreeView t = view(tree);
  hipStream_t st = as_stream(stream);
  const int nm = tree->n_max_vert, d = tree->depth;
  if (nm <= 3) return launch_locate_bary_v<3>(t, d, points, n, tol, face_out, weights, st);
  if (nm <= 4) return launch_locate_bary_v<4>(t, d, points, n, tol, face_out, weights, st);
  if (nm <= 8) return launch_locate_bary_v<8>(t, d, points, n, tol, face_out, weights, st);
  if (nm <= 16) return launch_locate_bary_v<16>(t, d, points, n, tol, face_out, weights, st);
  return launch_locate_bary_v<32>(t, d, points, n, tol, face_out, weights, st);
}

int ct_collect_node_bounds(const ct_node* nodes, int64_t n_nodes, const double* bbox,
                           int32_t* parent_scratch, double* out, void* stream) {
  CT_REQUIRE(n_nodes >= 1 && nodes && bbox && parent_scratch && out, "bad arguments");
  hipStream_t st = as_stream(stream);
  int rc = fill<int32_t>(parent_scratch, n_nodes, -1, st);
  if (rc) return rc;
  k_parents<<<grid_for(n_nodes), kBlock, 0, st>>>(nodes, n_nodes, parent_scratch);
  CT_LAUNCH_CHECK();
  k_node_bounds<<<grid_for(n_nodes), kBlock, 0, st>>>(nodes, n_nodes, parent_scratch, bbox[0], bbox[1],
                                                      bbox[2], bbox[3], out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_validate_node_bounds(const ct_tree* tree, const double* node_bounds, uint8_t* validity,
                            void* stream) {
  CT_REQUIRE(tree && tree->n_nodes >= 1, "bad arguments");
  k_validate<<<grid_for(tree->n_nodes), kBlock, 0, as_stream(stream)>>>(view(tree), tree->n_nodes,
                                                                         node_bounds, validity);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // extern "C"
