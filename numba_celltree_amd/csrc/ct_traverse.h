// ct_traverse.h -- per-query tree traversals shared by the single-tree kernels (ct_query.hip)
// and the many-tree persistent kernels (ct_batch.hip).
//
// One lane per query, depth first.  The reference pushes both viable children and pops the
// near one at once (query.py:93-105, 236-242, 447-453); here the near child is entered
// directly and only the far child is deferred on the stack, which visits nodes in exactly
// the same order with half the stack traffic.  At most one deferred node per tree level is
// ever pending, so a stack of STACK >= depth entries cannot overflow; STACK is chosen per
// tree from {16, 24, 32, 48, 64, 128, 256}.  Leaf predicates are the reference's, so
// first-hit answers and the order of variable-length results match it exactly.
#pragma once

#include "ct_internal.h"

namespace ct {
namespace {

struct TreeView {
  const double* vertices;
  const int32_t* elements;
  const double* bb;
  const ct_node* nodes;
  const int32_t* bb_indices;
  const double* leaf;  // packed leaf-ordered coordinates, or null
  const double* nbox;  // tight node boxes, node j at row j + 1, or null
  int nmax;
  int n_elements;
  Box bbox;
  const int32_t* grid;  // start-node grid (grid_n x grid_n over bbox), or null
  int grid_n;
  double gsx, gsy;      // grid_n / bbox extent, per axis
  const double* lbb;    // cell bboxes in bb_indices order, or null
};

__host__ __device__ __forceinline__ TreeView view(const ct_tree& t) {
  const int g = t.start_grid ? t.start_grid_n : 0;
  return TreeView{t.vertices, t.elements, t.bb_coords, t.nodes, t.bb_indices, t.leaf_coords, t.node_boxes,
                  t.n_max_vert, (int)t.n_elements, Box{t.bbox[0], t.bbox[1], t.bbox[2], t.bbox[3]},
                  t.start_grid, g, g / (t.bbox[1] - t.bbox[0]), g / (t.bbox[3] - t.bbox[2]), t.leaf_bb};
}

inline TreeView view(const ct_tree* t) { return view(*t); }

// Calls f.template run<S>() with the smallest stack class S >= depth.
template <typename F>
int with_stack(int depth, F&& f) {
  if (depth <= 16) return f.template run<16>();
  if (depth <= 24) return f.template run<24>();
  if (depth <= 32) return f.template run<32>();
  if (depth <= 48) return f.template run<48>();
  if (depth <= 64) return f.template run<64>();
  if (depth <= 128) return f.template run<128>();
  return f.template run<256>();
}

// Per-lane traversal stack of STACK entries.  The first kLdsStackSlots entries live in LDS,
// lane-interleaved (slot s of lane l at base[s * stride + l]: a wavefront's pushes and pops
// touch 64 consecutive dwords); deeper entries go to a private array in scratch memory,
// which point queries rarely reach.  A VGPR array indexed by a per-lane depth would compile
// to a select over every entry per push and pop and cost occupancy; a full LDS stack
// (STACK ints per lane) would cap the workgroups per CU.  With 4 LDS slots the traversal
// kernels keep 8 wavefronts per SIMD.
//
// The kernel owns both arrays and hands this view two pointers, which stay in registers, so
// the compiler keeps LDS accesses as ds_* and spill accesses as scratch_* instructions (a
// struct holding the array itself would live in scratch and turn both into flat accesses).
constexpr int kLdsStackSlots = 4;
constexpr int kStackStride = 256;  // lanes per workgroup of every traversal kernel
static_assert(kBlock == kStackStride, "single-tree traversal workgroups are kStackStride lanes");

template <int STACK>
struct LaneStack {
  static_assert(STACK > kLdsStackSlots, "stack classes start at 16");
  static constexpr int kSpill = STACK - kLdsStackSlots;
  int* base;             // this lane's slot 0 in the workgroup's LDS array
  volatile int* spill;   // kSpill private ints (volatile: scratch, not VGPRs)
  __device__ __forceinline__ void put(int i, int v) {
    if (i < kLdsStackSlots) base[i * kStackStride] = v;
    else spill[i - kLdsStackSlots] = v;
  }
  __device__ __forceinline__ int get(int i) {
    int v;
    if (i < kLdsStackSlots) v = base[i * kStackStride];
    else v = spill[i - kLdsStackSlots];
    return v;
  }
};

// Declares the LDS and private arrays of a traversal stack and its view `name`.
#define CT_LANE_STACK(STACK, name)                                \
  __shared__ int name##_lds[kLdsStackSlots * kStackStride];       \
  volatile int name##_spill[LaneStack<STACK>::kSpill];            \
  LaneStack<STACK> name { name##_lds + threadIdx.x, name##_spill }

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

__device__ __forceinline__ int64_t query_index(const int32_t* __restrict__ order, int64_t i) {
  return order ? (int64_t)order[i] : i;
}

// Processing order of a query batch: lane i answers query order[i] (i when `order` is null).
// `sorted`, when not null, holds the queries already in that order (sorted[i] =
// queries[order[i]]), so lane i reads its query contiguously instead of gathering it.
struct QueryOrder {
  const int32_t* order;
  const double* sorted;
};

// Row k of the packed leaf coordinates: the cell's polygon (polygon_length's rule,
// geometry_utils.py:72-79) in slots [0, len), slots [len, nmax) repeating vertex len - 1.
// Registers past nmax (MAXV > nmax) repeat slot nmax - 1 as well, so `poly` always holds a
// MAXV-vertex row in point_in_cell's padded form.
// EXACT: the caller knows nmax == MAXV (rows of compile-time width, no selects).
template <int MAXV, bool EXACT>
__device__ __forceinline__ void load_leaf_poly(const double* __restrict__ leaf, int nmax, int k,
                                               P2 (&poly)[MAXV]) {
  if constexpr (EXACT) {
    const double2* row = reinterpret_cast<const double2*>(leaf) + (int64_t)k * MAXV;
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      const double2 d = row[i];
      poly[i] = {d.x, d.y};
    }
    return;
  }
  const double2* row = reinterpret_cast<const double2*>(leaf) + (int64_t)k * nmax;
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    if (i < 3 || i < nmax) {
      const double2 d = row[i];
      poly[i] = {d.x, d.y};
    } else {
      poly[i] = poly[i - 1];
    }
  }
}

// ----------------------------------------------------------------- points
// query.py:62-107 locate_point (faces) / :120-164 locate_point_on_edge (edges).
//
// LEAF says where a leaf's cells come from: the faces and vertices (kLeafFaces), the packed
// rows (kLeafPacked, rows of width nmax <= MAXV), or packed rows known to be exactly MAXV
// wide (kLeafPackedExact).
enum : int { kLeafFaces = 0, kLeafPacked = 1, kLeafPackedExact = 2 };
//
// Loop shape: an inner loop walks internal nodes until the lane stands on a leaf (or its
// stack runs dry), and only then tests that leaf's cells.  Lanes of a wavefront reach
// leaves at different iterations; with one loop per node the costly cell test would run
// (masked) once for every iteration in which any lane sits on a leaf.  Here lanes that
// reach a leaf early wait in the inner loop's exit, so the wavefront runs the cell test
// for all of them together.  Each lane still visits its nodes in the reference's order.
template <int MAXV, int STACK, bool ON_EDGE, int LEAF>
__device__ __forceinline__ bool leaf_hit(const TreeView& t, P2 p, double tol, const NodeLd& nd, int* found) {
  for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
    if constexpr (ON_EDGE) {
      const int c = t.bb_indices[k];
      const int2 e = reinterpret_cast<const int2*>(t.elements)[c];
      if (point_on_edge(p, ld_vertex(t.vertices, e.x), ld_vertex(t.vertices, e.y), tol)) {
        *found = c;
        return true;
      }
    } else if constexpr (LEAF != kLeafFaces) {
      P2 poly[MAXV];
      load_leaf_poly<MAXV, LEAF == kLeafPackedExact>(t.leaf, t.nmax, k, poly);
      if (point_in_cell<MAXV>(p, poly, tol)) {
        *found = t.bb_indices[k];
        return true;
      }
    } else {
      const int c = t.bb_indices[k];
      P2 poly[MAXV];
      const int len = load_poly<MAXV>(t.elements, t.nmax, c, t.vertices, poly);
      if (point_in_poly_or_edge<MAXV>(p, poly, len, tol)) {
        *found = c;
        return true;
      }
    }
  }
  return false;
}

// ----------------------------------------------------------------- start-node grid
// An optional grid_n x grid_n grid over the tree bbox (ct_start_grid) holds, per grid cell,
// the deepest node that every point walk starting in that cell reaches with nothing
// deferred: from the root, while the cell's whole span along the split axis lies on one
// side of the split (only the left child viable, or only the right), the walk's choice is
// forced.  A walk that starts at that node visits exactly the nodes the walk from the root
// visits after it, in the same order, so its answer is the same.  The cell's span is
// recomputed here by the same arithmetic as in the grid build, and a point outside it (a
// rounding at a cell border, a NaN) starts at the root.
__host__ __device__ __forceinline__ void grid_span(double lo, double hi, int n, int g, double* a, double* b) {
  const double w = (hi - lo) / n;
  *a = g == 0 ? -__builtin_inf() : lo + g * w;
  *b = g == n - 1 ? __builtin_inf() : lo + (g + 1) * w;
}

__device__ __forceinline__ int grid_index(double v, double lo, double scale, int n) {
  double g = (v - lo) * scale;
  g = g > 0.0 ? (g < n - 1.0 ? g : n - 1.0) : 0.0;  // NaN -> 0
  return (int)g;
}

__device__ __forceinline__ int start_node(const TreeView& t, P2 p) {
  if (t.grid == nullptr) return 0;
  const int gx = grid_index(p.x, t.bbox.xmin, t.gsx, t.grid_n);
  const int gy = grid_index(p.y, t.bbox.ymin, t.gsy, t.grid_n);
  double x0, x1, y0, y1;
  grid_span(t.bbox.xmin, t.bbox.xmax, t.grid_n, gx, &x0, &x1);
  grid_span(t.bbox.ymin, t.bbox.ymax, t.grid_n, gy, &y0, &y1);
  const bool inside = p.x >= x0 && p.x <= x1 && p.y >= y0 && p.y <= y1;
  return inside ? t.grid[(int64_t)gy * t.grid_n + gx] : 0;
}

template <int MAXV, int STACK, bool ON_EDGE, int LEAF>
__device__ __forceinline__ int locate_one(const TreeView& t, P2 p, double tol, LaneStack<STACK>& stack) {
  int sp = 0;
  int node = start_node(t, p);
  while (true) {
    NodeLd nd = ld_node(t.nodes, node);
    bool exhausted = false;
    while (nd.child != -1) {
      const double v = nd.dim ? p.y : p.x;
      const bool left = v <= nd.Lmax;
      const bool right = v >= nd.Rmin;
      const int lc = nd.child, rc = lc + 1;
      if (left && right) {
        // Both viable: the side with the larger margin goes first (query.py:93-101).
        const bool right_first = (nd.Lmax - v) < (v - nd.Rmin);
        stack.put(sp++, right_first ? lc : rc);
        node = right_first ? rc : lc;
      } else if (left) {
        node = lc;
      } else if (right) {
        node = rc;
      } else {
        if (sp == 0) {
          exhausted = true;
          break;
        }
        node = stack.get(--sp);
      }
      nd = ld_node(t.nodes, node);
    }
    if (exhausted) return -1;
    int found;
    if (leaf_hit<MAXV, STACK, ON_EDGE, LEAF>(t, p, tol, nd, &found)) return found;
    if (sp == 0) return -1;
    node = stack.get(--sp);
  }
}

// ----------------------------------------------------------------- segments
// query.py:309-328 compute_edge_face_intersect: Cohen-Sutherland on the cell bbox, then
// Cyrus-Beck on the cell with tol = max(1e-15, 1e-12 * max(w, h)).
// `box` is cell c's bbox.
template <int MAXV>
__device__ __forceinline__ bool edge_face_box(const TreeView& t, int c, const Box& box, P2 a, P2 b,
                                              P2* cc, P2* dd) {
  if (!cohen_sutherland(a, b, box, cc, dd)) return false;
  P2 poly[MAXV];
  const int len = load_poly<MAXV>(t.elements, t.nmax, c, t.vertices, poly);
  const double tol = pmax(kMinTolerance, kToleranceFactor * pmax(box.xmax - box.xmin, box.ymax - box.ymin));
  return cyrus_beck<MAXV>(a, b, poly, len, tol, cc, dd);
}

template <int MAXV>
__device__ __forceinline__ bool edge_face(const TreeView& t, int c, P2 a, P2 b, P2* cc, P2* dd) {
  return edge_face_box<MAXV>(t, c, ld_box(t.bb, c), a, b, cc, dd);
}

// The bbox of the cell at leaf position k (cell c = bb_indices[k]): the leaf-ordered copy
// when the tree has one (a read that does not wait for c).
__device__ __forceinline__ Box leaf_box(const TreeView& t, int k, int c) {
  return t.lbb ? ld_box(t.lbb, k) : ld_box(t.bb, c);
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

// Subtree pruning for segment-face queries (no reference counterpart).  A face is a hit only
// if Cohen-Sutherland accepts the segment against the face's bbox, and every face bbox of a
// subtree lies inside the subtree's tight box.  Cohen-Sutherland's clipped points stay
// within a few ulps of the true segment (each clip interpolates inside [a, b]), so when
// the segment misses the subtree box grown by m = 1e-9 * (largest coordinate magnitude),
// no face below can be accepted: skipping the subtree changes no result, only the number
// of faces tested (O(length) instead of O(length^2) for long segments).  Tested by
// Liang-Barsky slab clipping of t in [0, 1]; a NaN anywhere makes the test accept.
__device__ __forceinline__ bool seg_meets_box(P2 a, double dx, double dy, double mseg, double4 nb) {
  const double mb = pmax(pmax(fabs(nb.x), fabs(nb.y)), pmax(fabs(nb.z), fabs(nb.w)));
  const double m = 1e-9 * pmax(mseg, mb);
  const double xlo = nb.x - m, xhi = nb.y + m, ylo = nb.z - m, yhi = nb.w + m;
  double t0 = 0.0, t1 = 1.0;
  if (dx == 0.0) {
    if (a.x < xlo || a.x > xhi) return false;
  } else {
    double ta = (xlo - a.x) / dx, tb = (xhi - a.x) / dx;
    if (dx < 0.0) {
      const double s = ta;
      ta = tb;
      tb = s;
    }
    t0 = ta > t0 ? ta : t0;
    t1 = tb < t1 ? tb : t1;
  }
  if (dy == 0.0) {
    if (a.y < ylo || a.y > yhi) return false;
  } else {
    double ta = (ylo - a.y) / dy, tb = (yhi - a.y) / dy;
    if (dy < 0.0) {
      const double s = ta;
      ta = tb;
      tb = s;
    }
    t0 = ta > t0 ? ta : t0;
    t1 = tb < t1 ? tb : t1;
  }
  return !(t0 > t1);
}

// Which children of internal node nd segment (a, b) (v = b - a) may reach: query.py:407-441's
// slab test on the split dimension (the t-tests always hold once the slab test passed, but
// are kept), then, when nbox is set (tight node boxes, the segment finite), seg_meets_box.
__device__ __forceinline__ void seg_children(const NodeLd& nd, P2 a, P2 b, double vx, double vy, double mseg,
                                             const double4* __restrict__ nbox, bool* left_out, bool* right_out) {
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
  if (nbox != nullptr && (left || right)) {
    // the two children's boxes are rows child + 1, child + 2: one aligned 64-byte pair
    if (left) left = seg_meets_box(a, vx, vy, mseg, nbox[nd.child]);
    if (right) right = seg_meets_box(a, vx, vy, mseg, nbox[nd.child + 1]);
  }
  *left_out = left;
  *right_out = right;
}

// Per-segment setup of the subtree pruning: the node-box rows (or null when the tree has none,
// the query is a segment-edge query, or the segment is not finite) and the segment's largest
// coordinate magnitude.
template <bool FACES>
__device__ __forceinline__ const double4* seg_prune_boxes(const TreeView& t, P2 a, P2 b, double vx, double vy,
                                                          double* mseg) {
  *mseg = pmax(pmax(fabs(a.x), fabs(a.y)), pmax(fabs(b.x), fabs(b.y)));
  const bool prune = FACES && t.nbox != nullptr && isfinite(*mseg) && isfinite(vx) && isfinite(vy);
  return prune ? reinterpret_cast<const double4*>(t.nbox) + 1 : nullptr;
}

// The first `slots` hit cells of one query, kept by a count pass so that the write pass can
// replay them instead of walking the tree again: hit c at cells[c * stride].
struct HitRecord {
  int32_t* cells = nullptr;
  int slots = 0;
  int64_t stride = 0;
};

// query.py:357-455 locate_edge for segment q = (a, b): Cohen-Sutherland against the tree
// bbox, then slab pruning per node (the t-tests of :428-441 always hold once the slab test
// passed, but are kept).  Returns the number of hits; with WRITE, hit rows go to o, o+1, ...;
// without, the first rec.slots hit cells go to rec (when rec.cells is set).
// Segment-face queries on a tree with node boxes also skip subtrees by seg_meets_box.
// `start` != 0: walk only the subtree of node `start`, which the caller reached by this same
// walk's tests (a subquery, see ct_split.hip); the tree-bbox test was passed then.
template <bool WRITE, int MAXV, int STACK, bool FACES>
__device__ __forceinline__ int64_t edge_query(const TreeView& t, P2 a, P2 b, int64_t q, int64_t o,
                                              int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                              double* __restrict__ xy, LaneStack<STACK>& stack,
                                              HitRecord rec = HitRecord{}, int start = 0) {
  int64_t cnt = 0;
  P2 c0, d0;
  if (start == 0 && !cohen_sutherland(a, b, t.bbox, &c0, &d0)) return 0;
  const double vx = b.x - a.x, vy = b.y - a.y;
  // Non-finite segments keep the reference traversal unchanged (no pruning).
  double mseg;
  const double4* __restrict__ nbox = seg_prune_boxes<FACES>(t, a, b, vx, vy, &mseg);
  // Same loop shape as locate_one: internal nodes in an inner loop, the leaf's cell tests
  // once every lane of the wavefront stands on a leaf or has finished.
  int sp = 0;
  int node = start;
  while (true) {
    NodeLd nd = ld_node(t.nodes, node);
    bool exhausted = false;
    while (nd.child != -1) {
      bool left, right;
      seg_children(nd, a, b, vx, vy, mseg, nbox, &left, &right);
      if (left && right) {
        stack.put(sp++, nd.child);
        node = nd.child + 1;
      } else if (left || right) {
        node = nd.child + (right ? 1 : 0);
      } else {
        if (sp == 0) {
          exhausted = true;
          break;
        }
        node = stack.get(--sp);
      }
      nd = ld_node(t.nodes, node);
    }
    if (exhausted) break;
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
        } else if (rec.cells != nullptr && cnt < rec.slots) {
          rec.cells[cnt * rec.stride] = c;
        }
        ++cnt;
      }
    }
    if (sp == 0) break;
    node = stack.get(--sp);
  }
  return cnt;
}

// geometry_utils.py:564-574 sort_intersections_by_edge on grouped rows, np.lexsort's order
// (stable, NaN last).  One lane per row: the row's rank inside its segment's range = rows
// with a smaller t plus earlier rows with an equal t; the row is written once, to its rank.
// t = (c - a) . (b - a) with c the clipped start point, x-term + y-term as the reference's
// np.sum over the last axis.
__device__ __forceinline__ bool t_less(double x, double y) {
  return x < y || (!isnan(x) && isnan(y));
}

__device__ __forceinline__ void sort_row(const double* __restrict__ segs, const int64_t* __restrict__ offsets,
                                         const int64_t* __restrict__ ii, const int64_t* __restrict__ jj_in,
                                         const double* __restrict__ xy_in, int64_t r,
                                         int64_t* __restrict__ jj_out, double* __restrict__ xy_out) {
  const int64_t s = ii[r];
  const int64_t lo = offsets[s], hi = offsets[s + 1];
  const double4 s4 = reinterpret_cast<const double4*>(segs)[s];
  const double abx = s4.z - s4.x, aby = s4.w - s4.y;
  const double4 mine = reinterpret_cast<const double4*>(xy_in)[r];
  const double t = (mine.x - s4.x) * abx + (mine.y - s4.y) * aby;
  int64_t rank = lo;
  for (int64_t k = lo; k < hi; ++k) {
    if (k == r) continue;
    const double2 c = reinterpret_cast<const double2*>(xy_in)[2 * k];
    const double tk = (c.x - s4.x) * abx + (c.y - s4.y) * aby;
    if (t_less(tk, t) || (k < r && !t_less(t, tk))) ++rank;
  }
  jj_out[rank] = jj_in[r];
  reinterpret_cast<double4*>(xy_out)[rank] = mine;
}

}  // namespace
}  // namespace ct
