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
// The second half of edge_face_box: the Cyrus-Beck clip of a cell whose bbox Cohen-Sutherland
// accepted.
template <int MAXV>
__device__ __forceinline__ bool edge_face_clip(const TreeView& t, int c, const Box& box, P2 a, P2 b, P2* cc,
                                               P2* dd) {
  P2 poly[MAXV];
  const int len = load_poly<MAXV>(t.elements, t.nmax, c, t.vertices, poly);
  const double tol = pmax(kMinTolerance, kToleranceFactor * pmax(box.xmax - box.xmin, box.ymax - box.ymin));
  return cyrus_beck<MAXV>(a, b, poly, len, tol, cc, dd);
}

// edge_face_clip for the cell at leaf position k of a tree with packed leaf rows: its vertices
// from the row (a read that does not wait for c; slots past the polygon's length repeat its
// last vertex and are never read by the clip), its length from its face row, and the cell box
// behind the tolerance as the min / max of the row's vertices -- the values the face bboxes
// hold, so the same tolerance bit for bit, without loading the box.
template <int MAXV>
__device__ __forceinline__ bool edge_face_clip_row(const TreeView& t, int k, int c, P2 a, P2 b, P2* cc, P2* dd) {
  P2 poly[MAXV];
  load_leaf_poly<MAXV, false>(t.leaf, t.nmax, k, poly);
  const int32_t* row = t.elements + (int64_t)c * t.nmax;
  int len = t.nmax < MAXV ? t.nmax : MAXV;
#pragma unroll
  for (int i = MAXV - 1; i >= 3; --i)
    if (i < t.nmax && row[i] == kFill) len = i;
  double xmin = poly[0].x, xmax = poly[0].x, ymin = poly[0].y, ymax = poly[0].y;
#pragma unroll
  for (int i = 1; i < MAXV; ++i) {
    xmin = pmin(xmin, poly[i].x);
    xmax = pmax(xmax, poly[i].x);
    ymin = pmin(ymin, poly[i].y);
    ymax = pmax(ymax, poly[i].y);
  }
  const double tol = pmax(kMinTolerance, kToleranceFactor * pmax(xmax - xmin, ymax - ymin));
  return cyrus_beck<MAXV>(a, b, poly, len, tol, cc, dd);
}

template <int MAXV>
__device__ __forceinline__ bool edge_face_box(const TreeView& t, int c, const Box& box, P2 a, P2 b,
                                              P2* cc, P2* dd) {
  if (!cohen_sutherland(a, b, box, cc, dd)) return false;
  return edge_face_clip<MAXV>(t, c, box, a, b, cc, dd);
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

// Per-segment setup of the node tests below.
struct SegPrune {
  const double4* nbox;  // node-box rows (node j's children at child + 1, child + 2), or null
  double mseg;          // the segment's largest coordinate magnitude
  double ix, iy;        // 1 / (b - a) per axis (used only where the component is not 0)
  bool finite_v;        // b - a is finite
};

// Subtree pruning for segment-face queries (no reference counterpart).  A face is a hit only
// if Cohen-Sutherland accepts the segment against the face's bbox, and every face bbox of a
// subtree lies inside the subtree's tight box.  Cohen-Sutherland's clipped points stay
// within a few ulps of the true segment (each clip interpolates inside [a, b]), so when
// the segment misses the subtree box grown by m = 1e-9 * (largest coordinate magnitude),
// no face below can be accepted: skipping the subtree changes no result, only the number
// of faces tested (O(length) instead of O(length^2) for long segments).  Tested by
// Liang-Barsky slab clipping of t in [0, 1], with t from the segment's reciprocal
// components (a relative error of an ulp or two in t moves the tested point by a few ulps of
// the coordinates, far inside m); a NaN anywhere makes the test accept.
__device__ __forceinline__ bool seg_meets_box(P2 a, double dx, double dy, const SegPrune& sp, double4 nb) {
  const double mb = pmax(pmax(fabs(nb.x), fabs(nb.y)), pmax(fabs(nb.z), fabs(nb.w)));
  const double m = 1e-9 * pmax(sp.mseg, mb);
  const double xlo = nb.x - m, xhi = nb.y + m, ylo = nb.z - m, yhi = nb.w + m;
  double t0 = 0.0, t1 = 1.0;
  if (dx == 0.0) {
    if (a.x < xlo || a.x > xhi) return false;
  } else {
    double ta = (xlo - a.x) * sp.ix, tb = (xhi - a.x) * sp.ix;
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
    double ta = (ylo - a.y) * sp.iy, tb = (yhi - a.y) * sp.iy;
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
// slab test on the split dimension, then its t-tests, then, when sp.nbox is set (tight node
// boxes, the segment finite), seg_meets_box.  The t-tests divide a distance of the tested sign
// by the segment's component along the split axis: for a finite, nonzero component the
// quotient always has the sign (or the bound) they test, so they can only fail when both are
// infinite, and are evaluated only for segments whose components are not finite.
__device__ __forceinline__ void seg_children(const NodeLd& nd, P2 a, P2 b, double vx, double vy, const SegPrune& sp,
                                             bool* left_out, bool* right_out) {
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
  if (!sp.finite_v) {
    if (dx > 0.0) {
      if (left) left = (dl / dx) >= 0.0;
      if (right) right = (dr / dx) <= 1.0;
    } else if (dx < 0.0) {
      if (left) left = (1.0 - (dl / dx)) >= 0.0;
      if (right) right = (1.0 - (dr / dx)) <= 1.0;
    }
  }
  if (sp.nbox != nullptr && (left || right)) {
    // the two children's boxes are rows child + 1, child + 2: one aligned 64-byte pair
    if (left) left = seg_meets_box(a, vx, vy, sp, sp.nbox[nd.child]);
    if (right) right = seg_meets_box(a, vx, vy, sp, sp.nbox[nd.child + 1]);
  }
  *left_out = left;
  *right_out = right;
}

// The node-test setup of segment (a, b): node-box rows only for segment-face queries on a tree
// that has them and a finite segment.
template <bool FACES>
__device__ __forceinline__ SegPrune seg_prune(const TreeView& t, P2 a, P2 b, double vx, double vy) {
  SegPrune sp;
  sp.mseg = pmax(pmax(fabs(a.x), fabs(a.y)), pmax(fabs(b.x), fabs(b.y)));
  sp.finite_v = isfinite(vx) && isfinite(vy);
  const bool prune = FACES && t.nbox != nullptr && isfinite(sp.mseg) && sp.finite_v;
  sp.nbox = prune ? reinterpret_cast<const double4*>(t.nbox) + 1 : nullptr;
  sp.ix = vx != 0.0 ? 1.0 / vx : 0.0;
  sp.iy = vy != 0.0 ? 1.0 / vy : 0.0;
  return sp;
}

// The first `slots` hit cells of one query, kept by a count pass so that the write pass can
// replay them instead of walking the tree again: hit h at cells[h * stride].  Segment queries
// may also keep the first `xy_slots` (<= slots) clipped segments, hit h at xy[h * stride] as
// (x0, y0, x1, y1), so that the write pass copies them instead of clipping again.
struct HitRecord {
  int32_t* cells = nullptr;
  int slots = 0;
  int64_t stride = 0;
  double4* xy = nullptr;
  int xy_slots = 0;
};

// geometry_utils.py:564-574's sort key of a row of segment s4 = (a, b) whose clipped start
// point is c: t = (c - a) . (b - a), x-term + y-term as the reference's np.sum over the last
// axis.
__device__ __forceinline__ double row_t(double4 s4, double2 c) {
  const double abx = s4.z - s4.x, aby = s4.w - s4.y;
  return (c.x - s4.x) * abx + (c.y - s4.y) * aby;
}

// Where a writing walk puts its rows: the output arrays (row r = (q, cell, clipped segment)) ...
struct FinalRows {
  int64_t* ii;
  int64_t* jj;
  double* xy;
  __device__ __forceinline__ void put(int64_t r, int64_t q, int c, P2 cc, P2 dd) const {
    ii[r] = q;
    jj[r] = c;
    reinterpret_cast<double4*>(xy)[r] = make_double4(cc.x, cc.y, dd.x, dd.y);
  }
};

// ... or staging rows in processing order (ct_split.hip), with each row's sort key t for
// segment `seg`.
struct StagedRows {
  int32_t* query;
  int32_t* cell;
  double4* xy;
  double* t;
  double4 seg;
  __device__ __forceinline__ void put(int64_t r, int64_t q, int c, P2 cc, P2 dd) const {
    query[r] = (int32_t)q;
    cell[r] = c;
    xy[r] = make_double4(cc.x, cc.y, dd.x, dd.y);
    t[r] = row_t(seg, make_double2(cc.x, cc.y));
  }
};

// query.py:357-455 locate_edge for segment q = (a, b): Cohen-Sutherland against the tree
// bbox, then slab pruning per node (the t-tests of :428-441 always hold once the slab test
// passed, but are kept).  Returns the number of hits; with WRITE, hit rows go to o, o+1, ...;
// without, the first rec.slots hit cells go to rec (when rec.cells is set).
// Segment-face queries on a tree with node boxes also skip subtrees by seg_meets_box.
// `start` != 0: walk only the subtree of node `start`, which the caller reached by this same
// walk's tests (a subquery, see ct_split.hip); the tree-bbox test was passed then.
// CAND (segment-face count pass only): a cell counts once Cohen-Sutherland accepts its bbox,
// and the record keeps its leaf position k; the Cyrus-Beck clip of these candidates runs
// afterwards, one candidate list per lane, without the walk's divergence around it.
template <bool WRITE, int MAXV, int STACK, bool FACES, bool CAND = false, class OUT = FinalRows>
__device__ __forceinline__ int64_t edge_query(const TreeView& t, P2 a, P2 b, int64_t q, int64_t o, const OUT& out,
                                              LaneStack<STACK>& stack, HitRecord rec = HitRecord{}, int start = 0) {
  int64_t cnt = 0;
  P2 c0, d0;
  if (start == 0 && !cohen_sutherland(a, b, t.bbox, &c0, &d0)) return 0;
  const double vx = b.x - a.x, vy = b.y - a.y;
  // Non-finite segments keep the reference traversal unchanged (no pruning).
  const SegPrune pr = seg_prune<FACES>(t, a, b, vx, vy);
  // Same loop shape as locate_one: internal nodes in an inner loop, the leaf's cell tests
  // once every lane of the wavefront stands on a leaf or has finished.
  int sp = 0;
  int node = start;
  while (true) {
    NodeLd nd = ld_node(t.nodes, node);
    bool exhausted = false;
    while (nd.child != -1) {
      bool left, right;
      seg_children(nd, a, b, vx, vy, pr, &left, &right);
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
      if constexpr (CAND) {
        static_assert(FACES && !WRITE, "candidate lists are for segment-face count passes");
        hit = cohen_sutherland(a, b, leaf_box(t, k, c), &cc, &dd);
        if (hit && rec.cells != nullptr && cnt < rec.slots) rec.cells[cnt * rec.stride] = k;
        cnt += hit ? 1 : 0;
        continue;
      } else if constexpr (FACES) {
        hit = edge_face_box<MAXV>(t, c, leaf_box(t, k, c), a, b, &cc, &dd);
      } else {
        hit = edge_edge(t, c, a, b, &cc, &dd);
      }
      if (hit) {
        if (WRITE) {
          out.put(o + cnt, q, c, cc, dd);
        } else if (rec.cells != nullptr && cnt < rec.slots) {
          rec.cells[cnt * rec.stride] = c;
          if (cnt < rec.xy_slots) rec.xy[cnt * rec.stride] = make_double4(cc.x, cc.y, dd.x, dd.y);
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
// with a smaller t plus earlier rows with an equal t (row_t); the row is written once, to its
// rank.
__device__ __forceinline__ bool t_less(double x, double y) {
  return x < y || (!isnan(x) && isnan(y));
}

// Rows are ranked a tile of kRankTile at a time by one 256-lane workgroup: the t of every
// row within kRankHalo rows of the tile is computed once into LDS, and a row whose segment's
// rows all lie in that window is ranked from LDS; a longer segment's row reads its
// neighbours' start points from global memory.  Rows before r (equal t counts) and after r
// (only smaller t counts) are two loops, so no lane tests k == r.
constexpr int kRankTile = 1024;
constexpr int kRankHalo = 256;
constexpr int kRankWindow = kRankTile + 2 * kRankHalo;

// Every lane of the workgroup calls this with the same r0 (it synchronises the workgroup);
// tw holds kRankWindow doubles of LDS.
__device__ __forceinline__ void sort_tile(const double* __restrict__ segs, const int64_t* __restrict__ offsets,
                                          const int64_t* __restrict__ ii, const int64_t* __restrict__ jj_in,
                                          const double* __restrict__ xy_in, int64_t n_rows, int64_t r0,
                                          int64_t* __restrict__ jj_out, double* __restrict__ xy_out, double* tw) {
  const double4* __restrict__ s4v = reinterpret_cast<const double4*>(segs);
  const double2* __restrict__ xy2 = reinterpret_cast<const double2*>(xy_in);
  const int64_t w0 = r0 > kRankHalo ? r0 - kRankHalo : 0;
  const int64_t w1 = r0 + kRankTile + kRankHalo < n_rows ? r0 + kRankTile + kRankHalo : n_rows;
  for (int64_t r = w0 + threadIdx.x; r < w1; r += blockDim.x) tw[r - w0] = row_t(s4v[ii[r]], xy2[2 * r]);
  __syncthreads();
  for (int k = threadIdx.x; k < kRankTile; k += blockDim.x) {
    const int64_t r = r0 + k;
    if (r >= n_rows) break;
    const int64_t s = ii[r];
    const int64_t lo = offsets[s], hi = offsets[s + 1];
    const double t = tw[r - w0];
    int64_t rank = lo;
    if (lo >= w0 && hi <= w1) {
      for (int64_t m = lo; m < r; ++m) rank += t_less(t, tw[m - w0]) ? 0 : 1;
      for (int64_t m = r + 1; m < hi; ++m) rank += t_less(tw[m - w0], t) ? 1 : 0;
    } else {
      const double4 s4 = s4v[s];
      for (int64_t m = lo; m < r; ++m) rank += t_less(t, row_t(s4, xy2[2 * m])) ? 0 : 1;
      for (int64_t m = r + 1; m < hi; ++m) rank += t_less(row_t(s4, xy2[2 * m]), t) ? 1 : 0;
    }
    jj_out[rank] = jj_in[r];
    reinterpret_cast<double4*>(xy_out)[rank] = reinterpret_cast<const double4*>(xy_in)[r];
  }
  __syncthreads();  // tw is refilled for the workgroup's next tile
}

}  // namespace
}  // namespace ct
