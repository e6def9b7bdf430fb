// ct_query.hip -- tree traversals on the device.
//
// One lane per query, depth first.  The reference pushes both viable children and pops the
// near one at once (query.py:93-105, 236-242, 447-453); here the near child is entered
// directly and only the far child is deferred on the stack, which visits nodes in exactly
// the same order with half the stack traffic.  At most one deferred node per tree level is
// ever pending, so a stack of STACK >= depth entries cannot overflow; STACK is chosen per
// tree from {16, 24, 32, 48, 64, 128, 256}.  Leaf predicates are the reference's, so
// first-hit answers and the order of variable-length results match it exactly.
//
// Variable-length queries run twice: a count pass, a device exclusive scan
// (ct_exclusive_scan), and a write pass at the scanned offsets -- no atomic append; output
// order = query order, then DFS order.  Queries may be processed in any lane order
// (`order`, from ct_query_order): every query writes only its own count / rows.
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
  const double* leaf;  // packed leaf-ordered coordinates, or null
  int nmax;
  int n_elements;
  Box bbox;
};

TreeView view(const ct_tree* t) {
  return TreeView{t->vertices, t->elements, t->bb_coords, t->nodes, t->bb_indices, t->leaf_coords,
                  t->n_max_vert, (int)t->n_elements,
                  Box{t->bbox[0], t->bbox[1], t->bbox[2], t->bbox[3]}};
}

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

// Row k of the packed leaf coordinates: the cell's vertices, unused slots NaN.  The length
// rule is polygon_length's (geometry_utils.py:72-79): the first empty slot at index >= 3.
template <int MAXV>
__device__ __forceinline__ int load_leaf_poly(const double* __restrict__ leaf, int nmax, int k,
                                              P2 (&poly)[MAXV]) {
  const double2* row = reinterpret_cast<const double2*>(leaf) + (int64_t)k * nmax;
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    if (i < nmax) {
      const double2 d = row[i];
      poly[i] = {d.x, d.y};
    } else {
      poly[i] = {0.0, 0.0};
    }
  }
  int len = nmax < MAXV ? nmax : MAXV;
#pragma unroll
  for (int i = MAXV - 1; i >= 3; --i)
    if (i < nmax && isnan(poly[i].x)) len = i;
  return len;
}

// ----------------------------------------------------------------- points
// query.py:62-107 locate_point (faces) / :120-164 locate_point_on_edge (edges).
template <int MAXV, int STACK, bool ON_EDGE, bool PACKED>
__device__ __forceinline__ int locate_one(const TreeView& t, P2 p, double tol) {
  int stack[STACK];
  int sp = 0;
  int node = 0;
  while (true) {
    const NodeLd nd = ld_node(t.nodes, node);
    if (nd.child == -1) {
      for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
        if constexpr (ON_EDGE) {
          const int c = t.bb_indices[k];
          const int2 e = reinterpret_cast<const int2*>(t.elements)[c];
          if (point_on_edge(p, ld_vertex(t.vertices, e.x), ld_vertex(t.vertices, e.y), tol)) return c;
        } else if constexpr (PACKED) {
          P2 poly[MAXV];
          const int len = load_leaf_poly<MAXV>(t.leaf, t.nmax, k, poly);
          if (point_in_poly_or_edge<MAXV>(p, poly, len, tol)) return t.bb_indices[k];
        } else {
          const int c = t.bb_indices[k];
          P2 poly[MAXV];
          const int len = load_poly<MAXV>(t.elements, t.nmax, c, t.vertices, poly);
          if (point_in_poly_or_edge<MAXV>(p, poly, len, tol)) return c;
        }
      }
      if (sp == 0) return -1;
      node = stack[--sp];
      continue;
    }
    const double v = nd.dim ? p.y : p.x;
    const bool left = v <= nd.Lmax;
    const bool right = v >= nd.Rmin;
    const int lc = nd.child, rc = lc + 1;
    if (left && right) {
      // Both viable: the side with the larger margin goes first (query.py:93-101).
      const bool right_first = (nd.Lmax - v) < (v - nd.Rmin);
      stack[sp++] = right_first ? lc : rc;
      node = right_first ? rc : lc;
    } else if (left) {
      node = lc;
    } else if (right) {
      node = rc;
    } else {
      if (sp == 0) return -1;
      node = stack[--sp];
    }
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8).
// With ordered queries, workgroup b takes tile (b % 8) * (grid / 8) + b / 8, so every XCD
// sweeps one contiguous range of the order -- one region of the tree -- through its own L2.
// The grid is a multiple of 8, which makes the map a bijection.
constexpr unsigned kXcds = 8;

__device__ __forceinline__ int64_t xcd_tile(unsigned b, unsigned grid) {
  return (int64_t)(b % kXcds) * (grid / kXcds) + b / kXcds;
}

inline unsigned grid_xcd(int64_t n) {
  const unsigned g = grid_for(n);
  return (g + kXcds - 1) / kXcds * kXcds;
}

// One lane per query; lane i answers query order[i] (or i).
template <int MAXV, int STACK, bool ON_EDGE, bool PACKED>
__global__ void __launch_bounds__(kBlock) k_locate_points(TreeView t, const double* __restrict__ pts,
                                                          const int32_t* __restrict__ order,
                                                          int64_t n, double tol,
                                                          int64_t* __restrict__ out) {
  const int64_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  out[q] = locate_one<MAXV, STACK, ON_EDGE, PACKED>(t, ld_point(pts, q), tol);
}

// ----------------------------------------------------------------- query ordering
// Sort of the queries by a 20-bit key.  Points: the Morton key of a 1024 x 1024 grid over
// the tree bbox.  Boxes and segments: a 4-bit work class (log2 of the query's bbox area
// relative to the tree bbox; equal work per lane within a wavefront) above the Morton key
// of a 256 x 256 grid of the query's centre.  Coordinates outside the tree bbox clamp to
// the border.  The order among equal keys is whatever the LDS atomics give: results do not
// depend on the processing order.
__device__ __forceinline__ unsigned spread_bits(unsigned v) {  // 10 bits -> even positions
  v &= 0x3ffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__device__ __forceinline__ unsigned grid_coord(double x, double lo, double scale, double cells) {
  double g = (x - lo) * scale * cells;
  g = g > 0.0 ? (g < cells - 1.0 ? g : cells - 1.0) : 0.0;  // NaN -> 0
  return (unsigned)g;
}

struct OrderGeom {
  Box bb;
  double sx, sy;     // 1 / extent
  double inv_area;   // 1 / (extent_x * extent_y)
};

__device__ __forceinline__ unsigned order_key(int kind, const double* __restrict__ q, int64_t i,
                                              const OrderGeom& g) {
  if (kind == 0) {
    const double2 p = reinterpret_cast<const double2*>(q)[i];
    return spread_bits(grid_coord(p.x, g.bb.xmin, g.sx, 1024.0)) |
           (spread_bits(grid_coord(p.y, g.bb.ymin, g.sy, 1024.0)) << 1);
  }
  const double4 d = reinterpret_cast<const double4*>(q)[i];
  double cx, cy, w, h;
  if (kind == 1) {  // (xmin, xmax, ymin, ymax)
    cx = 0.5 * (d.x + d.y);
    cy = 0.5 * (d.z + d.w);
    w = d.y - d.x;
    h = d.w - d.z;
  } else {  // ((x0, y0), (x1, y1))
    cx = 0.5 * (d.x + d.z);
    cy = 0.5 * (d.y + d.w);
    w = fabs(d.z - d.x);
    h = fabs(d.w - d.y);
  }
  const double rel = (w * h) * g.inv_area;  // fraction of the tree bbox covered
  int cls;                                   // 2^-24 .. 2^-9 -> 0 .. 15
  if (!(rel > 0.0)) {
    cls = (rel == rel) ? 0 : 15;  // empty -> 0, NaN -> 15
  } else if (rel >= 1.0) {
    cls = 15;
  } else {
    cls = ilogb(rel) + 24;
    cls = cls < 0 ? 0 : (cls > 15 ? 15 : cls);
  }
  const unsigned m = spread_bits(grid_coord(cx, g.bb.xmin, g.sx, 256.0)) |
                     (spread_bits(grid_coord(cy, g.bb.ymin, g.sy, 256.0)) << 1);
  return ((unsigned)cls << 16) | (m & 0xffffu);
}

// Two-level MSD bucket sort on the 20-bit key, without global atomics:
//   k_sort_hist     per tile of kSortTile queries: LDS histogram of the top 8 key bits,
//                   written bin-major to T[bin][tile]
//   scan of T       -> the start of every (bin, tile) run
//   k_sort_scatter  per tile: LDS cursors, (key, index) written to the coarse runs
//   k_sort_fine     one workgroup per coarse bin: LDS histogram / scan / scatter of the
//                   low 12 key bits -> the final order
constexpr int kSortItems = 32;
constexpr int kSortBatch = 8;  // loads in flight per lane
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kCoarse = 256;  // == kBlock
constexpr int kFineBits = 12;
constexpr int kFine = 1 << kFineBits;
constexpr int kFineBlock = 1024;

__global__ void __launch_bounds__(kBlock) k_sort_hist(int kind, const double* __restrict__ q, int64_t n,
                                                      OrderGeom g, int64_t ntiles,
                                                      int64_t* __restrict__ T) {
  __shared__ int hist[kCoarse];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int j0 = 0; j0 < kSortItems; j0 += kSortBatch) {
    unsigned k[kSortBatch];
#pragma unroll
    for (int j = 0; j < kSortBatch; ++j) {
      const int64_t i = base + (int64_t)(j0 + j) * kBlock + threadIdx.x;
      k[j] = (i < n) ? order_key(kind, q, i, g) : ~0u;
    }
#pragma unroll
    for (int j = 0; j < kSortBatch; ++j)
      if (k[j] != ~0u) atomicAdd(&hist[k[j] >> kFineBits], 1);
  }
  __syncthreads();
  T[(int64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_sort_scatter(int kind, const double* __restrict__ q, int64_t n,
                                                         OrderGeom g, int64_t ntiles,
                                                         const int64_t* __restrict__ Ts,
                                                         unsigned* __restrict__ keys,
                                                         int32_t* __restrict__ idx) {
  __shared__ int cur[kCoarse];
  cur[threadIdx.x] = (int)Ts[(int64_t)threadIdx.x * ntiles + blockIdx.x];
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int j0 = 0; j0 < kSortItems; j0 += kSortBatch) {
    unsigned k[kSortBatch];
#pragma unroll
    for (int j = 0; j < kSortBatch; ++j) {
      const int64_t i = base + (int64_t)(j0 + j) * kBlock + threadIdx.x;
      k[j] = (i < n) ? order_key(kind, q, i, g) : ~0u;
    }
#pragma unroll
    for (int j = 0; j < kSortBatch; ++j) {
      if (k[j] != ~0u) {
        const int pos = atomicAdd(&cur[k[j] >> kFineBits], 1);
        keys[pos] = k[j];
        idx[pos] = (int32_t)(base + (int64_t)(j0 + j) * kBlock + threadIdx.x);
      }
    }
  }
}

__global__ void __launch_bounds__(kFineBlock) k_sort_fine(int64_t ntiles, const int64_t* __restrict__ Ts,
                                                          int64_t n, const unsigned* __restrict__ keys,
                                                          const int32_t* __restrict__ idx,
                                                          int32_t* __restrict__ order) {
  __shared__ int hist[kFine];
  __shared__ int lds[kFineBlock / 64];
  const int cb = blockIdx.x;
  const int lo = (int)Ts[(int64_t)cb * ntiles];
  const int hi = (cb + 1 < kCoarse) ? (int)Ts[(int64_t)(cb + 1) * ntiles] : (int)n;
  for (int b = threadIdx.x; b < kFine; b += kFineBlock) hist[b] = 0;
  __syncthreads();
  constexpr int B = 4;
  for (int p0 = lo + threadIdx.x; p0 < hi; p0 += B * kFineBlock) {
    unsigned k[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int p = p0 + j * kFineBlock;
      k[j] = (p < hi) ? keys[p] : ~0u;
    }
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (k[j] != ~0u) atomicAdd(&hist[k[j] & (kFine - 1)], 1);
  }
  __syncthreads();
  constexpr int per = kFine / kFineBlock;
  int s = 0;
#pragma unroll
  for (int k = 0; k < per; ++k) s += hist[threadIdx.x * per + k];
  int total;
  int run = lo + block_exclusive_scan<int, kFineBlock>(s, lds, &total);
#pragma unroll
  for (int k = 0; k < per; ++k) {
    const int c = hist[threadIdx.x * per + k];
    hist[threadIdx.x * per + k] = run;
    run += c;
  }
  __syncthreads();
  for (int p0 = lo + threadIdx.x; p0 < hi; p0 += B * kFineBlock) {
    unsigned k[B];
    int v[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int p = p0 + j * kFineBlock;
      k[j] = (p < hi) ? keys[p] : ~0u;
      v[j] = (p < hi) ? idx[p] : 0;
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
      if (k[j] != ~0u) {
        const int pos = atomicAdd(&hist[k[j] & (kFine - 1)], 1);
        order[pos] = v[j];
      }
    }
  }
}

struct OrderWs {
  int64_t* T;
  int64_t* Ts;
  void* scan_ws;
  size_t scan_bytes;
  unsigned* keys;
  int32_t* idx;
  int32_t* order;  // only used by ct_locate_points_ordered
  int64_t ntiles;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

OrderWs carve_order(char* base, int64_t n, bool with_order, size_t* total) {
  size_t off = 0;
  OrderWs w{};
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off = align256(off + bytes);
    return p;
  };
  const int64_t nn = n > 0 ? n : 1;
  w.ntiles = (nn + kSortTile - 1) / kSortTile;
  const int64_t nt = (int64_t)kCoarse * w.ntiles;
  w.T = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * nt));
  w.Ts = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * (nt + 1)));
  w.scan_bytes = ct_scan_workspace_bytes(nt);
  w.scan_ws = take(w.scan_bytes);
  w.keys = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * nn));
  w.idx = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * nn));
  w.order = with_order ? reinterpret_cast<int32_t*>(take(sizeof(int32_t) * nn)) : nullptr;
  *total = off + 256;
  return w;
}

int query_order(int kind, const double* q, int64_t n, const double* bbox, const OrderWs& w,
                int32_t* order, hipStream_t st) {
  OrderGeom g;
  g.bb = Box{bbox[0], bbox[1], bbox[2], bbox[3]};
  const double ex = g.bb.xmax - g.bb.xmin, ey = g.bb.ymax - g.bb.ymin;
  g.sx = 1.0 / ex;
  g.sy = 1.0 / ey;
  g.inv_area = 1.0 / (ex * ey);
  const unsigned ntiles = (unsigned)w.ntiles;
  k_sort_hist<<<ntiles, kBlock, 0, st>>>(kind, q, n, g, w.ntiles, w.T);
  CT_LAUNCH_CHECK();
  int rc = exclusive_scan(w.T, (int64_t)kCoarse * w.ntiles, w.Ts, w.scan_ws, w.scan_bytes, st);
  if (rc) return rc;
  k_sort_scatter<<<ntiles, kBlock, 0, st>>>(kind, q, n, g, w.ntiles, w.Ts, w.keys, w.idx);
  CT_LAUNCH_CHECK();
  k_sort_fine<<<kCoarse, kFineBlock, 0, st>>>(w.ntiles, w.Ts, n, w.keys, w.idx, order);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

// Packed leaf coordinates (see ct_pack_leaf_coords).
__global__ void __launch_bounds__(kBlock) k_pack_leaf(TreeView t, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.n_elements) return;
  const int c = t.bb_indices[k];
  const int32_t* row = t.elements + (int64_t)c * t.nmax;
  const int len = poly_len(row, t.nmax);
  double2* dst = reinterpret_cast<double2*>(out) + k * t.nmax;
  const double nan = __builtin_nan("");
  for (int i = 0; i < t.nmax; ++i) {
    if (i < len) {
      const P2 v = ld_vertex(t.vertices, row[i]);
      dst[i] = make_double2(v.x, v.y);
    } else {
      dst[i] = make_double2(nan, nan);
    }
  }
}

// ----------------------------------------------------------------- boxes
// query.py:177-244 locate_box: strict overlap with the tree bbox first; the left child is
// pushed before the right, so the right child is visited first.
template <bool WRITE, int STACK>
__global__ void __launch_bounds__(kBlock) k_boxes(TreeView t, const double* __restrict__ boxes,
                                                  const int32_t* __restrict__ order, int64_t n,
                                                  int64_t* __restrict__ counts,
                                                  const int64_t* __restrict__ offsets,
                                                  int64_t* __restrict__ ii, int64_t* __restrict__ jj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  const Box box = ld_box(boxes, q);
  int64_t cnt = 0;
  int64_t o = WRITE ? offsets[q] : 0;
  if (boxes_intersect(box, t.bbox)) {
    int stack[STACK];
    int sp = 0;
    int node = 0;
    while (true) {
      const NodeLd nd = ld_node(t.nodes, node);
      bool pop = true;
      if (nd.child == -1) {
        for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
          const int c = t.bb_indices[k];
          if (boxes_intersect(box, ld_box(t.bb, c))) {
            if (WRITE) {
              ii[o] = q;
              jj[o] = c;
              ++o;
            } else {
              ++cnt;
            }
          }
        }
      } else {
        const double bmin = nd.dim ? box.ymin : box.xmin;
        const double bmax = nd.dim ? box.ymax : box.xmax;
        const bool left = bmin <= nd.Lmax;
        const bool right = bmax >= nd.Rmin;
        if (left && right) {
          stack[sp++] = nd.child;
          node = nd.child + 1;
          pop = false;
        } else if (left || right) {
          node = nd.child + (right ? 1 : 0);
          pop = false;
        }
      }
      if (pop) {
        if (sp == 0) break;
        node = stack[--sp];
      }
    }
  }
  if (!WRITE) counts[q] = cnt;
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
                                                  const int32_t* __restrict__ order, int64_t n,
                                                  int64_t* __restrict__ counts,
                                                  const int64_t* __restrict__ offsets,
                                                  int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                                  double* __restrict__ xy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  const double4 s4 = reinterpret_cast<const double4*>(segs)[q];
  const P2 a = {s4.x, s4.y}, b = {s4.z, s4.w};
  int64_t cnt = 0;
  int64_t o = WRITE ? offsets[q] : 0;
  P2 c0, d0;
  if (cohen_sutherland(a, b, t.bbox, &c0, &d0)) {
    const double vx = b.x - a.x, vy = b.y - a.y;
    int stack[STACK];
    int sp = 0;
    int node = 0;
    while (true) {
      const NodeLd nd = ld_node(t.nodes, node);
      bool pop = true;
      if (nd.child == -1) {
        for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
          const int c = t.bb_indices[k];
          P2 cc, dd;
          const bool hit = FACES ? edge_face<MAXV>(t, c, a, b, &cc, &dd) : edge_edge(t, c, a, b, &cc, &dd);
          if (hit) {
            if (WRITE) {
              ii[o] = q;
              jj[o] = c;
              reinterpret_cast<double4*>(xy)[o] = make_double4(cc.x, cc.y, dd.x, dd.y);
              ++o;
            } else {
              ++cnt;
            }
          }
        }
      } else {
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
        if (left && right) {
          stack[sp++] = nd.child;
          node = nd.child + 1;
          pop = false;
        } else if (left || right) {
          node = nd.child + (right ? 1 : 0);
          pop = false;
        }
      }
      if (pop) {
        if (sp == 0) break;
        node = stack[--sp];
      }
    }
  }
  if (!WRITE) counts[q] = cnt;
}

// geometry_utils.py:564-574 sort_intersections_by_edge on grouped rows, np.lexsort's order
// (stable, NaN last).  One lane per row: the row's rank inside its segment's range = rows
// with a smaller t plus earlier rows with an equal t; the row is written once, to its rank.
// t = (c - a) . (b - a) with c the clipped start point, x-term + y-term as the reference's
// np.sum over the last axis.
__device__ __forceinline__ bool t_less(double x, double y) {
  return x < y || (!isnan(x) && isnan(y));
}

__global__ void __launch_bounds__(kBlock) k_sort_by_edge(const double* __restrict__ segs,
                                                         const int64_t* __restrict__ offsets,
                                                         const int64_t* __restrict__ ii,
                                                         const int64_t* __restrict__ jj_in,
                                                         const double* __restrict__ xy_in,
                                                         int64_t n_rows, int64_t* __restrict__ jj_out,
                                                         double* __restrict__ xy_out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
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

// barycentric_wachspress.py:26-85 compute_weights (w has nw columns).
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

template <int MAXV, int STACK, bool PACKED>
__global__ void __launch_bounds__(kBlock) k_locate_bary(TreeView t, const double* __restrict__ pts,
                                                        int64_t n, double tol,
                                                        int64_t* __restrict__ face_out,
                                                        double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P2 p = ld_point(pts, i);
  const int f = locate_one<MAXV, STACK, false, PACKED>(t, p, tol);
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
int check_tree(const ct_tree* tree) {
  CT_REQUIRE(tree != nullptr && tree->nodes && tree->bb_indices && tree->elements && tree->vertices,
             "incomplete tree");
  if (tree->depth > CT_MAX_DEPTH) {
    set_error("tree depth %d exceeds the supported %d", tree->depth, CT_MAX_DEPTH);
    return CT_ERR_DEPTH;
  }
  return CT_OK;
}

template <int MAXV, bool ON_EDGE, bool PACKED>
struct LocateLaunch {
  const TreeView& t;
  const double* pts;
  const int32_t* order;
  int64_t n;
  double tol;
  int64_t* out;
  hipStream_t st;
  template <int S>
  int run() {
    k_locate_points<MAXV, S, ON_EDGE, PACKED><<<grid_xcd(n), kBlock, 0, st>>>(t, pts, order, n, tol, out);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <int MAXV, bool ON_EDGE, bool PACKED>
int launch_locate_v(const TreeView& t, int depth, const double* pts, const int32_t* order, int64_t n,
                    double tol, int64_t* out, hipStream_t st) {
  return with_stack(depth, LocateLaunch<MAXV, ON_EDGE, PACKED>{t, pts, order, n, tol, out, st});
}

// Face trees: picks the vertex-count class, and the packed leaf path when present.
int launch_locate_faces(const ct_tree* tree, const double* pts, const int32_t* order, int64_t n,
                        double tol, int64_t* out, hipStream_t st) {
  const TreeView t = view(tree);
  const int nm = tree->n_max_vert, d = tree->depth;
  if (t.leaf != nullptr && nm <= 8) {
    if (nm <= 3) return launch_locate_v<3, false, true>(t, d, pts, order, n, tol, out, st);
    if (nm <= 4) return launch_locate_v<4, false, true>(t, d, pts, order, n, tol, out, st);
    return launch_locate_v<8, false, true>(t, d, pts, order, n, tol, out, st);
  }
  if (nm <= 3) return launch_locate_v<3, false, false>(t, d, pts, order, n, tol, out, st);
  if (nm <= 4) return launch_locate_v<4, false, false>(t, d, pts, order, n, tol, out, st);
  if (nm <= 8) return launch_locate_v<8, false, false>(t, d, pts, order, n, tol, out, st);
  if (nm <= 16) return launch_locate_v<16, false, false>(t, d, pts, order, n, tol, out, st);
  return launch_locate_v<32, false, false>(t, d, pts, order, n, tol, out, st);
}

template <bool WRITE>
struct BoxesLaunch {
  const TreeView& t;
  const double* boxes;
  const int32_t* order;
  int64_t n;
  int64_t* counts;
  const int64_t* offsets;
  int64_t* ii;
  int64_t* jj;
  hipStream_t st;
  template <int S>
  int run() {
    k_boxes<WRITE, S><<<grid_for(n), kBlock, 0, st>>>(t, boxes, order, n, counts, offsets, ii, jj);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <bool WRITE, int MAXV, bool FACES>
struct EdgesLaunch {
  const TreeView& t;
  const double* segs;
  const int32_t* order;
  int64_t n;
  int64_t* counts;
  const int64_t* offsets;
  int64_t* ii;
  int64_t* jj;
  double* xy;
  hipStream_t st;
  template <int S>
  int run() {
    k_edges<WRITE, MAXV, S, FACES><<<grid_for(n), kBlock, 0, st>>>(t, segs, order, n, counts, offsets, ii, jj, xy);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <bool WRITE>
int launch_edges(const ct_tree* tree, const double* segs, const int32_t* order, int64_t n, int kind,
                 int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy,
                 hipStream_t st) {
  const TreeView t = view(tree);
  const int d = tree->depth, nm = tree->n_max_vert;
  if (kind == 0)
    return with_stack(d, EdgesLaunch<WRITE, 4, false>{t, segs, order, n, counts, offsets, ii, jj, xy, st});
  if (nm <= 4)
    return with_stack(d, EdgesLaunch<WRITE, 4, true>{t, segs, order, n, counts, offsets, ii, jj, xy, st});
  if (nm <= 8)
    return with_stack(d, EdgesLaunch<WRITE, 8, true>{t, segs, order, n, counts, offsets, ii, jj, xy, st});
  if (nm <= 16)
    return with_stack(d, EdgesLaunch<WRITE, 16, true>{t, segs, order, n, counts, offsets, ii, jj, xy, st});
  return with_stack(d, EdgesLaunch<WRITE, 32, true>{t, segs, order, n, counts, offsets, ii, jj, xy, st});
}

template <int MAXV, bool PACKED>
struct LocateBaryLaunch {
  const TreeView& t;
  const double* pts;
  int64_t n;
  double tol;
  int64_t* fo;
  double* w;
  hipStream_t st;
  template <int S>
  int run() {
    k_locate_bary<MAXV, S, PACKED><<<grid_for(n), kBlock, 0, st>>>(t, pts, n, tol, fo, w);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

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
  CT_REQUIRE(n >= 0 && tree->n_max_vert >= 3, "bad arguments");
  if (n == 0) return CT_OK;
  return launch_locate_faces(tree, points, nullptr, n, tol, out, as_stream(stream));
}

int ct_pack_leaf_coords(const ct_tree* tree, double* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(tree->n_max_vert >= 3 && tree->n_max_vert <= 8 && out, "n_max_vert must be in [3, 8]");
  k_pack_leaf<<<grid_for(tree->n_elements), kBlock, 0, as_stream(stream)>>>(view(tree), out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

size_t ct_order_workspace_bytes(int64_t n) {
  size_t total = 0;
  carve_order(nullptr, n, false, &total);
  return total;
}

int ct_query_order(int32_t kind, const double* queries, int64_t n, const double* bbox, void* workspace,
                   size_t workspace_bytes, int32_t* order, void* stream) {
  CT_REQUIRE(kind >= 0 && kind <= 2 && n >= 0 && n < ((int64_t)1 << 31) && bbox && order,
             "bad arguments");
  if (n == 0) return CT_OK;
  size_t need = 0;
  OrderWs w = carve_order(static_cast<char*>(workspace), n, false, &need);
  if (workspace_bytes < need) {
    set_error("order workspace too small: %zu < %zu", workspace_bytes, need);
    return CT_ERR_WORKSPACE;
  }
  return query_order(kind, queries, n, bbox, w, order, as_stream(stream));
}

size_t ct_locate_workspace_bytes(int64_t n) {
  size_t total = 0;
  carve_order(nullptr, n, true, &total);
  return total;
}

int ct_locate_points_ordered(const ct_tree* tree, const double* points, int64_t n, double tol,
                             void* workspace, size_t workspace_bytes, int64_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && n < ((int64_t)1 << 31) && tree->n_max_vert >= 3, "bad arguments");
  if (n == 0) return CT_OK;
  size_t need = 0;
  OrderWs w = carve_order(static_cast<char*>(workspace), n, true, &need);
  if (workspace_bytes < need) {
    set_error("locate workspace too small: %zu < %zu", workspace_bytes, need);
    return CT_ERR_WORKSPACE;
  }
  hipStream_t st = as_stream(stream);
  rc = query_order(0, points, n, tree->bbox, w, w.order, st);
  if (rc) return rc;
  return launch_locate_faces(tree, points, w.order, n, tol, out, st);
}

int ct_locate_points_in_order(const ct_tree* tree, const double* points, int64_t n,
                              const int32_t* order, double tol, int64_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && n < ((int64_t)1 << 31) && tree->n_max_vert >= 3, "bad arguments");
  if (n == 0) return CT_OK;
  return launch_locate_faces(tree, points, order, n, tol, out, as_stream(stream));
}

int ct_locate_points_on_edge(const ct_tree* tree, const double* points, int64_t n, double tol,
                             int64_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  CT_REQUIRE(tree->n_max_vert == 2, "edge tree expected (n_max_vert == 2)");
  if (n == 0) return CT_OK;
  return launch_locate_v<4, true, false>(view(tree), tree->depth, points, nullptr, n, tol, out,
                                         as_stream(stream));
}

int ct_locate_boxes_count(const ct_tree* tree, const double* boxes, int64_t n, const int32_t* order,
                          int64_t* counts, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  const TreeView t = view(tree);
  return with_stack(tree->depth, BoxesLaunch<false>{t, boxes, order, n, counts, nullptr, nullptr, nullptr,
                                                    as_stream(stream)});
}

int ct_locate_boxes_write(const ct_tree* tree, const double* boxes, int64_t n, const int32_t* order,
                          const int64_t* offsets, int64_t* ii, int64_t* jj, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  const TreeView t = view(tree);
  return with_stack(tree->depth, BoxesLaunch<true>{t, boxes, order, n, nullptr, offsets, ii, jj,
                                                   as_stream(stream)});
}

int ct_locate_edges_count(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int32_t* order, int64_t* counts, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 || tree->n_max_vert == 2, "edge-edge queries need an edge tree");
  if (n == 0) return CT_OK;
  return launch_edges<false>(tree, segments, order, n, kind, counts, nullptr, nullptr, nullptr, nullptr,
                             as_stream(stream));
}

int ct_locate_edges_write(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int32_t* order, const int64_t* offsets, int64_t* ii, int64_t* jj,
                          double* xy, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 || tree->n_max_vert == 2, "edge-edge queries need an edge tree");
  if (n == 0) return CT_OK;
  return launch_edges<true>(tree, segments, order, n, kind, nullptr, offsets, ii, jj, xy, as_stream(stream));
}

int ct_sort_intersections_by_edge(const double* segments, const int64_t* offsets, const int64_t* ii,
                                  const int64_t* jj, const double* xy, int64_t n_rows,
                                  int64_t* jj_out, double* xy_out, void* stream) {
  CT_REQUIRE(n_rows >= 0, "n_rows must be >= 0");
  if (n_rows == 0) return CT_OK;
  k_sort_by_edge<<<grid_for(n_rows), kBlock, 0, as_stream(stream)>>>(segments, offsets, ii, jj, xy, n_rows,
                                                                     jj_out, xy_out);
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
  if (t.leaf != nullptr && nm <= 8) {
    if (nm <= 3) return with_stack(d, LocateBaryLaunch<3, true>{t, points, n, tol, face_out, weights, st});
    if (nm <= 4) return with_stack(d, LocateBaryLaunch<4, true>{t, points, n, tol, face_out, weights, st});
    return with_stack(d, LocateBaryLaunch<8, true>{t, points, n, tol, face_out, weights, st});
  }
  if (nm <= 3) return with_stack(d, LocateBaryLaunch<3, false>{t, points, n, tol, face_out, weights, st});
  if (nm <= 4) return with_stack(d, LocateBaryLaunch<4, false>{t, points, n, tol, face_out, weights, st});
  if (nm <= 8) return with_stack(d, LocateBaryLaunch<8, false>{t, points, n, tol, face_out, weights, st});
  if (nm <= 16) return with_stack(d, LocateBaryLaunch<16, false>{t, points, n, tol, face_out, weights, st});
  return with_stack(d, LocateBaryLaunch<32, false>{t, points, n, tol, face_out, weights, st});
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
