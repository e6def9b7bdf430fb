// ct_query.hip -- single-tree query kernels on the device (traversals in ct_traverse.h).
//
// Variable-length queries run twice: a count pass, a device exclusive scan
// (ct_exclusive_scan), and a write pass at the scanned offsets -- no atomic append; output
// order = query order, then DFS order.  Queries may be processed in any lane order
// (`order`, from ct_query_order): every query writes only its own count / rows.
#include <mutex>

#include "ct_deep.h"
#include "ct_internal.h"
#include "ct_primitives.h"
#include "ct_traverse.h"

namespace ct {
namespace {

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
template <int MAXV, int STACK, bool ON_EDGE, int LEAF>
__global__ void __launch_bounds__(kBlock) k_locate_points(TreeView t, const double* __restrict__ pts,
                                                          QueryOrder ord, int64_t n, double tol,
                                                          int64_t* __restrict__ out) {
  const int64_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // The query and its result are touched once: streaming (non-temporal) accesses, which
  // leave the L2 to the tree (traversal 0.772 -> 0.759 ms at 1e7 points).  Result rows of
  // box and segment queries do not gain from it: there streaming stores were slower.
  const int64_t q = ord.order ? (int64_t)__builtin_nontemporal_load(ord.order + i) : i;
  const double* src = ord.sorted ? ord.sorted + 2 * i : pts + 2 * q;
  const P2 p = {__builtin_nontemporal_load(src), __builtin_nontemporal_load(src + 1)};
  CT_LANE_STACK(STACK, stack);
  __builtin_nontemporal_store((int64_t)locate_one<MAXV, STACK, ON_EDGE, LEAF>(t, p, tol, stack), out + q);
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
    // streaming: the sort reads each point once, and a cached copy would push the tree out
    // of the Infinity Cache before the traversal
    const double px = __builtin_nontemporal_load(q + 2 * i), py = __builtin_nontemporal_load(q + 2 * i + 1);
    const double2 p = make_double2(px, py);
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
  // Work proxy relative to the tree bbox: a box's area (its hits grow with it); a segment's
  // squared length (its walk grows with its length; its bbox area is 0 when axis-parallel).
  const double rel = (kind == 1 ? w * h : w * w + h * h) * g.inv_area;
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
  // The heaviest class first: workgroups start roughly in order, so the longest walks start
  // early and the light ones fill in behind them (longest-processing-time first).
  return ((unsigned)(15 - cls) << 16) | (m & 0xffffu);
}

// Two-level bucket sort on the 20-bit key, without global atomics:
//   k_sort_hist     per tile of kSortTile queries: the keys (written out) and an LDS
//                   histogram of their top 8 bits, stored bin-major in T[bin][tile]
//   scan of T       -> Ts, the start of every (bin, tile) run of the coarse order
//   k_sort_scatter  per tile: the tile's (key, index) pairs grouped by coarse bin in LDS,
//                   then copied out as one contiguous run per bin (coalesced stores)
//   k_sort_chunks   per chunk of at most kChunk entries of one coarse bin: LDS histogram,
//                   scan and ranks of the low 12 key bits -> that stretch of the order
// Every stage reads and writes each entry once.  A coarse bin (1/256 of the bbox) larger
// than kChunk is ordered chunk by chunk, so the order is exact within chunks of kChunk
// queries and interleaves them within the bin.  Order only schedules lanes: the traversal
// needs a wavefront's queries to be close, and 64 consecutive entries of a 32768-query
// chunk of one coarse cell still span about 1/512 of that cell.  Chunks of 32768 (128 KB of
// staged indices, one workgroup per CU) ordered 1e7 points 6 us slower than chunks of 16384
// but the traversal ran 12 us faster; chunks of 8192 were slower in both.
constexpr int kSortItems = 32;
constexpr int kSortBatch = 8;  // loads in flight per lane
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kCoarse = 256;  // == kBlock
constexpr int kFineBits = 12;
constexpr int kFine = 1 << kFineBits;
constexpr int kChunkBlock = 1024;
constexpr int kChunkItems = 32;  // entries per lane, held in registers
constexpr int kChunk = kChunkBlock * kChunkItems;

__global__ void __launch_bounds__(kBlock) k_sort_hist(int kind, const double* __restrict__ q, int64_t n,
                                                      OrderGeom g, int64_t ntiles,
                                                      unsigned* __restrict__ keys,
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
    for (int j = 0; j < kSortBatch; ++j) {
      if (k[j] != ~0u) {
        atomicAdd(&hist[k[j] >> kFineBits], 1);
        keys[base + (int64_t)(j0 + j) * kBlock + threadIdx.x] = k[j];
      }
    }
  }
  __syncthreads();
  T[(int64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

// The query rows (2 or 4 doubles) go with their key into the coarse runs when `cq` is set,
// so the fine pass can write the queries in final order without gathering them.
__device__ __forceinline__ void copy_row(int kind, const double* __restrict__ from, int64_t i,
                                         double* __restrict__ to, int64_t j) {
  if (kind == 0) {
    reinterpret_cast<double2*>(to)[j] = reinterpret_cast<const double2*>(from)[i];
  } else {
    reinterpret_cast<double4*>(to)[j] = reinterpret_cast<const double4*>(from)[i];
  }
}

__global__ void __launch_bounds__(kBlock) k_sort_scatter(int kind, const double* __restrict__ q, int64_t n,
                                                         int64_t ntiles, const unsigned* __restrict__ keys_in,
                                                         const int64_t* __restrict__ T,
                                                         const int64_t* __restrict__ Ts,
                                                         unsigned* __restrict__ keys,
                                                         int32_t* __restrict__ idx, double* __restrict__ cq) {
  __shared__ unsigned sk[kSortTile];
  __shared__ int si[kSortTile];
  __shared__ int cur[kCoarse];
  __shared__ int64_t shift[kCoarse];  // global start - tile-local start of every bin's run
  __shared__ int lds[kBlock / 64];
  const int64_t tile = blockIdx.x;
  const int b = threadIdx.x;
  int total;
  const int local = block_exclusive_scan<int, kBlock>((int)T[(int64_t)b * ntiles + tile], lds, &total);
  cur[b] = local;
  shift[b] = Ts[(int64_t)b * ntiles + tile] - local;
  __syncthreads();
  const int64_t base = tile * kSortTile;
  for (int j0 = 0; j0 < kSortItems; j0 += kSortBatch) {
    unsigned k[kSortBatch];
#pragma unroll
    for (int j = 0; j < kSortBatch; ++j) {
      const int64_t i = base + (int64_t)(j0 + j) * kBlock + threadIdx.x;
      k[j] = (i < n) ? keys_in[i] : ~0u;
    }
#pragma unroll
    for (int j = 0; j < kSortBatch; ++j) {
      if (k[j] != ~0u) {
        const int pos = atomicAdd(&cur[k[j] >> kFineBits], 1);
        sk[pos] = k[j];
        si[pos] = (j0 + j) * kBlock + threadIdx.x;
      }
    }
  }
  __syncthreads();
  // LDS slot pos of bin c goes to shift[c] + pos: consecutive slots, consecutive addresses.
  for (int pos = threadIdx.x; pos < total; pos += kBlock) {
    const unsigned key = sk[pos];
    const int64_t dst = shift[key >> kFineBits] + pos;
    const int64_t i = base + si[pos];
    keys[dst] = key;
    idx[dst] = (int32_t)i;
    if (cq) copy_row(kind, q, i, cq, dst);
  }
}

__global__ void __launch_bounds__(kChunkBlock) k_sort_chunks(int kind, int64_t ntiles, const int64_t* __restrict__ Ts,
                                                             int64_t n, const unsigned* __restrict__ keys,
                                                             const int32_t* __restrict__ idx,
                                                             const double* __restrict__ cq,
                                                             int32_t* __restrict__ order,
                                                             double* __restrict__ sorted) {
  __shared__ int hist[kFine];
  __shared__ int staged[kChunk];
  __shared__ int lds[kChunkBlock / 64];
  __shared__ int64_t span[2];
  // Chunk c of the grid: coarse bins in order, ceil(count / kChunk) chunks each.
  int64_t blo = 0, bhi = 0;
  int nch = 0;
  if (threadIdx.x < kCoarse) {
    const int c = threadIdx.x;
    blo = Ts[(int64_t)c * ntiles];
    bhi = (c + 1 < kCoarse) ? Ts[(int64_t)(c + 1) * ntiles] : n;
    nch = (int)((bhi - blo + kChunk - 1) / kChunk);
  }
  if (threadIdx.x == 0) span[0] = span[1] = 0;
  int all;
  const int first = block_exclusive_scan<int, kChunkBlock>(nch, lds, &all);
  if ((int)blockIdx.x >= first && (int)blockIdx.x < first + nch) {
    const int64_t lo = blo + (int64_t)((int)blockIdx.x - first) * kChunk;
    span[0] = lo;
    span[1] = lo + kChunk < bhi ? lo + kChunk : bhi;
  }
  for (int f = threadIdx.x; f < kFine; f += kChunkBlock) hist[f] = 0;
  __syncthreads();
  const int64_t lo = span[0], hi = span[1];
  if (lo >= hi) return;  // past the last chunk (the grid is an upper bound); block-uniform
  // v: the query index of each entry, staged by rank so the order is written with one
  // coalesced copy; with `sorted` (rows to gather too) the entry's position instead.
  unsigned k[kChunkItems];
  int v[kChunkItems];
#pragma unroll
  for (int j = 0; j < kChunkItems; ++j) {
    const int64_t p = lo + (int64_t)j * kChunkBlock + threadIdx.x;
    k[j] = p < hi ? keys[p] : ~0u;
    v[j] = sorted ? j * kChunkBlock + (int)threadIdx.x : (p < hi ? idx[p] : 0);
  }
#pragma unroll
  for (int j = 0; j < kChunkItems; ++j)
    if (k[j] != ~0u) atomicAdd(&hist[k[j] & (kFine - 1)], 1);
  __syncthreads();
  constexpr int per = kFine / kChunkBlock;
  int s = 0;
#pragma unroll
  for (int f = 0; f < per; ++f) s += hist[threadIdx.x * per + f];
  int total;
  int run = block_exclusive_scan<int, kChunkBlock>(s, lds, &total);
#pragma unroll
  for (int f = 0; f < per; ++f) {
    const int c = hist[threadIdx.x * per + f];
    hist[threadIdx.x * per + f] = run;
    run += c;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kChunkItems; ++j) {
    if (k[j] != ~0u) {
      const int pos = atomicAdd(&hist[k[j] & (kFine - 1)], 1);
      staged[pos] = v[j];
    }
  }
  __syncthreads();
  const int m = (int)(hi - lo);
  if (sorted) {
    for (int pos = threadIdx.x; pos < m; pos += kChunkBlock) {
      const int64_t src = lo + staged[pos];
      order[lo + pos] = idx[src];
      copy_row(kind, cq, src, sorted, lo + pos);
    }
  } else {
    // streaming store: the traversal reads the order once, with a streaming load
    for (int pos = threadIdx.x; pos < m; pos += kChunkBlock) __builtin_nontemporal_store(staged[pos], order + lo + pos);
  }
}

struct OrderWs {
  int64_t* T;
  int64_t* Ts;
  void* scan_ws;
  size_t scan_bytes;
  unsigned* kraw;  // keys in query order
  unsigned* keys;  // keys in coarse-bin order
  int32_t* idx;
  double* cq;      // query rows in coarse-bin order (up to 4 doubles per query)
  int32_t* order;  // only used by ct_locate_points_ordered
  double* sorted;  // only used by ct_locate_points_ordered
  int64_t ntiles;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// with_rows: also the query rows in coarse-bin order (cq) and the sorted queries, which a
// caller needs only when it asks query_order for `sorted`.
OrderWs carve_order(char* base, int64_t n, bool with_order, size_t* total, bool with_rows = true) {
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
  w.kraw = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * nn));
  w.keys = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * nn));
  w.idx = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * nn));
  w.cq = with_rows ? reinterpret_cast<double*>(take(sizeof(double) * 4 * nn)) : nullptr;
  w.order = with_order ? reinterpret_cast<int32_t*>(take(sizeof(int32_t) * nn)) : nullptr;
  w.sorted = with_order && with_rows ? reinterpret_cast<double*>(take(sizeof(double) * 2 * nn)) : nullptr;
  *total = off + 256;
  return w;
}

// `sorted` (may be null) receives the queries in the final order.
int query_order(int kind, const double* q, int64_t n, const double* bbox, const OrderWs& w,
                int32_t* order, double* sorted, hipStream_t st) {
  OrderGeom g;
  g.bb = Box{bbox[0], bbox[1], bbox[2], bbox[3]};
  const double ex = g.bb.xmax - g.bb.xmin, ey = g.bb.ymax - g.bb.ymin;
  g.sx = 1.0 / ex;
  g.sy = 1.0 / ey;
  g.inv_area = 1.0 / (ex * ey);
  const unsigned ntiles = (unsigned)w.ntiles;
  k_sort_hist<<<ntiles, kBlock, 0, st>>>(kind, q, n, g, w.ntiles, w.kraw, w.T);
  CT_LAUNCH_CHECK();
  int rc = exclusive_scan(w.T, (int64_t)kCoarse * w.ntiles, w.Ts, w.scan_ws, w.scan_bytes, st);
  if (rc) return rc;
  double* cq = sorted ? w.cq : nullptr;
  k_sort_scatter<<<ntiles, kBlock, 0, st>>>(kind, q, n, w.ntiles, w.kraw, w.T, w.Ts, w.keys, w.idx, cq);
  CT_LAUNCH_CHECK();
  // sum over bins of ceil(count / kChunk) <= ceil(n / kChunk) + kCoarse
  const unsigned nchunks = (unsigned)((n + kChunk - 1) / kChunk + kCoarse);
  k_sort_chunks<<<nchunks, kChunkBlock, 0, st>>>(kind, w.ntiles, w.Ts, n, w.keys, w.idx, cq, order, sorted);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

// Start-node grid (see ct_start_grid): one lane per grid cell.
__global__ void __launch_bounds__(kBlock) k_start_grid(TreeView t, int n, int32_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)n * n) return;
  const int gx = (int)(k % n), gy = (int)(k / n);
  double x0, x1, y0, y1;
  grid_span(t.bbox.xmin, t.bbox.xmax, n, gx, &x0, &x1);
  grid_span(t.bbox.ymin, t.bbox.ymax, n, gy, &y0, &y1);
  int node = 0;
  while (true) {
    const NodeLd nd = ld_node(t.nodes, node);
    if (nd.child == -1) break;
    const double lo = nd.dim ? y0 : x0;
    const double hi = nd.dim ? y1 : x1;
    if (hi <= nd.Lmax && hi < nd.Rmin) node = nd.child;               // only the left child, always
    else if (lo > nd.Lmax && lo >= nd.Rmin) node = nd.child + 1;      // only the right child, always
    else break;
  }
  out[k] = node;
}

// Leaf-ordered cell bboxes (see ct_pack_leaf_bboxes).
__global__ void __launch_bounds__(kBlock) k_pack_leaf_bb(TreeView t, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.n_elements) return;
  const int c = t.bb_indices[k];
  reinterpret_cast<double4*>(out)[k] = reinterpret_cast<const double4*>(t.bb)[c];
}

// Packed leaf coordinates (see ct_pack_leaf_coords).
__global__ void __launch_bounds__(kBlock) k_pack_leaf(TreeView t, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.n_elements) return;
  const int c = t.bb_indices[k];
  const int32_t* row = t.elements + (int64_t)c * t.nmax;
  const int len = poly_len(row, t.nmax);
  double2* dst = reinterpret_cast<double2*>(out) + k * t.nmax;
  P2 v = {0.0, 0.0};
  for (int i = 0; i < t.nmax; ++i) {
    if (i < len) v = ld_vertex(t.vertices, row[i]);
    dst[i] = make_double2(v.x, v.y);  // slots past the polygon repeat its last vertex
  }
}

// ----------------------------------------------------------------- boxes
// query.py:177-244 locate_box: strict overlap with the tree bbox first; the left child is
// pushed before the right, so the right child is visited first.  With a hit record (hits !=
// null, slot-major as in k_edges) the count pass keeps the first `slots` hit cells, and the
// write pass copies them for boxes with at most `slots` hits instead of walking again.
template <bool WRITE, int STACK>
__global__ void __launch_bounds__(kBlock) k_boxes(TreeView t, const double* __restrict__ boxes,
                                                  QueryOrder ord, int64_t n,
                                                  int64_t* __restrict__ counts,
                                                  const int64_t* __restrict__ offsets,
                                                  int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                                  int32_t* __restrict__ hits, int slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(ord.order, i);
  if (WRITE && hits != nullptr) {
    const int64_t o = offsets[q], total = offsets[q + 1] - o;
    if (total <= slots) {
      for (int64_t h = 0; h < total; ++h) {
        ii[o + h] = q;
        jj[o + h] = hits[h * n + i];
      }
      return;
    }
  }
  const Box box = ord.sorted ? ld_box(ord.sorted, i) : ld_box(boxes, q);
  // Box queries keep deep stacks (a box often overlaps both children), which would spill
  // past LaneStack's LDS slots on most queries: a VGPR array is faster here.
  int stack[STACK];
  int64_t cnt = 0;
  int64_t o = WRITE ? offsets[q] : 0;
  if (boxes_intersect(box, t.bbox)) {
    int sp = 0;
    int node = 0;
    // Internal nodes in an inner loop; the leaf's cells once the whole wavefront is at a
    // leaf or finished (the loop shape of locate_one).
    while (true) {
      NodeLd nd = ld_node(t.nodes, node);
      bool exhausted = false;
      while (nd.child != -1) {
        const double bmin = nd.dim ? box.ymin : box.xmin;
        const double bmax = nd.dim ? box.ymax : box.xmax;
        const bool left = bmin <= nd.Lmax;
        const bool right = bmax >= nd.Rmin;
        if (left && right) {
          stack[sp++] = nd.child;
          node = nd.child + 1;
        } else if (left || right) {
          node = nd.child + (right ? 1 : 0);
        } else {
          if (sp == 0) {
            exhausted = true;
            break;
          }
          node = stack[--sp];
        }
        nd = ld_node(t.nodes, node);
      }
      if (exhausted) break;
      for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
        const int c = t.bb_indices[k];
        if (boxes_intersect(box, leaf_box(t, k, c))) {
          if (WRITE) {
            ii[o] = q;
            jj[o] = c;
            ++o;
          } else {
            if (hits != nullptr && cnt < slots) hits[cnt * n + i] = c;
            ++cnt;
          }
        }
      }
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
  if (!WRITE) counts[q] = cnt;
}

// ----------------------------------------------------------------- segments
// With a hit record (hits != null, slot-major: hit c of the lane at position i in
// hits[c * n + i]), the count pass keeps the first `slots` hit cells of every query, and the
// write pass replays the queries with at most `slots` hits: the same cell predicates on the
// recorded cells, in the recorded (DFS) order, so the rows are identical.  Queries with more
// hits walk the tree again.
template <bool WRITE, int MAXV, int STACK, bool FACES>
__global__ void __launch_bounds__(kBlock) k_edges(TreeView t, const double* __restrict__ segs,
                                                  QueryOrder ord, int64_t n,
                                                  int64_t* __restrict__ counts,
                                                  const int64_t* __restrict__ offsets,
                                                  int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                                  double* __restrict__ xy, int32_t* __restrict__ hits,
                                                  int slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(ord.order, i);
  const double4 s4 = ord.sorted ? reinterpret_cast<const double4*>(ord.sorted)[i]
                                : reinterpret_cast<const double4*>(segs)[q];
  const P2 a = {s4.x, s4.y}, b = {s4.z, s4.w};
  CT_LANE_STACK(STACK, stack);
  if (WRITE) {
    const int64_t o = offsets[q];
    const int64_t cnt = offsets[q + 1] - o;
    if (hits != nullptr && cnt <= slots) {
      for (int64_t h = 0; h < cnt; ++h) {
        const int c = hits[h * n + i];
        P2 cc, dd;
        if (FACES) {
          edge_face<MAXV>(t, c, a, b, &cc, &dd);
        } else {
          edge_edge(t, c, a, b, &cc, &dd);
        }
        ii[o + h] = q;
        jj[o + h] = c;
        reinterpret_cast<double4*>(xy)[o + h] = make_double4(cc.x, cc.y, dd.x, dd.y);
      }
    } else {
      edge_query<true, MAXV, STACK, FACES>(t, a, b, q, o, FinalRows{ii, jj, xy}, stack);
    }
  } else {
    HitRecord rec;
    if (hits != nullptr) rec = HitRecord{hits + i, slots, n};
    counts[q] = edge_query<false, MAXV, STACK, FACES>(t, a, b, q, 0, FinalRows{}, stack, rec);
  }
}

__global__ void __launch_bounds__(kBlock) k_sort_by_edge(const double* __restrict__ segs,
                                                         const int64_t* __restrict__ offsets,
                                                         const int64_t* __restrict__ ii,
                                                         const int64_t* __restrict__ jj_in,
                                                         const double* __restrict__ xy_in,
                                                         int64_t n_rows, int64_t* __restrict__ jj_out,
                                                         double* __restrict__ xy_out) {
  __shared__ double tw[kRankWindow];
  sort_tile(segs, offsets, ii, jj_in, xy_in, n_rows, (int64_t)blockIdx.x * kRankTile, jj_out, xy_out, tw);
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

template <int MAXV, int STACK, int LEAF>
__global__ void __launch_bounds__(kBlock) k_locate_bary(TreeView t, const double* __restrict__ pts,
                                                        int64_t n, double tol,
                                                        int64_t* __restrict__ face_out,
                                                        double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P2 p = ld_point(pts, i);
  CT_LANE_STACK(STACK, stack);
  const int f = locate_one<MAXV, STACK, false, LEAF>(t, p, tol, stack);
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

// Tight node boxes, bottom-up in rounds without inter-workgroup synchronisation: the leaf
// pass reduces every leaf's cell boxes (stamp 0); round r (r = 1 .. depth) forms the union of
// the two children of every internal node whose children were both stamped before round r
// (stamp < r) and stamps it r.  A node of height h is finished in round h, so `depth` rounds
// finish the root; children's rows are read only after the kernel that wrote them has ended.
// NaN poisons the union.
constexpr int kUnstamped = 0x7fffffff;

__device__ __forceinline__ double4 box_union(double4 a, double4 b) {
  const double nan = __builtin_nan("");
  const bool bad = isnan(a.x) || isnan(a.y) || isnan(a.z) || isnan(a.w) || isnan(b.x) || isnan(b.y) ||
                   isnan(b.z) || isnan(b.w);
  if (bad) return make_double4(nan, nan, nan, nan);
  return make_double4(b.x < a.x ? b.x : a.x, b.y > a.y ? b.y : a.y, b.z < a.z ? b.z : a.z,
                      b.w > a.w ? b.w : a.w);
}

__global__ void __launch_bounds__(kBlock) k_node_boxes_leaves(TreeView t, int64_t n_nodes,
                                                              int32_t* __restrict__ stamp,
                                                              double* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n_nodes) return;
  const ct_node nd = t.nodes[x];
  if (nd.child != -1) {
    stamp[x] = kUnstamped;
    return;
  }
  const double inf = __builtin_inf();
  double4 u = make_double4(inf, -inf, inf, -inf);
  for (int k = nd.ptr; k < nd.ptr + nd.size; ++k) {
    const int c = t.bb_indices[k];
    u = box_union(u, reinterpret_cast<const double4*>(t.bb)[c]);
  }
  reinterpret_cast<double4*>(out)[x + 1] = u;
  stamp[x] = 0;
}

__global__ void __launch_bounds__(kBlock) k_node_boxes_round(TreeView t, int64_t n_nodes, int r,
                                                             int32_t* __restrict__ stamp,
                                                             double* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n_nodes || stamp[x] != kUnstamped) return;
  const int c = t.nodes[x].child;
  if (c < 0 || (int64_t)c + 1 >= n_nodes) return;
  if (stamp[c] >= r || stamp[c + 1] >= r) return;
  double4* rows = reinterpret_cast<double4*>(out) + 1;
  rows[x] = box_union(rows[c], rows[c + 1]);
  stamp[x] = r;
}

// Longest parent chain k_node_bounds follows: deeper than any tree it is given (a chain that
// long means edited child links formed a cycle; the climb then stops).
constexpr int kClimbLimit = 1 << 16;

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
  for (int guard = 0; guard < kClimbLimit && y > 0; ++guard) {
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
// Trees deeper than CT_MAX_DEPTH are walked without a stack (ct_deep.hip); every traversal
// launcher below hands them over first.
int check_tree(const ct_tree* tree) {
  CT_REQUIRE(tree != nullptr && tree->nodes && tree->bb_indices && tree->elements && tree->vertices,
             "incomplete tree");
  CT_REQUIRE(tree->depth >= 0, "negative tree depth");
  return CT_OK;
}

bool is_deep(const ct_tree* tree) { return tree->depth > CT_MAX_DEPTH; }

template <int MAXV, bool ON_EDGE, int LEAF>
struct LocateLaunch {
  const TreeView& t;
  const double* pts;
  QueryOrder order;
  int64_t n;
  double tol;
  int64_t* out;
  hipStream_t st;
  template <int S>
  int run() {
    k_locate_points<MAXV, S, ON_EDGE, LEAF><<<grid_xcd(n), kBlock, 0, st>>>(t, pts, order, n, tol, out);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <int MAXV, bool ON_EDGE, int LEAF>
int launch_locate_v(const TreeView& t, int depth, const double* pts, QueryOrder order, int64_t n,
                    double tol, int64_t* out, hipStream_t st) {
  return with_stack(depth, LocateLaunch<MAXV, ON_EDGE, LEAF>{t, pts, order, n, tol, out, st});
}

// Face trees: picks the vertex-count class, and the packed leaf path when present.
int launch_locate_faces(const ct_tree* tree, const double* pts, QueryOrder order, int64_t n,
                        double tol, int64_t* out, hipStream_t st) {
  if (is_deep(tree)) return deep_locate_points(tree, pts, order.order, n, tol, false, out, st);
  const TreeView t = view(tree);
  const int nm = tree->n_max_vert, d = tree->depth;
  if (t.leaf != nullptr && nm <= 8) {
    if (nm <= 3) return launch_locate_v<3, false, kLeafPackedExact>(t, d, pts, order, n, tol, out, st);
    if (nm <= 4) return launch_locate_v<4, false, kLeafPackedExact>(t, d, pts, order, n, tol, out, st);
    return launch_locate_v<8, false, kLeafPacked>(t, d, pts, order, n, tol, out, st);
  }
  if (nm <= 3) return launch_locate_v<3, false, kLeafFaces>(t, d, pts, order, n, tol, out, st);
  if (nm <= 4) return launch_locate_v<4, false, kLeafFaces>(t, d, pts, order, n, tol, out, st);
  if (nm <= 8) return launch_locate_v<8, false, kLeafFaces>(t, d, pts, order, n, tol, out, st);
  if (nm <= 16) return launch_locate_v<16, false, kLeafFaces>(t, d, pts, order, n, tol, out, st);
  return launch_locate_v<32, false, kLeafFaces>(t, d, pts, order, n, tol, out, st);
}

template <bool WRITE>
struct BoxesLaunch {
  const TreeView& t;
  const double* boxes;
  QueryOrder order;
  int64_t n;
  int64_t* counts;
  const int64_t* offsets;
  int64_t* ii;
  int64_t* jj;
  int32_t* hits;
  int slots;
  hipStream_t st;
  template <int S>
  int run() {
    k_boxes<WRITE, S><<<grid_for(n), kBlock, 0, st>>>(t, boxes, order, n, counts, offsets, ii, jj, hits, slots);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <bool WRITE, int MAXV, bool FACES>
struct EdgesLaunch {
  const TreeView& t;
  const double* segs;
  QueryOrder order;
  int64_t n;
  int64_t* counts;
  const int64_t* offsets;
  int64_t* ii;
  int64_t* jj;
  double* xy;
  int32_t* hits;
  int slots;
  hipStream_t st;
  template <int S>
  int run() {
    k_edges<WRITE, MAXV, S, FACES><<<grid_for(n), kBlock, 0, st>>>(t, segs, order, n, counts, offsets, ii, jj, xy,
                                                                    hits, slots);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <bool WRITE>
int launch_edges(const ct_tree* tree, const double* segs, QueryOrder order, int64_t n, int kind,
                 int64_t* counts, const int64_t* offsets, int64_t* ii, int64_t* jj, double* xy,
                 int32_t* hits, int slots, hipStream_t st) {
  if (is_deep(tree))  // no hit record: both passes walk
    return deep_edges(tree, segs, order.order, n, kind == 1, WRITE, counts, offsets, ii, jj, xy, st);
  const TreeView t = view(tree);
  const int d = tree->depth, nm = tree->n_max_vert;
  if (kind == 0)
    return with_stack(d, EdgesLaunch<WRITE, 4, false>{t, segs, order, n, counts, offsets, ii, jj, xy, hits, slots, st});
  if (nm <= 4)
    return with_stack(d, EdgesLaunch<WRITE, 4, true>{t, segs, order, n, counts, offsets, ii, jj, xy, hits, slots, st});
  if (nm <= 8)
    return with_stack(d, EdgesLaunch<WRITE, 8, true>{t, segs, order, n, counts, offsets, ii, jj, xy, hits, slots, st});
  if (nm <= 16)
    return with_stack(d, EdgesLaunch<WRITE, 16, true>{t, segs, order, n, counts, offsets, ii, jj, xy, hits, slots,
                                                      st});
  return with_stack(d, EdgesLaunch<WRITE, 32, true>{t, segs, order, n, counts, offsets, ii, jj, xy, hits, slots, st});
}

template <int MAXV, int LEAF>
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
    k_locate_bary<MAXV, S, LEAF><<<grid_for(n), kBlock, 0, st>>>(t, pts, n, tol, fo, w);
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

// ------------------------------------------------------------ allocation-free locate
// The workspace of ct_locate_points_ordered_noalloc: the order kernels' keys, indices, tile
// histograms and the order itself, 16 bytes per point plus about 1 MiB at the maximum of
// CT_NOALLOC_MAX_POINTS points (checked on every call).  Each device holds its own copy for
// the life of the process, and every call on that device reuses it: concurrent calls on one
// device race on it.  Nothing in it needs clearing (every byte a call reads, the same call
// writes first).
static constexpr size_t kNoAllocScratchBytes = size_t(36) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of the current device's `scratch`, resolved once per device (the host
// name `scratch` is only a stand-in and must never reach a kernel).
int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("hipGetDevice: %s", hipGetErrorString(e));
    return CT_ERR_HIP;
  }
  CT_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "device %d beyond the %d supported", dev, kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return CT_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CT_OK;
}
}  // namespace

extern "C" {

int ct_locate_points_ordered_noalloc(const ct_tree* tree, const double* points, int64_t n, double tol,
                                     int64_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && tree->n_max_vert >= 3 && (n == 0 || (points && out)), "bad arguments");
  CT_REQUIRE(n <= CT_NOALLOC_MAX_POINTS, "%lld points exceed the allocation-free maximum of %d",
             (long long)n, CT_NOALLOC_MAX_POINTS);
  if (n == 0) return CT_OK;
  void* ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  size_t need = 0;
  const OrderWs w = carve_order(static_cast<char*>(ws), n, true, &need, /*with_rows=*/false);
  if (need > kNoAllocScratchBytes) {
    set_error("allocation-free workspace too small: %zu < %zu", kNoAllocScratchBytes, need);
    return CT_ERR_WORKSPACE;
  }
  hipStream_t st = as_stream(stream);
  rc = query_order(0, points, n, tree->bbox, w, w.order, nullptr, st);
  if (rc) return rc;
  return launch_locate_faces(tree, points, QueryOrder{w.order, nullptr}, n, tol, out, st);
}

int ct_locate_points(const ct_tree* tree, const double* points, int64_t n, double tol, int64_t* out,
                     void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && tree->n_max_vert >= 3, "bad arguments");
  if (n == 0) return CT_OK;
  return launch_locate_faces(tree, points, QueryOrder{nullptr, nullptr}, n, tol, out, as_stream(stream));
}

int ct_pack_leaf_coords(const ct_tree* tree, double* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(tree->n_max_vert >= 3 && tree->n_max_vert <= 8 && out, "n_max_vert must be in [3, 8]");
  k_pack_leaf<<<grid_for(tree->n_elements), kBlock, 0, as_stream(stream)>>>(view(tree), out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_pack_leaf_bboxes(const ct_tree* tree, double* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(out != nullptr, "null output");
  if (tree->n_elements == 0) return CT_OK;
  k_pack_leaf_bb<<<grid_for(tree->n_elements), kBlock, 0, as_stream(stream)>>>(view(tree), out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_start_grid(const ct_tree* tree, int32_t n, int32_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 1 && n <= 4096 && out != nullptr, "grid size must be in [1, 4096]");
  ct_tree t = *tree;
  t.start_grid = nullptr;
  k_start_grid<<<grid_for((int64_t)n * n), kBlock, 0, as_stream(stream)>>>(view(t), n, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

size_t ct_order_workspace_bytes(int64_t n) {
  size_t total = 0;
  carve_order(nullptr, n, false, &total);
  return total;
}

int ct_query_order(int32_t kind, const double* queries, int64_t n, const double* bbox, void* workspace,
                   size_t workspace_bytes, int32_t* order, double* sorted, void* stream) {
  CT_REQUIRE(kind >= 0 && kind <= 2 && n >= 0 && n < ((int64_t)1 << 31) && bbox && order,
             "bad arguments");
  if (n == 0) return CT_OK;
  size_t need = 0;
  OrderWs w = carve_order(static_cast<char*>(workspace), n, false, &need);
  if (workspace_bytes < need) {
    set_error("order workspace too small: %zu < %zu", workspace_bytes, need);
    return CT_ERR_WORKSPACE;
  }
  return query_order(kind, queries, n, bbox, w, order, sorted, as_stream(stream));
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
  rc = query_order(0, points, n, tree->bbox, w, w.order, w.sorted, st);
  if (rc) return rc;
  return launch_locate_faces(tree, points, QueryOrder{w.order, w.sorted}, n, tol, out, st);
}

int ct_locate_points_in_order(const ct_tree* tree, const double* points, int64_t n,
                              const int32_t* order, const double* sorted, double tol, int64_t* out,
                              void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && n < ((int64_t)1 << 31) && tree->n_max_vert >= 3, "bad arguments");
  CT_REQUIRE(sorted == nullptr || order != nullptr, "sorted queries need their order");
  if (n == 0) return CT_OK;
  return launch_locate_faces(tree, points, QueryOrder{order, sorted}, n, tol, out, as_stream(stream));
}

int ct_locate_points_on_edge(const ct_tree* tree, const double* points, int64_t n, double tol,
                             int64_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  CT_REQUIRE(tree->n_max_vert == 2, "edge tree expected (n_max_vert == 2)");
  if (n == 0) return CT_OK;
  if (is_deep(tree)) return deep_locate_points(tree, points, nullptr, n, tol, true, out, as_stream(stream));
  return launch_locate_v<4, true, kLeafFaces>(view(tree), tree->depth, points, QueryOrder{nullptr, nullptr}, n, tol,
                                         out, as_stream(stream));
}

int ct_locate_boxes_count_rec(const ct_tree* tree, const double* boxes, int64_t n, const int32_t* order,
                              const double* sorted, int64_t* counts, int32_t* hits, int32_t hit_slots,
                              void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  CT_REQUIRE(sorted == nullptr || order != nullptr, "sorted queries need their order");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n == 0) return CT_OK;
  if (is_deep(tree))  // no hit record: both passes walk
    return deep_boxes(tree, boxes, order, n, false, counts, nullptr, nullptr, nullptr, as_stream(stream));
  const TreeView t = view(tree);
  return with_stack(tree->depth, BoxesLaunch<false>{t, boxes, QueryOrder{order, sorted}, n, counts, nullptr,
                                                    nullptr, nullptr, hits, hits ? hit_slots : 0,
                                                    as_stream(stream)});
}

int ct_locate_boxes_write_rec(const ct_tree* tree, const double* boxes, int64_t n, const int32_t* order,
                              const double* sorted, const int64_t* offsets, const int32_t* hits,
                              int32_t hit_slots, int64_t* ii, int64_t* jj, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0, "n must be >= 0");
  CT_REQUIRE(sorted == nullptr || order != nullptr, "sorted queries need their order");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n == 0) return CT_OK;
  if (is_deep(tree)) return deep_boxes(tree, boxes, order, n, true, nullptr, offsets, ii, jj, as_stream(stream));
  const TreeView t = view(tree);
  return with_stack(tree->depth, BoxesLaunch<true>{t, boxes, QueryOrder{order, sorted}, n, nullptr, offsets, ii,
                                                   jj, const_cast<int32_t*>(hits), hits ? hit_slots : 0,
                                                   as_stream(stream)});
}

int ct_locate_boxes_count(const ct_tree* tree, const double* boxes, int64_t n, const int32_t* order,
                          const double* sorted, int64_t* counts, void* stream) {
  return ct_locate_boxes_count_rec(tree, boxes, n, order, sorted, counts, nullptr, 0, stream);
}

int ct_locate_boxes_write(const ct_tree* tree, const double* boxes, int64_t n, const int32_t* order,
                          const double* sorted, const int64_t* offsets, int64_t* ii, int64_t* jj,
                          void* stream) {
  return ct_locate_boxes_write_rec(tree, boxes, n, order, sorted, offsets, nullptr, 0, ii, jj, stream);
}

int ct_locate_edges_count_rec(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                              const int32_t* order, const double* sorted, int64_t* counts, int32_t* hits,
                              int32_t hit_slots, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 || tree->n_max_vert == 2, "edge-edge queries need an edge tree");
  CT_REQUIRE(sorted == nullptr || order != nullptr, "sorted queries need their order");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n == 0) return CT_OK;
  return launch_edges<false>(tree, segments, QueryOrder{order, sorted}, n, kind, counts, nullptr, nullptr,
                             nullptr, nullptr, hits, hits ? hit_slots : 0, as_stream(stream));
}

int ct_locate_edges_write_rec(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                              const int32_t* order, const double* sorted, const int64_t* offsets,
                              const int32_t* hits, int32_t hit_slots, int64_t* ii, int64_t* jj, double* xy,
                              void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n >= 0 && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 || tree->n_max_vert == 2, "edge-edge queries need an edge tree");
  CT_REQUIRE(sorted == nullptr || order != nullptr, "sorted queries need their order");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n == 0) return CT_OK;
  return launch_edges<true>(tree, segments, QueryOrder{order, sorted}, n, kind, nullptr, offsets, ii, jj, xy,
                            const_cast<int32_t*>(hits), hits ? hit_slots : 0, as_stream(stream));
}

int ct_locate_edges_count(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int32_t* order, const double* sorted, int64_t* counts, void* stream) {
  return ct_locate_edges_count_rec(tree, segments, n, kind, order, sorted, counts, nullptr, 0, stream);
}

int ct_locate_edges_write(const ct_tree* tree, const double* segments, int64_t n, int32_t kind,
                          const int32_t* order, const double* sorted, const int64_t* offsets, int64_t* ii,
                          int64_t* jj, double* xy, void* stream) {
  return ct_locate_edges_write_rec(tree, segments, n, kind, order, sorted, offsets, nullptr, 0, ii, jj, xy,
                                   stream);
}

int ct_sort_intersections_by_edge(const double* segments, const int64_t* offsets, const int64_t* ii,
                                  const int64_t* jj, const double* xy, int64_t n_rows,
                                  int64_t* jj_out, double* xy_out, void* stream) {
  CT_REQUIRE(n_rows >= 0, "n_rows must be >= 0");
  if (n_rows == 0) return CT_OK;
  k_sort_by_edge<<<grid_for(n_rows, kRankTile), kBlock, 0, as_stream(stream)>>>(segments, offsets, ii, jj, xy,
                                                                                n_rows, jj_out, xy_out);
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
  hipStream_t st = as_stream(stream);
  const int nm = tree->n_max_vert, d = tree->depth;
  if (is_deep(tree)) {  // locate, then the weights of the located faces (same arithmetic)
    rc = deep_locate_points(tree, points, nullptr, n, tol, false, face_out, st);
    if (rc) return rc;
    return launch_bary(points, face_out, n, tree->elements, nm, tree->vertices, tol, nm > 3, weights, st);
  }
  const TreeView t = view(tree);
  if (t.leaf != nullptr && nm <= 8) {
    if (nm <= 3) return with_stack(d, LocateBaryLaunch<3, kLeafPackedExact>{t, points, n, tol, face_out, weights, st});
    if (nm <= 4) return with_stack(d, LocateBaryLaunch<4, kLeafPackedExact>{t, points, n, tol, face_out, weights, st});
    return with_stack(d, LocateBaryLaunch<8, kLeafPacked>{t, points, n, tol, face_out, weights, st});
  }
  if (nm <= 3) return with_stack(d, LocateBaryLaunch<3, kLeafFaces>{t, points, n, tol, face_out, weights, st});
  if (nm <= 4) return with_stack(d, LocateBaryLaunch<4, kLeafFaces>{t, points, n, tol, face_out, weights, st});
  if (nm <= 8) return with_stack(d, LocateBaryLaunch<8, kLeafFaces>{t, points, n, tol, face_out, weights, st});
  if (nm <= 16) return with_stack(d, LocateBaryLaunch<16, kLeafFaces>{t, points, n, tol, face_out, weights, st});
  return with_stack(d, LocateBaryLaunch<32, kLeafFaces>{t, points, n, tol, face_out, weights, st});
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

int ct_node_parents(const ct_tree* tree, int32_t* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(tree->n_nodes >= 1 && out, "bad arguments");
  hipStream_t st = as_stream(stream);
  rc = fill<int32_t>(out, tree->n_nodes, -1, st);
  if (rc) return rc;
  k_parents<<<grid_for(tree->n_nodes), kBlock, 0, st>>>(tree->nodes, tree->n_nodes, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_node_boxes(const ct_tree* tree, int32_t* scratch, double* out, void* stream) {
  int rc = check_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(tree->n_nodes >= 1 && tree->bb_coords && scratch && out, "bad arguments");
  const int64_t n = tree->n_nodes;
  hipStream_t st = as_stream(stream);
  const double nan = __builtin_nan("");
  rc = fill<double>(out, 4, nan, st);  // row 0 (unused)
  if (rc) return rc;
  const TreeView t = view(tree);
  k_node_boxes_leaves<<<grid_for(n), kBlock, 0, st>>>(t, n, scratch, out);
  CT_LAUNCH_CHECK();
  for (int r = 1; r <= tree->depth; ++r) {
    k_node_boxes_round<<<grid_for(n), kBlock, 0, st>>>(t, n, r, scratch, out);
    CT_LAUNCH_CHECK();
  }
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
