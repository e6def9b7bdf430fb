// ct_split.hip -- work splitting for box and segment queries: subqueries by subtree.
//
// A box or segment query walks its viable subtrees depth first (query.py:177-244, 357-455)
// and emits the hit cells of each leaf in visit order.  The largest queries (wide boxes,
// long segments) walk hundreds of leaves in one lane; a few thousand such lanes cannot fill
// the GPU and set the kernel's tail.  Here such a query is cut into subqueries by subtree:
// a frontier walk from the root repeatedly replaces the frontier subtree holding the most
// cells by its viable children, in the walk's visit order (right child first, as the
// reference's push order gives), until every frontier subtree holds at most `leaf_cells`
// cells or the frontier has `max_parts` entries.  Visiting a node visits its viable
// children's subtrees one after the other, so the frontier subtrees, walked one by one in
// frontier order, emit exactly the query's rows in the query's order: one lane per subquery,
// rows concatenated per query, no deduplication and no sort.  Queries whose expected work is
// small keep one subquery (the root).
//
// Passes (the Python layer chains them; every array is per processing position i, or per
// subquery j in processing order):
//   ct_split_queries (count)   parts[i] = number of subqueries of query order[i]
//   ct_split_queries (write)   subs[part_offsets[i] + k] = {order[i], frontier node k}
// boxes:
//   ct_locate_boxes_count_sub  counts[j] (+ hit record per subquery)
//   ct_subquery_offsets        counts[j] -> the exclusive prefix within its query;
//                              per_query[q] = the query's total (then ct_exclusive_scan)
//   ct_locate_boxes_write_sub  rows of subquery j at query_offsets[q] + prefix[j]
// segments (every hit is clipped once, in the count pass, and its piece kept):
//   ct_locate_edges_count_sub  walk + candidate list, wave-cooperative clip, counts[j], the
//                              hit cells and clipped pieces of each subquery
//   (exclusive scan of counts) staging offsets sub_offsets[j]
//   ct_subquery_spans          per_query[q] and each query's first subquery
//   (exclusive scan)           query_start[q]
//   ct_locate_edges_write_sub  rows (query, cell, piece, sort key t) into staging
//   ct_sort_staged_rows        each query's rows ranked by t into the output
#include "ct_internal.h"
#include "ct_primitives.h"
#include "ct_traverse.h"

namespace ct {
namespace {

constexpr int kMaxParts = 16;

// query.py:228-242: which children of internal node nd box may reach.
__device__ __forceinline__ void box_children(const NodeLd& nd, const Box& box, bool* left, bool* right) {
  const double bmin = nd.dim ? box.ymin : box.xmin;
  const double bmax = nd.dim ? box.ymax : box.xmax;
  *left = bmin <= nd.Lmax;
  *right = bmax >= nd.Rmin;
}

struct SplitParams {
  double cell;     // typical cell side: sqrt(tree bbox area / n_elements)
  int min_cells;   // split queries expected to meet more cells than this
  int leaf_cells;  // stop expanding subtrees of at most this many cells
  int max_parts;   // at most this many subqueries per query (<= kMaxParts)
};

// Query q of kind 1 (boxes) or 2 (segments; FACES: segment-face, else segment-edge).
struct SplitQuery {
  int kind;
  Box box;      // kind 1
  P2 a, b;      // kind 2
  double vx, vy;
  SegPrune sp;
};

template <bool FACES>
__device__ __forceinline__ SplitQuery load_split_query(const TreeView& t, int kind, const double* __restrict__ qs,
                                                       int64_t q) {
  SplitQuery s{};
  s.kind = kind;
  if (kind == 1) {
    s.box = ld_box(qs, q);
  } else {
    const double4 d = reinterpret_cast<const double4*>(qs)[q];
    s.a = {d.x, d.y};
    s.b = {d.z, d.w};
    s.vx = s.b.x - s.a.x;
    s.vy = s.b.y - s.a.y;
    s.sp = seg_prune<FACES>(t, s.a, s.b, s.vx, s.vy);
  }
  return s;
}

// Expected cells met: a box (w / c + 1) (h / c + 1); a segment 1.5 |b - a| / c + 1.  NaN
// and infinite queries are never split.
__device__ __forceinline__ bool worth_splitting(const SplitQuery& s, const SplitParams& p) {
  double e;
  if (s.kind == 1) e = ((s.box.xmax - s.box.xmin) / p.cell + 1.0) * ((s.box.ymax - s.box.ymin) / p.cell + 1.0);
  else e = 1.5 * sqrt(s.vx * s.vx + s.vy * s.vy) / p.cell + 1.0;
  return e > (double)p.min_cells && e < 1e300;
}

template <bool FACES>
__device__ __forceinline__ void children(const SplitQuery& s, const NodeLd& nd, bool* left, bool* right) {
  if (s.kind == 1) box_children(nd, s.box, left, right);
  else seg_children(nd, s.a, s.b, s.vx, s.vy, s.sp, left, right);
}

// Frontier entries: node index, with bit 30 set once the node is known to stay (a leaf, or a
// subtree of at most leaf_cells cells), so that later rounds do not load it again.
constexpr int kFinal = 1 << 30;
constexpr int kNodeMask = kFinal - 1;

// The frontier of query s (see the file comment); returns m (0: the query misses the tree
// bbox) and leaves the entries in *out (one of the two LDS buffers a and b, this lane's slots
// at stride kBlock).  Queries not worth splitting get the root alone, untested (their walk
// tests the tree bbox itself).
//
// Expansion runs in rounds: every entry not yet known to stay is loaded once; an internal node
// of more than leaf_cells cells is replaced by its viable children (in visit order: right,
// then left), written straight into the other buffer; a round that would leave more than
// max_parts entries is dropped (the frontier stays as it was before it).
template <bool FACES>
__device__ int frontier(const TreeView& t, const SplitQuery& s, const SplitParams& p, int* a, int* b, int** out) {
  a[0] = 0;
  *out = a;
  if (!worth_splitting(s, p)) return 1;
  if (s.kind == 1) {
    if (!boxes_intersect(s.box, t.bbox)) return 0;
  } else {
    P2 c0, d0;
    if (!cohen_sutherland(s.a, s.b, t.bbox, &c0, &d0)) return 0;
  }
  // The top of the walk: while one child alone is viable, a round would only replace the
  // single entry by that child; follow the chain directly (one load per level).
  int node = 0;
  while (true) {
    const NodeLd nd = ld_node(t.nodes, node);
    if (nd.child == -1 || nd.size <= p.leaf_cells) {
      a[0] = node | kFinal;
      return 1;
    }
    bool left, right;
    children<FACES>(s, nd, &left, &right);
    if (left && right) break;
    if (!left && !right) return 0;
    node = nd.child + (right ? 1 : 0);
  }
  a[0] = node;
  int* cur = a;
  int* nxt = b;
  int m = 1;
  while (true) {
    int w = 0;
    bool any = false;
    for (int k = 0; k < m; ++k) {
      const int e = cur[k * kBlock];
      int e0 = e | kFinal, e1 = -1;
      if (!(e & kFinal)) {
        const NodeLd nd = ld_node(t.nodes, e);
        if (nd.child != -1 && nd.size > p.leaf_cells) {
          bool left, right;
          children<FACES>(s, nd, &left, &right);
          any = true;
          e0 = right ? nd.child + 1 : (left ? nd.child : -1);
          e1 = (right && left) ? nd.child : -1;
        }
      }
      if (e0 >= 0) {
        if (w < p.max_parts) nxt[w * kBlock] = e0;
        ++w;
      }
      if (e1 >= 0) {
        if (w < p.max_parts) nxt[w * kBlock] = e1;
        ++w;
      }
    }
    if (w > p.max_parts) break;  // keep cur
    int* tmp = cur;
    cur = nxt;
    nxt = tmp;
    m = w;
    if (!any || m == 0) break;
  }
  *out = cur;
  return m;
}

// Count call (subs null): parts[i] = m and, when `keep` is set, the frontier nodes in
// keep[i * max_parts + k].  Write call: the subqueries, from `keep` when set, else computed
// again (the same frontier).
template <bool FACES>
__global__ void __launch_bounds__(kBlock) k_split(TreeView t, int kind, const double* __restrict__ qs,
                                                  const int32_t* __restrict__ order, int64_t n, SplitParams p,
                                                  int64_t* __restrict__ parts, const int64_t* __restrict__ part_offsets,
                                                  ct_subquery* __restrict__ subs, int32_t* __restrict__ keep) {
  // per lane: two frontier buffers of p.max_parts entries (LDS, lane-interleaved), sized at
  // launch: 2 * max_parts * kBlock ints (32 KiB at 16 parts, 5 workgroups per CU; 16 KiB at 8)
  extern __shared__ int lds[];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  if (subs != nullptr && keep != nullptr) {
    const int64_t o = part_offsets[i], m = part_offsets[i + 1] - o;
    for (int64_t k = 0; k < m; ++k) subs[o + k] = ct_subquery{(int32_t)q, keep[i * p.max_parts + k]};
    return;
  }
  int* a = lds + threadIdx.x;
  int* b = a + p.max_parts * kBlock;
  const SplitQuery s = load_split_query<FACES>(t, kind, qs, q);
  int* fr;
  const int m = frontier<FACES>(t, s, p, a, b, &fr);
  if (subs == nullptr) {
    parts[i] = m;
    if (keep != nullptr)
      for (int k = 0; k < m; ++k) keep[i * p.max_parts + k] = fr[k * kBlock] & kNodeMask;
    return;
  }
  const int64_t o = part_offsets[i];
  for (int k = 0; k < m; ++k) subs[o + k] = ct_subquery{(int32_t)q, (int32_t)(fr[k * kBlock] & kNodeMask)};
}

// Per processing position i: the subqueries [part_offsets[i], part_offsets[i + 1]) of query
// q = order[i]; their counts become exclusive prefixes within the query, and per_query[q]
// the query's total.
__global__ void __launch_bounds__(kBlock) k_sub_offsets(const int32_t* __restrict__ order, int64_t n,
                                                        const int64_t* __restrict__ part_offsets,
                                                        int64_t* __restrict__ counts, int64_t* __restrict__ per_query) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  int64_t run = 0;
  for (int64_t j = part_offsets[i]; j < part_offsets[i + 1]; ++j) {
    const int64_t c = counts[j];
    counts[j] = run;
    run += c;
  }
  per_query[q] = run;
}

// Box subquery j = (q, start): the walk of query.py:177-244 from `start`; count pass
// (counts[j], hit record) or write pass (rows at query_offsets[q] + prefix[j]).
template <bool WRITE, int STACK>
__global__ void __launch_bounds__(kBlock) k_boxes_sub(TreeView t, const double* __restrict__ boxes,
                                                      const ct_subquery* __restrict__ subs, int64_t n,
                                                      int64_t* __restrict__ counts,
                                                      const int64_t* __restrict__ query_offsets,
                                                      const int64_t* __restrict__ prefix, int64_t* __restrict__ ii,
                                                      int64_t* __restrict__ jj, int32_t* __restrict__ hits,
                                                      int slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ct_subquery sq = subs[i];
  const int64_t q = sq.query;
  int64_t o = WRITE ? query_offsets[q] + prefix[i] : 0;
  if (WRITE && hits != nullptr) {
    const int64_t total = (i + 1 < n && subs[i + 1].query == sq.query ? prefix[i + 1] : query_offsets[q + 1] -
                           query_offsets[q]) - prefix[i];
    if (total <= slots) return;  // written from the record by k_boxes_replay
  }
  const Box box = ld_box(boxes, q);
  int stack[STACK];
  int64_t cnt = 0;
  if (sq.node != 0 || boxes_intersect(box, t.bbox)) {
    int sp = 0;
    int node = sq.node;
    while (true) {
      NodeLd nd = ld_node(t.nodes, node);
      bool exhausted = false;
      while (nd.child != -1) {
        bool left, right;
        box_children(nd, box, &left, &right);
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
  if (!WRITE) counts[i] = cnt;
}

// Box rows of the subqueries whose hits all fit the record, copied from it.  One wavefront
// per workgroup: it reads its lanes' record slots coalesced (slot-major: slot h of lane j at
// hits[h * n + j]) into LDS, kReplayRows slots at a time, then writes the rows of
// 64 / kReplayRows lanes per step as contiguous runs of up to kReplayRows, instead of 64
// scattered 8-byte stores per instruction.
constexpr int kReplayRows = 16;
constexpr int kReplayLanes = 64 / kReplayRows;  // lanes whose rows one step writes

__global__ void __launch_bounds__(64) k_boxes_replay(const ct_subquery* __restrict__ subs, int64_t n,
                                                     const int64_t* __restrict__ query_offsets,
                                                     const int64_t* __restrict__ prefix, const int32_t* __restrict__ hits,
                                                     int slots, int64_t* __restrict__ ii, int64_t* __restrict__ jj) {
  __shared__ int32_t buf[kReplayRows][64];
  const int lane = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  int64_t q = 0, o = 0, total = 0;
  if (i < n) {
    q = subs[i].query;
    o = query_offsets[q] + prefix[i];
    total = (i + 1 < n && subs[i + 1].query == q ? prefix[i + 1] : query_offsets[q + 1] - query_offsets[q]) - prefix[i];
    if (total > slots) total = 0;  // walked again by k_boxes_sub
  }
  int64_t H = total;
  for (int off = 32; off > 0; off >>= 1) {
    const int64_t other = __shfl_xor(H, off);
    H = other > H ? other : H;
  }
  for (int64_t h0 = 0; h0 < H; h0 += kReplayRows) {
    for (int c = 0; c < kReplayRows; ++c) {
      const int64_t h = h0 + c;
      buf[c][lane] = h < total ? hits[h * n + i] : 0;
    }
    __syncthreads();
    // kReplayLanes lanes' rows per step: lanes [p * kReplayRows, (p + 1) * kReplayRows)
    // write lane l + p's
    const int part = lane / kReplayRows, r = lane % kReplayRows;
    const int64_t h = h0 + r;
    for (int l = 0; l < 64; l += kReplayLanes) {
      const int64_t tl = __shfl(total, l + part);
      const int64_t ol = __shfl(o, l + part), ql = __shfl(q, l + part);
      if (h < tl) {
        jj[ol + h] = buf[r][l + part];
        ii[ol + h] = ql;
      }
    }
    __syncthreads();
  }
}

// Segment subquery i = (q, start): edge_query from `start`, count pass (counts[i], and the
// first hits in the record when there is one).
template <int MAXV, int STACK, bool FACES>
__global__ void __launch_bounds__(kBlock) k_edges_count(TreeView t, const double* __restrict__ segs,
                                                        const ct_subquery* __restrict__ subs, int64_t n,
                                                        int64_t* __restrict__ counts, int32_t* __restrict__ hits,
                                                        int slots, double4* __restrict__ hit_xy, int xy_slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ct_subquery sq = subs[i];
  const double4 s4 = reinterpret_cast<const double4*>(segs)[sq.query];
  CT_LANE_STACK(STACK, stack);
  HitRecord rec;
  if (hits != nullptr) rec = HitRecord{hits + i, slots, n, hit_xy != nullptr ? hit_xy + i : nullptr, xy_slots};
  counts[i] = edge_query<false, MAXV, STACK, FACES>(t, {s4.x, s4.y}, {s4.z, s4.w}, sq.query, 0, FinalRows{}, stack,
                                                    rec, sq.node);
}

// Write pass into the staging rows (processing order: subquery i's rows at sub_offsets[i] + h)
// for the subqueries whose rows are not all in the record with their clipped segments: cells
// recorded are clipped again, others walk again.  The rest are copied by k_edges_rows.
template <int MAXV, int STACK, bool FACES>
__global__ void __launch_bounds__(kBlock) k_edges_stage(TreeView t, const double* __restrict__ segs,
                                                        const ct_subquery* __restrict__ subs, int64_t n,
                                                        const int64_t* __restrict__ sub_offsets,
                                                        const int32_t* __restrict__ hits, int slots, int xy_slots,
                                                        StagedRows rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = sub_offsets[i], total = sub_offsets[i + 1] - o;
  if (total == 0 || (hits != nullptr && total <= xy_slots)) return;
  const ct_subquery sq = subs[i];
  const int64_t q = sq.query;
  const double4 s4 = reinterpret_cast<const double4*>(segs)[q];
  const P2 a = {s4.x, s4.y}, b = {s4.z, s4.w};
  rows.seg = s4;
  if (hits != nullptr && total <= slots) {
    for (int64_t h = 0; h < total; ++h) {
      const int c = hits[h * n + i];
      P2 cc, dd;
      if (FACES) edge_face<MAXV>(t, c, a, b, &cc, &dd);
      else edge_edge(t, c, a, b, &cc, &dd);
      rows.put(o + h, q, c, cc, dd);
    }
    return;
  }
  CT_LANE_STACK(STACK, stack);
  edge_query<true, MAXV, STACK, FACES>(t, a, b, q, o, rows, stack, HitRecord{}, sq.node);
}

// The staging rows of the subqueries whose hits all sit in the record with their clipped
// segments (sub_offsets[i + 1] - sub_offsets[i] <= xy_slots), copied by whole wavefronts: the
// 64 subqueries of a wavefront own one run of consecutive staging rows, written 64 rows per
// step (row g of the run by lane g mod 64, its owner found by a binary search over the
// wavefront's exclusive prefix), instead of each lane storing its own rows one by one.
__global__ void __launch_bounds__(kBlock) k_edges_rows(const double* __restrict__ segs,
                                                       const ct_subquery* __restrict__ subs, int64_t n,
                                                       const int64_t* __restrict__ sub_offsets,
                                                       const int32_t* __restrict__ hits,
                                                       const double4* __restrict__ hit_xy, int xy_slots,
                                                       StagedRows rows) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave0 = i - lane;
  int m = 0, q = 0;
  long long o = 0;
  double4 s4 = make_double4(0.0, 0.0, 0.0, 0.0);
  if (i < n) {
    o = sub_offsets[i];
    const int64_t total = sub_offsets[i + 1] - o;
    if (total <= xy_slots) m = (int)total;
    if (m > 0) {
      q = subs[i].query;
      s4 = reinterpret_cast<const double4*>(segs)[q];
    }
  }
  int incl = m;
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  const int excl = incl - m;
  const int total = __shfl(incl, 63);
  for (int base = 0; base < total; base += 64) {
    const int g = base + lane;
    int own = 0;
    for (int step = 32; step > 0; step >>= 1) {
      const int e = __shfl(excl, own + step);
      if (e <= g) own += step;
    }
    const int h = g - __shfl(excl, own);
    const long long r = __shfl(o, own) + h;
    const int oq = __shfl(q, own);
    const double4 seg = make_double4(__shfl(s4.x, own), __shfl(s4.y, own), __shfl(s4.z, own), __shfl(s4.w, own));
    if (g < total) {
      const int64_t src = (int64_t)h * n + wave0 + own;
      const double4 xy = hit_xy[src];
      rows.query[r] = oq;
      rows.cell[r] = hits[src];
      rows.xy[r] = xy;
      rows.t[r] = row_t(seg, make_double2(xy.x, xy.y));
    }
  }
}

// Per processing position i (query q = order[i], subqueries [part_offsets[i],
// part_offsets[i + 1])): the query's row count and its first staging row.
__global__ void __launch_bounds__(kBlock) k_query_spans(const int32_t* __restrict__ order, int64_t n,
                                                        const int64_t* __restrict__ part_offsets,
                                                        const int64_t* __restrict__ sub_offsets,
                                                        int64_t* __restrict__ per_query,
                                                        int64_t* __restrict__ query_start) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  const int64_t lo = sub_offsets[part_offsets[i]], hi = sub_offsets[part_offsets[i + 1]];
  per_query[q] = hi - lo;
  query_start[q] = lo;
}

// geometry_utils.py:564-574 on the staging rows: each row's rank by t within its query's
// staging run [query_start[q], + rows of q) (ranked as sort_tile does, from an LDS window of
// t), written to output row offsets[q] + rank.  A tile's rows of one query land in one output
// block, so output lines are written whole.
__global__ void __launch_bounds__(kBlock) k_sort_staged(const int64_t* __restrict__ offsets,
                                                        const int64_t* __restrict__ query_start,
                                                        const int32_t* __restrict__ rq,
                                                        const int32_t* __restrict__ rcell,
                                                        const double4* __restrict__ rxy,
                                                        const double* __restrict__ rt, int64_t n_rows,
                                                        int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                                                        double4* __restrict__ xy) {
  __shared__ double tw[kRankWindow];
  const int64_t r0 = (int64_t)blockIdx.x * kRankTile;
  const int64_t w0 = r0 > kRankHalo ? r0 - kRankHalo : 0;
  const int64_t w1 = r0 + kRankTile + kRankHalo < n_rows ? r0 + kRankTile + kRankHalo : n_rows;
  for (int64_t r = w0 + threadIdx.x; r < w1; r += blockDim.x) tw[r - w0] = rt[r];
  __syncthreads();
  for (int k = threadIdx.x; k < kRankTile; k += blockDim.x) {
    const int64_t r = r0 + k;
    if (r >= n_rows) break;
    const int64_t q = rq[r];
    const int64_t out0 = offsets[q];
    const int64_t lo = query_start[q], hi = lo + (offsets[q + 1] - out0);
    const double t = tw[r - w0];
    int64_t rank = 0;
    if (lo >= w0 && hi <= w1) {
      for (int64_t m = lo; m < r; ++m) rank += t_less(t, tw[m - w0]) ? 0 : 1;
      for (int64_t m = r + 1; m < hi; ++m) rank += t_less(tw[m - w0], t) ? 1 : 0;
    } else {
      for (int64_t m = lo; m < r; ++m) rank += t_less(t, rt[m]) ? 0 : 1;
      for (int64_t m = r + 1; m < hi; ++m) rank += t_less(rt[m], t) ? 1 : 0;
    }
    ii[out0 + rank] = q;
    jj[out0 + rank] = rcell[r];
    xy[out0 + rank] = rxy[r];
  }
}

// Segment-face count pass with a record, first half: the walk, counting and recording the
// candidates (cells whose bbox Cohen-Sutherland accepts, by leaf position) of subquery i.
template <int MAXV, int STACK>
__global__ void __launch_bounds__(kBlock) k_edges_cand(TreeView t, const double* __restrict__ segs,
                                                       const ct_subquery* __restrict__ subs, int64_t n,
                                                       int64_t* __restrict__ counts, int32_t* __restrict__ hits,
                                                       int slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ct_subquery sq = subs[i];
  const double4 s4 = reinterpret_cast<const double4*>(segs)[sq.query];
  CT_LANE_STACK(STACK, stack);
  counts[i] = edge_query<false, MAXV, STACK, true, true>(t, {s4.x, s4.y}, {s4.z, s4.w}, sq.query, 0, FinalRows{},
                                                         stack, HitRecord{hits + i, slots, n}, sq.node);
}

// Segment-face count pass, second half: subquery i's candidate leaf positions (the first
// counts[i] slots of the record) are clipped in order; the hits are compacted in place into
// cell indices (and clipped segments in the first xy_slots slots), counts[i] becomes the hit
// count.  A subquery with more candidates than slots was recorded in part and is left as it
// is (counts[i] > slots) for k_edges_overflow.
//
// Each wavefront clips its 64 subqueries' candidates together, 64 at a time: candidate g of
// the wavefront's concatenated lists (lane by lane) goes to lane g mod 64, so that the clips
// and their dependent loads run side by side instead of one list per lane.  A hit's slot is
// its owner's hits so far plus the hits among the owner's earlier candidates of the same
// batch (a ballot).  The lists are compacted in place: a hit's slot is at most its own
// candidate's slot, and every candidate of a batch is read before any hit of it is written.
template <int MAXV, bool ROWS>
__global__ void __launch_bounds__(kBlock) k_edges_clip(TreeView t, const double* __restrict__ segs,
                                                       const ct_subquery* __restrict__ subs, int64_t n,
                                                       int64_t* __restrict__ counts, int32_t* __restrict__ hits,
                                                       int slots, double4* __restrict__ hit_xy, int xy_slots) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave0 = i - lane;
  int m = 0;
  bool mine = false;  // this lane's counts[i] is finished here
  double4 s4 = make_double4(0.0, 0.0, 0.0, 0.0);
  if (i < n) {
    const int64_t mm = counts[i];
    mine = mm <= slots;
    if (mine) m = (int)mm;
    if (m > 0) s4 = reinterpret_cast<const double4*>(segs)[subs[i].query];
  }
  int incl = m;  // inclusive prefix of m over the wavefront
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  const int excl = incl - m;
  const int total = __shfl(incl, 63);
  int found = 0;  // hits of this lane's subquery so far
  for (int base = 0; base < total; base += 64) {
    const int g = base + lane;
    // owner o: the last lane with excl[o] <= g (lanes without candidates share the next
    // lane's excl and come before it)
    int o = 0;
    for (int step = 32; step > 0; step >>= 1) {
      const int e = __shfl(excl, o + step);
      if (e <= g) o += step;
    }
    const int oexcl = __shfl(excl, o);
    const int ofound = __shfl(found, o);
    const P2 a = {__shfl(s4.x, o), __shfl(s4.y, o)};
    const P2 b = {__shfl(s4.z, o), __shfl(s4.w, o)};
    const int64_t oi = wave0 + o;
    bool hit = false;
    int c = 0;
    P2 cc, dd;
    if (g < total) {
      const int k = hits[(int64_t)(g - oexcl) * n + oi];
      c = t.bb_indices[k];
      if constexpr (ROWS) hit = edge_face_clip_row<MAXV>(t, k, c, a, b, &cc, &dd);
      else hit = edge_face_clip<MAXV>(t, c, leaf_box(t, k, c), a, b, &cc, &dd);
    }
    const uint64_t mask = __ballot(hit);
    const int first = oexcl > base ? oexcl - base : 0;  // batch lane of the owner's first candidate
    const int local = __popcll(mask & ((1ull << lane) - 1)) - __popcll(mask & ((1ull << first) - 1));
    if (hit) {
      const int slot = ofound + local;
      hits[(int64_t)slot * n + oi] = c;
      if (slot < xy_slots) hit_xy[(int64_t)slot * n + oi] = make_double4(cc.x, cc.y, dd.x, dd.y);
    }
    const int lo = excl > base ? excl - base : 0;
    const int hi = excl + m - base < 64 ? excl + m - base : 64;
    if (hi > lo) {
      const uint64_t upto = hi == 64 ? ~0ull : ((1ull << hi) - 1);
      found += __popcll(mask & upto & ~((1ull << lo) - 1));
    }
  }
  if (mine) counts[i] = found;
}

// The subqueries k_edges_clip left (more candidates than slots) walk again with the whole test,
// recording their first hits.
template <int MAXV, int STACK>
__global__ void __launch_bounds__(kBlock) k_edges_overflow(TreeView t, const double* __restrict__ segs,
                                                           const ct_subquery* __restrict__ subs, int64_t n,
                                                           int64_t* __restrict__ counts, int32_t* __restrict__ hits,
                                                           int slots, double4* __restrict__ hit_xy, int xy_slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || counts[i] <= slots) return;
  const ct_subquery sq = subs[i];
  const double4 s4 = reinterpret_cast<const double4*>(segs)[sq.query];
  CT_LANE_STACK(STACK, stack);
  const HitRecord rec{hits + i, slots, n, hit_xy != nullptr ? hit_xy + i : nullptr, xy_slots};
  counts[i] = edge_query<false, MAXV, STACK, true>(t, {s4.x, s4.y}, {s4.z, s4.w}, sq.query, 0, FinalRows{}, stack,
                                                   rec, sq.node);
}

template <bool WRITE>
struct BoxesSubLaunch {
  const TreeView& t;
  const double* boxes;
  const ct_subquery* subs;
  int64_t n;
  int64_t* counts;
  const int64_t* qoff;
  const int64_t* prefix;
  int64_t* ii;
  int64_t* jj;
  int32_t* hits;
  int slots;
  hipStream_t st;
  template <int S>
  int run() {
    k_boxes_sub<WRITE, S><<<grid_for(n), kBlock, 0, st>>>(t, boxes, subs, n, counts, qoff, prefix, ii, jj, hits, slots);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

// Count pass (WRITE false): candidates then their clip for segment-face queries with a
// record, else the single walk.  Write pass: the staging rows (k_edges_rows for subqueries
// wholly in the record, k_edges_stage for the others).
template <bool WRITE, int MAXV, bool FACES>
struct EdgesSubLaunch {
  const TreeView& t;
  const double* segs;
  const ct_subquery* subs;
  int64_t n;
  int64_t* counts;             // count pass
  const int64_t* sub_offsets;  // write pass
  int32_t* hits;
  int slots;
  double4* hit_xy;
  int xy_slots;
  StagedRows rows;  // write pass
  hipStream_t st;
  template <int S>
  int run() {
    const unsigned g = grid_for(n);
    if constexpr (WRITE) {
      if (hits != nullptr && hit_xy != nullptr) {
        k_edges_rows<<<g, kBlock, 0, st>>>(segs, subs, n, sub_offsets, hits, hit_xy, xy_slots, rows);
        CT_LAUNCH_CHECK();
      }
      k_edges_stage<MAXV, S, FACES><<<g, kBlock, 0, st>>>(t, segs, subs, n, sub_offsets, hits, slots,
                                                           hit_xy != nullptr ? xy_slots : 0, rows);
      CT_LAUNCH_CHECK();
      return CT_OK;
    } else {
      if (FACES && hits != nullptr) {
        k_edges_cand<MAXV, S><<<g, kBlock, 0, st>>>(t, segs, subs, n, counts, hits, slots);
        CT_LAUNCH_CHECK();
        if (t.leaf != nullptr)
          k_edges_clip<MAXV, true><<<g, kBlock, 0, st>>>(t, segs, subs, n, counts, hits, slots, hit_xy, xy_slots);
        else
          k_edges_clip<MAXV, false><<<g, kBlock, 0, st>>>(t, segs, subs, n, counts, hits, slots, hit_xy, xy_slots);
        CT_LAUNCH_CHECK();
        k_edges_overflow<MAXV, S><<<g, kBlock, 0, st>>>(t, segs, subs, n, counts, hits, slots, hit_xy, xy_slots);
        CT_LAUNCH_CHECK();
        return CT_OK;
      }
      k_edges_count<MAXV, S, FACES><<<g, kBlock, 0, st>>>(t, segs, subs, n, counts, hits, slots, hit_xy, xy_slots);
      CT_LAUNCH_CHECK();
      return CT_OK;
    }
  }
};

template <bool WRITE>
int launch_edges_sub(const ct_tree* tree, const double* segs, const ct_subquery* subs, int64_t n, int kind,
                     int64_t* counts, const int64_t* sub_offsets, int32_t* hits, int slots, double4* hit_xy,
                     int xy_slots, StagedRows rows, hipStream_t st) {
  const TreeView t = view(tree);
  const int d = tree->depth, nm = tree->n_max_vert;
  if (kind == 0)
    return with_stack(d, EdgesSubLaunch<WRITE, 4, false>{t, segs, subs, n, counts, sub_offsets, hits, slots, hit_xy,
                                                         xy_slots, rows, st});
  if (nm <= 4)
    return with_stack(d, EdgesSubLaunch<WRITE, 4, true>{t, segs, subs, n, counts, sub_offsets, hits, slots, hit_xy,
                                                        xy_slots, rows, st});
  if (nm <= 8)
    return with_stack(d, EdgesSubLaunch<WRITE, 8, true>{t, segs, subs, n, counts, sub_offsets, hits, slots, hit_xy,
                                                        xy_slots, rows, st});
  if (nm <= 16)
    return with_stack(d, EdgesSubLaunch<WRITE, 16, true>{t, segs, subs, n, counts, sub_offsets, hits, slots, hit_xy,
                                                         xy_slots, rows, st});
  return with_stack(d, EdgesSubLaunch<WRITE, 32, true>{t, segs, subs, n, counts, sub_offsets, hits, slots, hit_xy,
                                                       xy_slots, rows, st});
}

int check_split_tree(const ct_tree* tree) {
  CT_REQUIRE(tree != nullptr && tree->nodes && tree->bb_indices && tree->elements && tree->vertices,
             "incomplete tree");
  CT_REQUIRE(tree->depth <= CT_MAX_DEPTH, "split queries take trees of depth <= %d", CT_MAX_DEPTH);
  return CT_OK;
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_split_queries(const ct_tree* tree, int32_t kind, const double* queries, int64_t n, const int32_t* order,
                     int32_t min_cells, int32_t leaf_cells, int32_t max_parts, int64_t* parts,
                     const int64_t* part_offsets, ct_subquery* subs, int32_t* frontier_buf, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(kind >= 1 && kind <= 3 && n >= 0, "kind must be 1 (boxes), 2 (segment-face) or 3 (segment-edge)");
  CT_REQUIRE(max_parts >= 1 && max_parts <= kMaxParts, "max_parts must be in [1, %d]", kMaxParts);
  CT_REQUIRE(leaf_cells >= 0 && min_cells >= 0, "negative split sizes");
  CT_REQUIRE(tree->n_nodes < ((int64_t)1 << 28), "split queries take trees of fewer than 2**28 nodes");
  CT_REQUIRE(subs == nullptr ? parts != nullptr : part_offsets != nullptr, "parts (count) or part_offsets + subs");
  if (n == 0) return CT_OK;
  const double ex = tree->bbox[1] - tree->bbox[0], ey = tree->bbox[3] - tree->bbox[2];
  double cell = sqrt(ex * ey / (double)tree->n_elements);
  if (!(cell > 0.0)) cell = 1.0;  // degenerate bbox: sizes compare against 1 (splitting never changes results)
  const SplitParams p{cell, min_cells, leaf_cells, max_parts};
  const TreeView t = view(tree);
  hipStream_t st = as_stream(stream);
  const int k = kind == 1 ? 1 : 2;  // box or segment
  const size_t lds = 2 * (size_t)max_parts * kBlock * sizeof(int);
  if (kind == 2)
    k_split<true><<<grid_for(n), kBlock, lds, st>>>(t, k, queries, order, n, p, parts, part_offsets, subs,
                                                    frontier_buf);
  else
    k_split<false><<<grid_for(n), kBlock, lds, st>>>(t, k, queries, order, n, p, parts, part_offsets, subs,
                                                     frontier_buf);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_subquery_offsets(const int32_t* order, int64_t n, const int64_t* part_offsets, int64_t* counts,
                        int64_t* per_query, void* stream) {
  CT_REQUIRE(n >= 0 && part_offsets && counts && per_query, "bad arguments");
  if (n == 0) return CT_OK;
  k_sub_offsets<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(order, n, part_offsets, counts, per_query);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_locate_boxes_count_sub(const ct_tree* tree, const double* boxes, const ct_subquery* subs, int64_t n_sub,
                              int64_t* counts, int32_t* hits, int32_t hit_slots, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n_sub >= 0 && counts && (n_sub == 0 || subs), "bad arguments");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n_sub == 0) return CT_OK;
  const TreeView t = view(tree);
  return with_stack(tree->depth, BoxesSubLaunch<false>{t, boxes, subs, n_sub, counts, nullptr, nullptr, nullptr, nullptr,
                                                       hits, hits ? hit_slots : 0, as_stream(stream)});
}

int ct_locate_boxes_write_sub(const ct_tree* tree, const double* boxes, const ct_subquery* subs, int64_t n_sub,
                              const int64_t* query_offsets, const int64_t* prefix, const int32_t* hits,
                              int32_t hit_slots, int64_t* ii, int64_t* jj, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n_sub >= 0 && (n_sub == 0 || (subs && query_offsets && prefix)), "bad arguments");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n_sub == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  if (hits != nullptr) {
    k_boxes_replay<<<(unsigned)((n_sub + 63) / 64), 64, 0, st>>>(subs, n_sub, query_offsets, prefix, hits, hit_slots,
                                                                 ii, jj);
    CT_LAUNCH_CHECK();
  }
  const TreeView t = view(tree);
  return with_stack(tree->depth, BoxesSubLaunch<true>{t, boxes, subs, n_sub, nullptr, query_offsets, prefix, ii, jj,
                                                      const_cast<int32_t*>(hits), hits ? hit_slots : 0, st});
}

int ct_locate_edges_count_sub(const ct_tree* tree, const double* segments, const ct_subquery* subs, int64_t n_sub,
                              int32_t kind, int64_t* counts, int32_t* hits, int32_t hit_slots, double* hit_xy,
                              int32_t xy_slots, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n_sub >= 0 && counts && (n_sub == 0 || subs) && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 ? tree->n_max_vert >= 3 : tree->n_max_vert == 2, "kind does not match the tree");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  CT_REQUIRE(hit_xy == nullptr || (hits != nullptr && xy_slots >= 1 && xy_slots <= hit_slots),
             "hit_xy needs hits and xy_slots in [1, hit_slots]");
  if (n_sub == 0) return CT_OK;
  return launch_edges_sub<false>(tree, segments, subs, n_sub, kind, counts, nullptr, hits, hits ? hit_slots : 0,
                                 reinterpret_cast<double4*>(hit_xy), hit_xy ? xy_slots : 0, StagedRows{},
                                 as_stream(stream));
}

int ct_subquery_spans(const int32_t* order, int64_t n, const int64_t* part_offsets, const int64_t* sub_offsets,
                      int64_t* per_query, int64_t* query_start, void* stream) {
  CT_REQUIRE(n >= 0 && (n == 0 || (part_offsets && sub_offsets && per_query && query_start)), "bad arguments");
  if (n == 0) return CT_OK;
  k_query_spans<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(order, n, part_offsets, sub_offsets, per_query,
                                                               query_start);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_locate_edges_write_sub(const ct_tree* tree, const double* segments, const ct_subquery* subs, int64_t n_sub,
                              int32_t kind, const int64_t* sub_offsets, const int32_t* hits, int32_t hit_slots,
                              const double* hit_xy, int32_t xy_slots, int32_t* rows_query, int32_t* rows_cell,
                              double* rows_xy, double* rows_t, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n_sub >= 0 && (n_sub == 0 || (subs && sub_offsets)) && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 ? tree->n_max_vert >= 3 : tree->n_max_vert == 2, "kind does not match the tree");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  CT_REQUIRE(hit_xy == nullptr || (hits != nullptr && xy_slots >= 1 && xy_slots <= hit_slots),
             "hit_xy needs hits and xy_slots in [1, hit_slots]");
  CT_REQUIRE(n_sub == 0 || (rows_query && rows_cell && rows_xy && rows_t), "null staging rows");
  if (n_sub == 0) return CT_OK;
  const StagedRows rows{rows_query, rows_cell, reinterpret_cast<double4*>(rows_xy), rows_t, {}};
  return launch_edges_sub<true>(tree, segments, subs, n_sub, kind, nullptr, sub_offsets, const_cast<int32_t*>(hits),
                                hits ? hit_slots : 0, reinterpret_cast<double4*>(const_cast<double*>(hit_xy)),
                                hit_xy ? xy_slots : 0, rows, as_stream(stream));
}

int ct_sort_staged_rows(const int64_t* offsets, const int64_t* query_start, const int32_t* rows_query,
                        const int32_t* rows_cell, const double* rows_xy, const double* rows_t, int64_t n_rows,
                        int64_t* ii, int64_t* jj, double* xy, void* stream) {
  CT_REQUIRE(n_rows >= 0, "n_rows must be >= 0");
  if (n_rows == 0) return CT_OK;
  CT_REQUIRE(offsets && query_start && rows_query && rows_cell && rows_xy && rows_t && ii && jj && xy,
             "bad arguments");
  k_sort_staged<<<grid_for(n_rows, kRankTile), kBlock, 0, as_stream(stream)>>>(
      offsets, query_start, rows_query, rows_cell, reinterpret_cast<const double4*>(rows_xy), rows_t, n_rows, ii, jj,
      reinterpret_cast<double4*>(xy));
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // extern "C"
