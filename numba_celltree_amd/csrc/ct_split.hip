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
//   *_count_sub                counts[j] (+ hit record per subquery)
//   ct_subquery_offsets        counts[j] -> the exclusive prefix within its query;
//                              per_query[q] = the query's total (then ct_exclusive_scan)
//   *_write_sub                rows of subquery j at query_offsets[q] + prefix[j]
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
  double vx, vy, mseg;
  const double4* nbox;
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
    s.nbox = seg_prune_boxes<FACES>(t, s.a, s.b, s.vx, s.vy, &s.mseg);
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
  else seg_children(nd, s.a, s.b, s.vx, s.vy, s.mseg, s.nbox, left, right);
}

// Frontier entries: node index, with bit 31 set when the subtree is worth expanding (an
// internal node holding more than leaf_cells cells).
constexpr int kExpand = 1 << 31;
constexpr int kNodeMask = 0x7fffffff;

__device__ __forceinline__ int frontier_entry(const ct_node* __restrict__ nodes, int node, int leaf_cells) {
  const NodeLd nd = ld_node(nodes, node);
  return node | ((nd.child != -1 && nd.size > leaf_cells) ? kExpand : 0);
}

// The frontier of query s (see the file comment); returns m (0: the query misses the tree
// bbox) and leaves the entries in *out (one of the two LDS buffers a and b, this lane's slots
// at stride kBlock).  Queries not worth splitting get the root alone, untested (their walk
// tests the tree bbox itself).
//
// Expansion runs in rounds: every expandable entry is replaced by its viable children (in
// visit order: right, then left) at once, so a round's node loads are independent of each
// other; a round that would exceed max_parts entries is not taken.  The second pass of a
// round tests an expanded entry's children again (its node record is in the cache by then)
// rather than keeping the first pass's answer in LDS.
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
    if (nd.child == -1 || nd.size <= p.leaf_cells) break;
    bool left, right;
    children<FACES>(s, nd, &left, &right);
    if (left && right) break;
    if (!left && !right) return 0;
    node = nd.child + (right ? 1 : 0);
  }
  a[0] = frontier_entry(t.nodes, node, p.leaf_cells);
  int* cur = a;
  int* nxt = b;
  int m = 1;
  while (true) {
    int m2 = 0;
    bool any = false;
    for (int k = 0; k < m; ++k) {
      const int e = cur[k * kBlock];
      int cnt = 1;
      if (e & kExpand) {
        const NodeLd nd = ld_node(t.nodes, e & kNodeMask);
        bool left, right;
        children<FACES>(s, nd, &left, &right);
        cnt = (left ? 1 : 0) + (right ? 1 : 0);
        any = true;
      }
      m2 += cnt;
    }
    if (!any || m2 > p.max_parts) break;
    int w = 0;
    for (int k = 0; k < m; ++k) {
      const int e = cur[k * kBlock];
      if (!(e & kExpand)) {
        nxt[w++ * kBlock] = e;
      } else {
        const NodeLd nd = ld_node(t.nodes, e & kNodeMask);
        bool left, right;
        children<FACES>(s, nd, &left, &right);
        if (right) nxt[w++ * kBlock] = frontier_entry(t.nodes, nd.child + 1, p.leaf_cells);
        if (left) nxt[w++ * kBlock] = frontier_entry(t.nodes, nd.child, p.leaf_cells);
      }
    }
    int* tmp = cur;
    cur = nxt;
    nxt = tmp;
    m = m2;
    if (m == 0) break;
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
  // per lane: two frontier buffers (LDS, lane-interleaved)
  __shared__ int lds[2 * kMaxParts * kBlock];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = query_index(order, i);
  if (subs != nullptr && keep != nullptr) {
    const int64_t o = part_offsets[i], m = part_offsets[i + 1] - o;
    for (int64_t k = 0; k < m; ++k) subs[o + k] = ct_subquery{(int32_t)q, keep[i * p.max_parts + k]};
    return;
  }
  int* a = lds + threadIdx.x;
  int* b = a + kMaxParts * kBlock;
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
// hits[h * n + j]) into LDS, 32 slots at a time, then writes the rows of two lanes per step as
// two contiguous runs of up to 32, instead of 64 scattered 8-byte stores per instruction.
constexpr int kReplayRows = 32;

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
    // two lanes' rows per step: lanes 0-31 write lane l's, lanes 32-63 lane l + 1's
    const int half = lane >> 5, r = lane & 31;
    const int64_t h = h0 + r;
    for (int l = 0; l < 64; l += 2) {
      const int64_t tl = __shfl(total, l + half);
      const int64_t ol = __shfl(o, l + half), ql = __shfl(q, l + half);
      if (h < tl) {
        jj[ol + h] = buf[r][l + half];
        ii[ol + h] = ql;
      }
    }
    __syncthreads();
  }
}

// Segment subquery j = (q, start): edge_query from `start`; count or write pass.
template <bool WRITE, int MAXV, int STACK, bool FACES>
__global__ void __launch_bounds__(kBlock) k_edges_sub(TreeView t, const double* __restrict__ segs,
                                                      const ct_subquery* __restrict__ subs, int64_t n,
                                                      int64_t* __restrict__ counts,
                                                      const int64_t* __restrict__ query_offsets,
                                                      const int64_t* __restrict__ prefix, int64_t* __restrict__ ii,
                                                      int64_t* __restrict__ jj, double* __restrict__ xy,
                                                      int32_t* __restrict__ hits, int slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ct_subquery sq = subs[i];
  const int64_t q = sq.query;
  const double4 s4 = reinterpret_cast<const double4*>(segs)[q];
  const P2 a = {s4.x, s4.y}, b = {s4.z, s4.w};
  CT_LANE_STACK(STACK, stack);
  if (WRITE) {
    const int64_t o = query_offsets[q] + prefix[i];
    const int64_t total = (i + 1 < n && subs[i + 1].query == sq.query ? prefix[i + 1] : query_offsets[q + 1] -
                           query_offsets[q]) - prefix[i];
    if (hits != nullptr && total <= slots) {
      for (int64_t h = 0; h < total; ++h) {
        const int c = hits[h * n + i];
        P2 cc, dd;
        if (FACES) edge_face<MAXV>(t, c, a, b, &cc, &dd);
        else edge_edge(t, c, a, b, &cc, &dd);
        ii[o + h] = q;
        jj[o + h] = c;
        reinterpret_cast<double4*>(xy)[o + h] = make_double4(cc.x, cc.y, dd.x, dd.y);
      }
    } else {
      edge_query<true, MAXV, STACK, FACES>(t, a, b, q, o, ii, jj, xy, stack, HitRecord{}, sq.node);
    }
  } else {
    HitRecord rec;
    if (hits != nullptr) rec = HitRecord{hits + i, slots, n};
    counts[i] = edge_query<false, MAXV, STACK, FACES>(t, a, b, q, 0, nullptr, nullptr, nullptr, stack, rec, sq.node);
  }
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

template <bool WRITE, int MAXV, bool FACES>
struct EdgesSubLaunch {
  const TreeView& t;
  const double* segs;
  const ct_subquery* subs;
  int64_t n;
  int64_t* counts;
  const int64_t* qoff;
  const int64_t* prefix;
  int64_t* ii;
  int64_t* jj;
  double* xy;
  int32_t* hits;
  int slots;
  hipStream_t st;
  template <int S>
  int run() {
    k_edges_sub<WRITE, MAXV, S, FACES><<<grid_for(n), kBlock, 0, st>>>(t, segs, subs, n, counts, qoff, prefix, ii, jj,
                                                                        xy, hits, slots);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <bool WRITE>
int launch_edges_sub(const ct_tree* tree, const double* segs, const ct_subquery* subs, int64_t n, int kind,
                     int64_t* counts, const int64_t* qoff, const int64_t* prefix, int64_t* ii, int64_t* jj,
                     double* xy, int32_t* hits, int slots, hipStream_t st) {
  const TreeView t = view(tree);
  const int d = tree->depth, nm = tree->n_max_vert;
  if (kind == 0)
    return with_stack(d, EdgesSubLaunch<WRITE, 4, false>{t, segs, subs, n, counts, qoff, prefix, ii, jj, xy, hits,
                                                         slots, st});
  if (nm <= 4)
    return with_stack(d, EdgesSubLaunch<WRITE, 4, true>{t, segs, subs, n, counts, qoff, prefix, ii, jj, xy, hits,
                                                        slots, st});
  if (nm <= 8)
    return with_stack(d, EdgesSubLaunch<WRITE, 8, true>{t, segs, subs, n, counts, qoff, prefix, ii, jj, xy, hits,
                                                        slots, st});
  if (nm <= 16)
    return with_stack(d, EdgesSubLaunch<WRITE, 16, true>{t, segs, subs, n, counts, qoff, prefix, ii, jj, xy, hits,
                                                         slots, st});
  return with_stack(d, EdgesSubLaunch<WRITE, 32, true>{t, segs, subs, n, counts, qoff, prefix, ii, jj, xy, hits,
                                                       slots, st});
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
  if (kind == 2)
    k_split<true><<<grid_for(n), kBlock, 0, st>>>(t, k, queries, order, n, p, parts, part_offsets, subs, frontier_buf);
  else
    k_split<false><<<grid_for(n), kBlock, 0, st>>>(t, k, queries, order, n, p, parts, part_offsets, subs, frontier_buf);
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
                              int32_t kind, int64_t* counts, int32_t* hits, int32_t hit_slots, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n_sub >= 0 && counts && (n_sub == 0 || subs) && (kind == 0 || kind == 1), "bad arguments");
  CT_REQUIRE(kind == 1 ? tree->n_max_vert >= 3 : tree->n_max_vert == 2, "kind does not match the tree");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n_sub == 0) return CT_OK;
  return launch_edges_sub<false>(tree, segments, subs, n_sub, kind, counts, nullptr, nullptr, nullptr, nullptr,
                                 nullptr, hits, hits ? hit_slots : 0, as_stream(stream));
}

int ct_locate_edges_write_sub(const ct_tree* tree, const double* segments, const ct_subquery* subs, int64_t n_sub,
                              int32_t kind, const int64_t* query_offsets, const int64_t* prefix, const int32_t* hits,
                              int32_t hit_slots, int64_t* ii, int64_t* jj, double* xy, void* stream) {
  int rc = check_split_tree(tree);
  if (rc) return rc;
  CT_REQUIRE(n_sub >= 0 && (n_sub == 0 || (subs && query_offsets && prefix)) && (kind == 0 || kind == 1),
             "bad arguments");
  CT_REQUIRE(kind == 1 ? tree->n_max_vert >= 3 : tree->n_max_vert == 2, "kind does not match the tree");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024), "hit_slots must be in [1, 1024]");
  if (n_sub == 0) return CT_OK;
  return launch_edges_sub<true>(tree, segments, subs, n_sub, kind, nullptr, query_offsets, prefix, ii, jj, xy,
                                const_cast<int32_t*>(hits), hits ? hit_slots : 0, as_stream(stream));
}

}  // extern "C"
