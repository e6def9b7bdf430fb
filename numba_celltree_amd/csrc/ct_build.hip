// ct_build.hip -- level-synchronous GPU build of the Garth-Joy bucket-split cell tree.
//
// The reference (creation.py:233-381) builds serially with a DFS stack.  Each node's
// work depends only on its own cell range, so all nodes of one level are split at once:
//
//   per level (A active nodes, one "slot" each; every cell position knows its slot):
//     k_range      per position: min bb[2d] / max bb[2d+1] per slot      (get_bounds :153-171)
//     k_bucket     per position: bucket id of the centroid               (centroid_test :44-51)
//     k_rank_*     global exclusive scan of one-hot bucket counts -> stable rank of every
//                  cell inside its (node, bucket) = the order sort_bbox_indices produces
//                  with its repeated stable_partition (:54-151)
//     k_scatter    per position: the stably partitioned permutation
//     k_bbounds    per position: per-bucket Rmin/Lmax (the get_bounds loop :300-306)
//     k_decide     per slot: special case (:312-320), empty-bucket merge (:322-341),
//                  retry-other-dimension / degenerate leaf (:343-358), split_plane (:174-213),
//                  children (:365-379)
//     k_reseg      per position: slot of the child that now owns the position
//
// Nodes are first numbered in creation order; a final pass renumbers them to the
// reference's DFS numbering: the children of the k-th internal node in DFS preorder
// (left first) sit at 1 + 2k (creation.py:373-379), so a bottom-up count of internal
// nodes per subtree and a top-down prefix give every node its reference index.
//
// Order-sensitive details kept from the reference: FP64 bucket bounds i*bl + Rmin and
// (i+1)*bl + Rmin (no FMA), first strict minimum of the split cost, a node that retries
// keeps its flipped dim even when it ends as a leaf.
#include <vector>

#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {
namespace {

constexpr int kRankItems = 4;
constexpr int kRankTile = kBlock * kRankItems;
constexpr int kMaxBuckets = 64;

enum Decision : int32_t { kLeaf = 0, kSplit = 1, kRetry = 2 };

struct Slots {  // active-node list, structure of arrays
  int32_t* id;     // creation-order node id
  int32_t* ptr;
  int32_t* size;
  int32_t* dim;
  int32_t* att;    // 0 first attempt, 1 after a dimension flip
  int32_t* depth;
};

struct Work {
  int32_t* perm[2];
  int32_t* seg;
  int32_t* bid[2];
  int32_t* rank;
  int32_t* blockcnt;
  Slots act[2];
  unsigned long long* rmin;
  unsigned long long* rmax;
  int32_t* sstart;
  int32_t* send;
  int32_t* boff;
  unsigned long long* bmin;
  unsigned long long* bmax;
  int32_t* decision;
  int32_t* right_index;
  int32_t* next_l;
  int32_t* next_r;
  ct_node* cnodes;  // creation-order nodes
  int32_t* icount;
  int32_t* drank;
  int32_t* newidx;
  int32_t* counters;  // [0] next node id, [1] next active count, [2] error, [3] max depth
};

struct Layout {
  size_t bytes = 0;
  template <typename T>
  size_t take(size_t count) {
    size_t off = (bytes + 255) & ~size_t(255);
    bytes = off + count * sizeof(T);
    return off;
  }
};

// Carves the workspace; with base == nullptr it only measures.
Work carve(char* base, int64_t n, int nb, size_t* total) {
  Layout L;
  Work w{};
  const int64_t amax = n / 2 + 2;
  const int64_t mmax = ct_max_nodes(n) + 2;
  const int64_t nblk = (n + kRankTile - 1) / kRankTile + 1;
  auto P = [&](size_t off) { return base ? base + off : nullptr; };
#define CT_TAKE(field, T, count) field = reinterpret_cast<T*>(P(L.take<T>(count)))
  CT_TAKE(w.perm[0], int32_t, n);
  CT_TAKE(w.perm[1], int32_t, n);
  CT_TAKE(w.seg, int32_t, n);
  CT_TAKE(w.bid[0], int32_t, n);
  CT_TAKE(w.bid[1], int32_t, n);
  CT_TAKE(w.rank, int32_t, n);
  CT_TAKE(w.blockcnt, int32_t, nblk * nb);
  for (int k = 0; k < 2; ++k) {
    CT_TAKE(w.act[k].id, int32_t, amax);
    CT_TAKE(w.act[k].ptr, int32_t, amax);
    CT_TAKE(w.act[k].size, int32_t, amax);
    CT_TAKE(w.act[k].dim, int32_t, amax);
    CT_TAKE(w.act[k].att, int32_t, amax);
    CT_TAKE(w.act[k].depth, int32_t, amax);
  }
  CT_TAKE(w.rmin, unsigned long long, amax);
  CT_TAKE(w.rmax, unsigned long long, amax);
  CT_TAKE(w.sstart, int32_t, amax * nb);
  CT_TAKE(w.send, int32_t, amax * nb);
  CT_TAKE(w.boff, int32_t, amax * nb);
  CT_TAKE(w.bmin, unsigned long long, amax * nb);
  CT_TAKE(w.bmax, unsigned long long, amax * nb);
  CT_TAKE(w.decision, int32_t, amax);
  CT_TAKE(w.right_index, int32_t, amax);
  CT_TAKE(w.next_l, int32_t, amax);
  CT_TAKE(w.next_r, int32_t, amax);
  CT_TAKE(w.cnodes, ct_node, mmax);
  CT_TAKE(w.icount, int32_t, mmax);
  CT_TAKE(w.drank, int32_t, mmax);
  CT_TAKE(w.newidx, int32_t, mmax);
  CT_TAKE(w.counters, int32_t, 8);
#undef CT_TAKE
  *total = L.bytes + 256;
  return w;
}

__global__ void k_init(int32_t* __restrict__ perm, int32_t* __restrict__ seg, int n, int root_active,
                       ct_node* __restrict__ cnodes, Slots act, int32_t* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    perm[i] = i;
    seg[i] = root_active ? 0 : -1;
  }
  if (i == 0) {
    cnodes[0] = ct_node{-1.0, -1.0, -1, 0, n, 0};  // creation.py:400-401
    act.id[0] = 0;
    act.ptr[0] = 0;
    act.size[0] = n;
    act.dim[0] = 0;
    act.att[0] = 0;
    act.depth[0] = 0;
    counters[0] = 1;
    counters[1] = 0;
    counters[2] = 0;
    counters[3] = 0;
  }
}

// Clears per-slot reduction targets for the A slots of this level and the next-level counter.
__global__ void k_level_reset(int A, int nb, unsigned long long* __restrict__ rmin,
                              unsigned long long* __restrict__ rmax,
                              unsigned long long* __restrict__ bmin,
                              unsigned long long* __restrict__ bmax, int32_t* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A) {
    rmin[i] = ~0ull;
    rmax[i] = 0ull;
  }
  if (i < A * nb) {
    bmin[i] = ~0ull;
    bmax[i] = 0ull;
  }
  if (i == 0) counters[1] = 0;
}

__global__ void __launch_bounds__(kBlock) k_range(int n, const int32_t* __restrict__ seg,
                                                  const int32_t* __restrict__ perm,
                                                  const int32_t* __restrict__ adim,
                                                  const double* __restrict__ bb,
                                                  unsigned long long* __restrict__ rmin,
                                                  unsigned long long* __restrict__ rmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int s = -1;
  unsigned long long lo = ~0ull, hi = 0ull;
  if (i < n) {
    s = seg[i];
    if (s >= 0) {
      const int d = adim[s];
      const double* row = bb + 4 * (int64_t)perm[i];
      lo = dkey(row[2 * d]);
      hi = dkey(row[2 * d + 1]);
    }
  }
  wave_run_minmax(s, lo, hi, rmin, rmax);
}

// creation.py:44-51 + the first-match bucket of sort_bbox_indices (:115-151).  Buckets
// [i*bl + Rmin, (i+1)*bl + Rmin) share their bounding expressions, so the first matching
// bucket is the only one.  No match: the reference indexes past its bucket list (error).
__global__ void __launch_bounds__(kBlock) k_bucket(int n, int nb, const int32_t* __restrict__ seg,
                                                   const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ adim,
                                                   const double* __restrict__ bb,
                                                   const unsigned long long* __restrict__ rmin,
                                                   const unsigned long long* __restrict__ rmax,
                                                   int32_t* __restrict__ bid,
                                                   int32_t* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = seg[i];
  if (s < 0) return;
  const int d = adim[s];
  const double range_Rmin = dval(rmin[s]);
  const double range_Lmax = dval(rmax[s]);
  const double bl = (range_Lmax - range_Rmin) / (double)nb;  // creation.py:278
  const double* row = bb + 4 * (int64_t)perm[i];
  const double lo = row[2 * d], hi = row[2 * d + 1];
  const double centroid = lo + 0.5 * (hi - lo);
  int b = -1;
  for (int k = 0; k < nb; ++k) {
    const double bmax = (double)(k + 1) * bl + range_Rmin;  // creation.py:286
    const double bmin = (double)k * bl + range_Rmin;        // creation.py:287
    if ((centroid >= bmin) && (centroid < bmax)) {
      b = k;
      break;
    }
  }
  if (b < 0) {
    counters[2] = 1;
    b = nb - 1;
  }
  bid[i] = b;
}

// Per-tile bucket histograms (positions of active nodes only).
__global__ void __launch_bounds__(kBlock) k_rank_count(int n, int nb, const int32_t* __restrict__ seg,
                                                       const int32_t* __restrict__ bid,
                                                       int32_t* __restrict__ blockcnt) {
  __shared__ int cnt[kMaxBuckets];
  for (int b = threadIdx.x; b < nb; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRankTile + threadIdx.x * kRankItems;
#pragma unroll
  for (int j = 0; j < kRankItems; ++j) {
    const int p = base + j;
    if (p < n && seg[p] >= 0) atomicAdd(&cnt[bid[p]], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) blockcnt[blockIdx.x * nb + b] = cnt[b];
}

// Exclusive scan of the tile histograms, per bucket, in one workgroup.
__global__ void __launch_bounds__(kBlock) k_rank_scan(int nblk, int nb, int32_t* __restrict__ blockcnt) {
  __shared__ int lds[kBlock / 64];
  for (int b = 0; b < nb; ++b) {
    int carry = 0;
    for (int base = 0; base < nblk; base += kBlock) {
      const int k = base + threadIdx.x;
      const int v = (k < nblk) ? blockcnt[k * nb + b] : 0;
      int total;
      const int ex = block_exclusive_scan<int>(v, lds, &total);
      if (k < nblk) blockcnt[k * nb + b] = carry + ex;
      carry += total;
    }
  }
}

// Global prefix count of each position's bucket; node heads / tails record the prefix
// vector at their segment start / end so k_scatter can rebase ranks per node.
__global__ void __launch_bounds__(kBlock) k_rank_write(int n, int nb, const int32_t* __restrict__ seg,
                                                       const int32_t* __restrict__ bid,
                                                       const int32_t* __restrict__ blockcnt,
                                                       const int32_t* __restrict__ aptr,
                                                       const int32_t* __restrict__ asize,
                                                       int32_t* __restrict__ rank,
                                                       int32_t* __restrict__ sstart,
                                                       int32_t* __restrict__ send) {
  __shared__ int lds[kBlock / 64];
  const int base = blockIdx.x * kRankTile + threadIdx.x * kRankItems;
  int ps[kRankItems], pb[kRankItems];
#pragma unroll
  for (int j = 0; j < kRankItems; ++j) {
    const int p = base + j;
    ps[j] = (p < n) ? seg[p] : -1;
    pb[j] = (ps[j] >= 0) ? bid[p] : -1;
  }
  for (int b = 0; b < nb; ++b) {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) c += (pb[j] == b);
    int total;
    int run = blockcnt[blockIdx.x * nb + b] + block_exclusive_scan<int>(c, lds, &total);
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
      const int p = base + j;
      const int s = ps[j];
      if (s < 0) continue;
      if (p == aptr[s]) sstart[s * nb + b] = run;
      if (pb[j] == b) {
        rank[p] = run;
        ++run;
      }
      if (p == aptr[s] + asize[s] - 1) send[s * nb + b] = run;
    }
  }
}

// Per node: bucket start offsets relative to the node's first position.
__global__ void k_offsets(int A, int nb, const int32_t* __restrict__ sstart,
                          const int32_t* __restrict__ send, int32_t* __restrict__ boff) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A) return;
  int off = 0;
  for (int b = 0; b < nb; ++b) {
    boff[s * nb + b] = off;
    off += send[s * nb + b] - sstart[s * nb + b];
  }
}

__global__ void __launch_bounds__(kBlock) k_scatter(int n, int nb, const int32_t* __restrict__ seg,
                                                    const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ bid,
                                                    const int32_t* __restrict__ rank,
                                                    const int32_t* __restrict__ aptr,
                                                    const int32_t* __restrict__ sstart,
                                                    const int32_t* __restrict__ boff,
                                                    int32_t* __restrict__ perm2,
                                                    int32_t* __restrict__ bid2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = seg[p];
  if (s < 0) {
    perm2[p] = perm[p];
    return;
  }
  const int b = bid[p];
  const int q = aptr[s] + boff[s * nb + b] + (rank[p] - sstart[s * nb + b]);
  perm2[q] = perm[p];
  bid2[q] = b;
}

__global__ void __launch_bounds__(kBlock) k_bbounds(int n, int nb, const int32_t* __restrict__ seg,
                                                    const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ bid,
                                                    const int32_t* __restrict__ adim,
                                                    const double* __restrict__ bb,
                                                    unsigned long long* __restrict__ bmin,
                                                    unsigned long long* __restrict__ bmax) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1;
  unsigned long long lo = ~0ull, hi = 0ull;
  if (p < n) {
    const int s = seg[p];
    if (s >= 0) {
      key = s * nb + bid[p];
      const int d = adim[s];
      const double* row = bb + 4 * (int64_t)perm[p];
      lo = dkey(row[2 * d]);
      hi = dkey(row[2 * d + 1]);
    }
  }
  wave_run_minmax(key, lo, hi, bmin, bmax);
}

__global__ void k_decide(int A, int nb, int cells_per_leaf, Slots cur, Slots nxt,
                         const unsigned long long* __restrict__ rmin,
                         const unsigned long long* __restrict__ rmax,
                         const int32_t* __restrict__ sstart, const int32_t* __restrict__ send,
                         const int32_t* __restrict__ boff,
                         const unsigned long long* __restrict__ bminv,
                         const unsigned long long* __restrict__ bmaxv,
                         ct_node* __restrict__ cnodes, int32_t* __restrict__ decision,
                         int32_t* __restrict__ right_index, int32_t* __restrict__ next_l,
                         int32_t* __restrict__ next_r, int32_t* __restrict__ counters) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A) return;
  const int X = cur.id[s], ptr = cur.ptr[s], size = cur.size[s], dim = cur.dim[s];
  const int att = cur.att[s], depth = cur.depth[s];
  const double range_Rmin = dval(rmin[s]);
  const double range_Lmax = dval(rmax[s]);
  const double bl = (range_Lmax - range_Rmin) / (double)nb;
  const int ndim = dim ? 0 : 1;

  auto push_child = [&](int id, int cptr, int csize) -> int {
    cnodes[id] = ct_node{-1.0, -1.0, -1, cptr, csize, ndim};  // creation.py:27-29
    if (csize <= cells_per_leaf) return -1;
    const int ns = atomicAdd(&counters[1], 1);
    nxt.id[ns] = id;
    nxt.ptr[ns] = cptr;
    nxt.size[ns] = csize;
    nxt.dim[ns] = ndim;
    nxt.att[ns] = 0;
    nxt.depth[ns] = depth + 1;
    return ns;
  };

  if (cells_per_leaf == 1 && size == 2) {  // creation.py:312-320
    cnodes[X].Lmax = range_Lmax;
    cnodes[X].Rmin = range_Rmin;
    const int c = atomicAdd(&counters[0], 2);
    cnodes[X].child = c;
    push_child(c, ptr, 1);
    push_child(c + 1, ptr + 1, 1);
    atomicMax(&counters[3], depth + 1);
    decision[s] = kLeaf;  // both children are leaves: no position stays active
    return;
  }

  // Non-empty buckets in order (creation.py:322-341 merges the empty ones away; only
  // Rmin, Lmax, index and size matter afterwards).
  double bR[kMaxBuckets], bL[kMaxBuckets];
  int bI[kMaxBuckets], bS[kMaxBuckets];
  int m = 0;
  for (int b = 0; b < nb; ++b) {
    const int cnt = send[s * nb + b] - sstart[s * nb + b];
    if (cnt > 0) {
      bR[m] = dval(bminv[s * nb + b]);
      bL[m] = dval(bmaxv[s * nb + b]);
      bI[m] = ptr + boff[s * nb + b];
      bS[m] = cnt;
      ++m;
    }
  }

  if (m == 1) {  // creation.py:343-358: every cell in one bucket
    if (att == 0) {
      cnodes[X].dim = ndim;
      const int ns = atomicAdd(&counters[1], 1);
      nxt.id[ns] = X;
      nxt.ptr[ns] = ptr;
      nxt.size[ns] = size;
      nxt.dim[ns] = ndim;
      nxt.att[ns] = 1;
      nxt.depth[ns] = depth;
      decision[s] = kRetry;
      next_l[s] = ns;
    } else {
      cnodes[X].Lmax = -1.0;
      cnodes[X].Rmin = -1.0;
      decision[s] = kLeaf;
    }
    return;
  }

  // creation.py:174-213 split_plane
  double plane_min_cost = 1.7976931348623157e308;
  int plane = -1;
  long long left = 0;
  for (int i = 1; i < m; ++i) {
    left += bS[i - 1];
    const long long right = (long long)size - left;
    const double left_volume = (bL[i - 1] - range_Rmin) / bl;
    const double right_volume = (range_Lmax - bR[i]) / bl;
    const double cost = left_volume * (double)left + right_volume * (double)right;
    if (cost < plane_min_cost) {
      plane_min_cost = cost;
      plane = i;
    }
  }
  if (plane < 0) {
    counters[2] = 2;
    decision[s] = kLeaf;
    return;
  }
  double Lmax = -1.7976931348623157e308, Rmin = 1.7976931348623157e308;
  for (int i = 0; i < plane; ++i)
    if (bL[i] > Lmax) Lmax = bL[i];
  for (int i = plane; i < m; ++i)
    if (bR[i] < Rmin) Rmin = bR[i];
  const int ridx = bI[plane];
  const int rsize = ptr + size - ridx;
  const int lsize = size - rsize;
  cnodes[X].Lmax = Lmax;
  cnodes[X].Rmin = Rmin;
  const int c = atomicAdd(&counters[0], 2);
  cnodes[X].child = c;
  next_l[s] = push_child(c, ptr, lsize);
  next_r[s] = push_child(c + 1, ridx, rsize);
  right_index[s] = ridx;
  decision[s] = kSplit;
  atomicMax(&counters[3], depth + 1);
}

__global__ void __launch_bounds__(kBlock) k_reseg(int n, int32_t* __restrict__ seg,
                                                  const int32_t* __restrict__ decision,
                                                  const int32_t* __restrict__ right_index,
                                                  const int32_t* __restrict__ next_l,
                                                  const int32_t* __restrict__ next_r) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = seg[p];
  if (s < 0) return;
  const int dcs = decision[s];
  int ns = -1;
  if (dcs == kSplit) ns = (p < right_index[s]) ? next_l[s] : next_r[s];
  else if (dcs == kRetry) ns = next_l[s];
  seg[p] = ns;
}

// Bottom-up: number of internal nodes in each subtree (creation-order groups, reversed).
__global__ void k_icount(int lo, int hi, const ct_node* __restrict__ cn, int32_t* __restrict__ icount) {
  const int x = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= hi) return;
  const int c = cn[x].child;
  icount[x] = (c >= 0) ? 1 + icount[c] + icount[c + 1] : 0;
}

// Top-down: DFS preorder rank of internal nodes -> reference indices of their children.
__global__ void k_dfs(int lo, int hi, const ct_node* __restrict__ cn, const int32_t* __restrict__ icount,
                      int32_t* __restrict__ drank, int32_t* __restrict__ newidx) {
  const int x = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= hi) return;
  if (x == 0) {
    newidx[0] = 0;
    drank[0] = 0;
  }
  const int c = cn[x].child;
  if (c < 0) return;
  const int r = (x == 0) ? 0 : drank[x];
  newidx[c] = 1 + 2 * r;
  newidx[c + 1] = 2 + 2 * r;
  drank[c] = r + 1;
  drank[c + 1] = r + 1 + icount[c];
}

__global__ void k_emit(int total, const ct_node* __restrict__ cn, const int32_t* __restrict__ drank,
                       const int32_t* __restrict__ newidx, ct_node* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= total) return;
  ct_node nd = cn[x];
  nd.child = (nd.child >= 0) ? 1 + 2 * drank[x] : -1;
  out[newidx[x]] = nd;
}

__global__ void k_copy_i32(int n, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int64_t ct_max_nodes(int64_t n_elements) {  // creation.py:216-230
  int64_t n_nodes = n_elements;
  int64_t nodes = (n_elements + 1) / 2;
  while (nodes > 1) {
    n_nodes += nodes;
    nodes = (nodes + 1) / 2;
  }
  return n_nodes + 1;
}

size_t ct_build_workspace_bytes(int64_t n_elements, int32_t n_buckets) {
  size_t total = 0;
  carve(nullptr, n_elements < 1 ? 1 : n_elements, n_buckets < 2 ? 2 : n_buckets, &total);
  return total;
}

int ct_build(const double* bb_coords, int64_t n_elements, int32_t n_buckets, int32_t cells_per_leaf,
             void* workspace, size_t workspace_bytes, ct_node* nodes_out, int32_t* bb_indices_out,
             int64_t* n_nodes, int32_t* depth, void* stream) {
  CT_REQUIRE(n_elements >= 1 && n_elements < (int64_t)1 << 30, "n_elements must be in [1, 2^30)");
  CT_REQUIRE(n_buckets >= 2 && n_buckets <= kMaxBuckets, "n_buckets must be in [2, %d]", kMaxBuckets);
  CT_REQUIRE(cells_per_leaf >= 1, "cells_per_leaf must be >= 1");
  CT_REQUIRE(n_nodes && depth && nodes_out && bb_indices_out && bb_coords, "null pointer");
  size_t need = 0;
  carve(nullptr, n_elements, n_buckets, &need);
  if (workspace_bytes < need) {
    set_error("build workspace too small: %zu < %zu", workspace_bytes, need);
    return CT_ERR_WORKSPACE;
  }
  hipStream_t st = as_stream(stream);
  Work w = carve(static_cast<char*>(workspace), n_elements, n_buckets, &need);
  const int n = (int)n_elements, nb = n_buckets, cpl = cells_per_leaf;
  const unsigned gp = grid_for(n);
  const int nblk = (n + kRankTile - 1) / kRankTile;

  int A = (n > cpl) ? 1 : 0;
  k_init<<<gp, kBlock, 0, st>>>(w.perm[0], w.seg, n, A, w.cnodes, w.act[0], w.counters);
  CT_LAUNCH_CHECK();

  std::vector<int> group_end = {1};  // creation-order id ranges, one per level
  int cur = 0;
  int32_t host_counters[4] = {1, 0, 0, 0};
  int levels = 0;
  while (A > 0) {
    Slots& ca = w.act[cur];
    Slots& na = w.act[cur ^ 1];
    k_level_reset<<<grid_for((int64_t)A * nb), kBlock, 0, st>>>(A, nb, w.rmin, w.rmax, w.bmin, w.bmax,
                                                                 w.counters);
    CT_LAUNCH_CHECK();
    k_range<<<gp, kBlock, 0, st>>>(n, w.seg, w.perm[0], ca.dim, bb_coords, w.rmin, w.rmax);
    CT_LAUNCH_CHECK();
    k_bucket<<<gp, kBlock, 0, st>>>(n, nb, w.seg, w.perm[0], ca.dim, bb_coords, w.rmin, w.rmax,
                                    w.bid[0], w.counters);
    CT_LAUNCH_CHECK();
    k_rank_count<<<nblk, kBlock, 0, st>>>(n, nb, w.seg, w.bid[0], w.blockcnt);
    CT_LAUNCH_CHECK();
    k_rank_scan<<<1, kBlock, 0, st>>>(nblk, nb, w.blockcnt);
    CT_LAUNCH_CHECK();
    k_rank_write<<<nblk, kBlock, 0, st>>>(n, nb, w.seg, w.bid[0], w.blockcnt, ca.ptr, ca.size, w.rank,
                                          w.sstart, w.send);
    CT_LAUNCH_CHECK();
    k_offsets<<<grid_for(A), kBlock, 0, st>>>(A, nb, w.sstart, w.send, w.boff);
    CT_LAUNCH_CHECK();
    k_scatter<<<gp, kBlock, 0, st>>>(n, nb, w.seg, w.perm[0], w.bid[0], w.rank, ca.ptr, w.sstart,
                                     w.boff, w.perm[1], w.bid[1]);
    CT_LAUNCH_CHECK();
    k_bbounds<<<gp, kBlock, 0, st>>>(n, nb, w.seg, w.perm[1], w.bid[1], ca.dim, bb_coords, w.bmin,
                                     w.bmax);
    CT_LAUNCH_CHECK();
    k_decide<<<grid_for(A), kBlock, 0, st>>>(A, nb, cpl, ca, na, w.rmin, w.rmax, w.sstart, w.send,
                                             w.boff, w.bmin, w.bmax, w.cnodes, w.decision,
                                             w.right_index, w.next_l, w.next_r, w.counters);
    CT_LAUNCH_CHECK();
    k_reseg<<<gp, kBlock, 0, st>>>(n, w.seg, w.decision, w.right_index, w.next_l, w.next_r);
    CT_LAUNCH_CHECK();
    std::swap(w.perm[0], w.perm[1]);
    cur ^= 1;
    CT_HIP(hipMemcpyAsync(host_counters, w.counters, sizeof(host_counters), hipMemcpyDeviceToHost, st));
    CT_HIP(hipStreamSynchronize(st));
    if (host_counters[2] != 0) {
      set_error(host_counters[2] == 1
                    ? "degenerate bucketing: a cell centroid falls outside every bucket"
                    : "degenerate bucketing: no finite split cost");
      return CT_ERR_DEGENERATE;
    }
    group_end.push_back(host_counters[0]);
    A = host_counters[1];
    // Children are strictly smaller than their parent and a node retries once, so the
    // loop ends within 2 * n_elements levels; the guard only protects the host loop.
    if (++levels > 2 * n + 2) {
      set_error("tree build did not converge");
      return CT_ERR_DEGENERATE;
    }
  }
  const int total = host_counters[0];

  // DFS renumbering.
  for (int g = (int)group_end.size() - 1; g >= 0; --g) {
    const int lo = (g == 0) ? 0 : group_end[g - 1], hi = group_end[g];
    if (hi > lo) {
      k_icount<<<grid_for(hi - lo), kBlock, 0, st>>>(lo, hi, w.cnodes, w.icount);
      CT_LAUNCH_CHECK();
    }
  }
  for (size_t g = 0; g < group_end.size(); ++g) {
    const int lo = (g == 0) ? 0 : group_end[g - 1], hi = group_end[g];
    if (hi > lo) {
      k_dfs<<<grid_for(hi - lo), kBlock, 0, st>>>(lo, hi, w.cnodes, w.icount, w.drank, w.newidx);
      CT_LAUNCH_CHECK();
    }
  }
  k_emit<<<grid_for(total), kBlock, 0, st>>>(total, w.cnodes, w.drank, w.newidx, nodes_out);
  CT_LAUNCH_CHECK();
  k_copy_i32<<<gp, kBlock, 0, st>>>(n, w.perm[0], bb_indices_out);
  CT_LAUNCH_CHECK();
  *n_nodes = total;
  *depth = host_counters[3];
  return CT_OK;
}

}  // extern "C"
