// ct_batch.hip -- many trees, one persistent launch (no reference counterpart: the reference
// answers one mesh per call, celltree.py:216-239 locate_points, :294-303 intersect_edges).
//
// A batch is a device array of ct_tree descriptors (each with its own device arrays), the
// queries of all trees concatenated (tree m owns [qoff[m], qoff[m + 1])), and a tile list:
// ranges of at most 64 * k queries of one tree, grouped into n_queues work queues
// (queue r owns tiles [qstart[r], qstart[r + 1])).  The host assigns whole trees to queues
// by estimated work, so the workgroups of one XCD (workgroup b runs on XCD b % 8, and takes
// queue b % n_queues first) keep one tree's nodes in their own L2.
//
// Every wavefront of the persistent grid pulls tiles on its own: lane 0 bumps the queue's
// counter, the index is broadcast, and the wave walks the tile's queries one per lane with
// the single-tree traversal (ct_traverse.h), so every answer equals the per-tree call's.
// A wave whose queue is drained moves to the next queue; after n_queues drained queues in
// a row it exits -- queues only drain, so every wave ends once all tiles are taken.  The
// counters are cleared by a fill kernel inside each call (no memset), so a sequence of
// batch calls can be captured in a HIP graph and replayed.
#include <type_traits>

#include "ct_internal.h"
#include "ct_primitives.h"
#include "ct_traverse.h"

namespace ct {
namespace {

constexpr int kWave = 64;
// Largest workgroup of the persistent kernels.  Without a bound the compiler assumes 1024
// threads (4 waves per SIMD), caps a lane at 128 VGPRs and spills the segment traversal to
// scratch; with 256 (one wave per SIMD) it may use what the traversal needs.
constexpr int kBatchMaxBlock = 256;
static_assert(kBatchMaxBlock == kStackStride, "batch workgroups are at most kStackStride lanes");

struct BatchArgs {
  const ct_tree* trees;
  const double* tols;
  const int64_t* qoff;
  const ct_batch_tile* tiles;
  const int32_t* qstart;
  int32_t n_queues;
  int32_t* counters;
};

// Element i of a read-only device array, for a wave-uniform i: through the constant address
// space the load becomes a scalar load into SGPRs.  A tile's descriptor, tree, tolerance
// and query offset are the same for the whole wavefront; held in VGPRs they cost ~12
// registers per lane and a wavefront per SIMD of occupancy.
// Structs are read as 32-bit words and reassembled.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p, int64_t i) {
  static_assert(sizeof(T) % 4 == 0 && std::is_trivially_copyable<T>::value, "plain records of whole words");
  constexpr int kWords = (int)(sizeof(T) / 4);
  const __attribute__((address_space(4))) uint32_t* src =
      (const __attribute__((address_space(4))) uint32_t*)(p + i);
  uint32_t w[kWords];
#pragma unroll
  for (int k = 0; k < kWords; ++k) w[k] = src[k];
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Next tile for this wavefront, or -1 when every queue is drained.  Wave-uniform.
__device__ __forceinline__ int next_tile(const BatchArgs& a, int& queue, int& misses) {
  const int lane = threadIdx.x & (kWave - 1);
  while (misses < a.n_queues) {
    int i = 0;
    if (lane == 0) i = atomicAdd(&a.counters[queue], 1);
    i = __builtin_amdgcn_readfirstlane(i);
    const int lo = a.qstart[queue], hi = a.qstart[queue + 1];
    if (i < hi - lo) return lo + i;
    queue = (queue + 1 == a.n_queues) ? 0 : queue + 1;
    ++misses;
  }
  return -1;
}

template <int MAXV, int STACK>
__global__ void __launch_bounds__(kBatchMaxBlock) k_batch_points(BatchArgs a, const double* __restrict__ pts, int64_t* __restrict__ out) {
  CT_LANE_STACK(STACK, stack);
  int queue = blockIdx.x % a.n_queues, misses = 0;
  const int lane = threadIdx.x & (kWave - 1);
  for (int tid = next_tile(a, queue, misses); tid >= 0; tid = next_tile(a, queue, misses)) {
    const ct_batch_tile tile = ld_uniform(a.tiles, __builtin_amdgcn_readfirstlane(tid));
    const int m = __builtin_amdgcn_readfirstlane(tile.tree);
    const TreeView t = view(ld_uniform(a.trees, m));
    const double tol = ld_uniform(a.tols, m);
    const int64_t base = ld_uniform(a.qoff, m) + tile.begin;
    for (int j = lane; j < tile.count; j += kWave) {
      const int64_t q = base + j;
      const P2 p = ld_point(pts, q);
      // wave-uniform choice of the leaf-row layout of this tree
      int r;
      if (t.leaf == nullptr) r = locate_one<MAXV, STACK, false, kLeafFaces>(t, p, tol, stack);
      else if (t.nmax == MAXV) r = locate_one<MAXV, STACK, false, kLeafPackedExact>(t, p, tol, stack);
      else r = locate_one<MAXV, STACK, false, kLeafPacked>(t, p, tol, stack);
      out[q] = r;
    }
  }
}

// Segment-face queries.  WRITE false: counts[q]; WRITE true: rows at offsets[q] when the
// query's rows end within `capacity` (ii = concatenated query index, jj = the tree's cell).
// With a hit record (hits != null, slot-major over the concatenated query index: hit c of
// query q at hits[c * n_queries + q]) the count pass keeps the first `slots` hit cells and
// the write pass rebuilds the rows of queries with at most `slots` hits from them (the same
// cell predicates in the same DFS order) instead of walking the tree again.
template <bool WRITE, int MAXV, int STACK>
__global__ void __launch_bounds__(kBatchMaxBlock) k_batch_edges(BatchArgs a, const double* __restrict__ segs, int64_t* __restrict__ counts,
                              const int64_t* __restrict__ offsets, int64_t capacity,
                              int64_t* __restrict__ ii, int64_t* __restrict__ jj,
                              double* __restrict__ xy, int32_t* __restrict__ hits, int slots,
                              int64_t n_queries) {
  CT_LANE_STACK(STACK, stack);
  int queue = blockIdx.x % a.n_queues, misses = 0;
  const int lane = threadIdx.x & (kWave - 1);
  for (int tid = next_tile(a, queue, misses); tid >= 0; tid = next_tile(a, queue, misses)) {
    const ct_batch_tile tile = ld_uniform(a.tiles, __builtin_amdgcn_readfirstlane(tid));
    const int m = __builtin_amdgcn_readfirstlane(tile.tree);
    const TreeView t = view(ld_uniform(a.trees, m));
    const int64_t base = ld_uniform(a.qoff, m) + tile.begin;
    for (int j = lane; j < tile.count; j += kWave) {
      const int64_t q = base + j;
      const double4 s4 = reinterpret_cast<const double4*>(segs)[q];
      const P2 pa = {s4.x, s4.y}, pb = {s4.z, s4.w};
      if (WRITE) {
        const int64_t o = offsets[q], end = offsets[q + 1];
        if (end <= capacity) {
          if (hits != nullptr && end - o <= slots) {
            for (int64_t h = 0; h < end - o; ++h) {
              const int c = hits[h * n_queries + q];
              P2 cc, dd;
              edge_face<MAXV>(t, c, pa, pb, &cc, &dd);
              ii[o + h] = q;
              jj[o + h] = c;
              reinterpret_cast<double4*>(xy)[o + h] = make_double4(cc.x, cc.y, dd.x, dd.y);
            }
          } else {
            edge_query<true, MAXV, STACK, true>(t, pa, pb, q, o, FinalRows{ii, jj, xy}, stack);
          }
        }
      } else {
        HitRecord rec;
        if (hits != nullptr) rec = HitRecord{hits + q, slots, n_queries};
        counts[q] = edge_query<false, MAXV, STACK, true>(t, pa, pb, q, 0, FinalRows{}, stack, rec);
      }
    }
  }
}

// Per-segment sort of the rows written by k_batch_edges.  The number of rows is read on the
// device (offsets[n_segments]); when it exceeds `capacity` the write pass was partial and
// nothing is sorted.  Grid-stride, so the launch does not depend on the row count.
__global__ void __launch_bounds__(kBlock) k_batch_sort(const double* __restrict__ segs, int64_t n_segments,
                                                       const int64_t* __restrict__ offsets,
                                                       const int64_t* __restrict__ ii,
                                                       const int64_t* __restrict__ jj_in,
                                                       const double* __restrict__ xy_in, int64_t capacity,
                                                       int64_t* __restrict__ jj_out, double* __restrict__ xy_out) {
  __shared__ double tw[kRankWindow];
  const int64_t total = offsets[n_segments];
  const int64_t rows = (total <= capacity) ? total : 0;
  for (int64_t r0 = (int64_t)blockIdx.x * kRankTile; r0 < rows; r0 += (int64_t)gridDim.x * kRankTile)
    sort_tile(segs, offsets, ii, jj_in, xy_in, rows, r0, jj_out, xy_out, tw);
}

struct PointsL {
  const BatchArgs& a;
  const double* pts;
  int64_t* out;
  unsigned grid, block;
  hipStream_t st;
  template <int MAXV, int S>
  int go() {
    k_batch_points<MAXV, S><<<grid, block, 0, st>>>(a, pts, out);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <bool WRITE>
struct EdgesL {
  const BatchArgs& a;
  const double* segs;
  int64_t* counts;
  const int64_t* offsets;
  int64_t capacity;
  int64_t* ii;
  int64_t* jj;
  double* xy;
  int32_t* hits;
  int slots;
  int64_t n_queries;
  unsigned grid, block;
  hipStream_t st;
  template <int MAXV, int S>
  int go() {
    k_batch_edges<WRITE, MAXV, S><<<grid, block, 0, st>>>(a, segs, counts, offsets, capacity, ii, jj, xy, hits, slots,
                                                           n_queries);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
};

template <int MV, typename L>
int by_stack(int stack, L& l) {
  switch (stack) {
    case 16: return l.template go<MV, 16>();
    case 24: return l.template go<MV, 24>();
    case 32: return l.template go<MV, 32>();
    case 48: return l.template go<MV, 48>();
    case 64: return l.template go<MV, 64>();
    case 128: return l.template go<MV, 128>();
    default: return l.template go<MV, 256>();
  }
}

// Picks the vertex class from max_n_vert and the compiled stack class from `stack`.
template <typename L>
int dispatch(int max_n_vert, int stack, L l) {
  if (max_n_vert <= 3) return by_stack<3>(stack, l);
  if (max_n_vert <= 4) return by_stack<4>(stack, l);
  return by_stack<8>(stack, l);
}

int check_batch(const ct_batch* b) {
  CT_REQUIRE(b != nullptr && b->trees && b->tolerances && b->query_offsets && b->tiles && b->queue_starts &&
                 b->counters,
             "incomplete batch");
  CT_REQUIRE(b->n_trees >= 1 && b->n_tiles >= 0, "bad batch sizes");
  CT_REQUIRE(b->n_queues >= 1 && b->n_queues <= 64, "n_queues must be in [1, 64]");
  CT_REQUIRE(b->max_n_vert >= 3 && b->max_n_vert <= 8, "batched queries support faces of 3..8 vertices");
  CT_REQUIRE(b->max_depth >= 0 && b->max_depth <= CT_MAX_DEPTH, "tree too deep");
  const int s = b->stack;
  CT_REQUIRE(s == 16 || s == 24 || s == 32 || s == 48 || s == 64 || s == 128 || s == 256,
             "stack must be one of 16, 24, 32, 48, 64, 128, 256");
  CT_REQUIRE(s >= b->max_depth, "stack smaller than the deepest tree");
  CT_REQUIRE(b->n_workgroups >= 1 && b->n_workgroups <= 65536, "n_workgroups must be in [1, 65536]");
  CT_REQUIRE(b->block == 64 || b->block == 128 || b->block == 256, "block must be 64, 128 or 256");
  return CT_OK;
}

BatchArgs args(const ct_batch* b) {
  return BatchArgs{b->trees, b->tolerances, b->query_offsets, b->tiles, b->queue_starts, b->n_queues,
                   b->counters};
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_batch_locate_points(const ct_batch* batch, const double* points, int64_t* out, void* stream) {
  int rc = check_batch(batch);
  if (rc) return rc;
  CT_REQUIRE(points && out, "bad arguments");
  if (batch->n_tiles == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  rc = fill<int32_t>(batch->counters, batch->n_queues, 0, st);
  if (rc) return rc;
  const BatchArgs a = args(batch);
  return dispatch(batch->max_n_vert, batch->stack,
                  PointsL{a, points, out, (unsigned)batch->n_workgroups, (unsigned)batch->block, st});
}

int ct_batch_edges_count_rec(const ct_batch* batch, const double* segments, int64_t n_queries, int64_t* counts,
                             int32_t* hits, int32_t hit_slots, void* stream) {
  int rc = check_batch(batch);
  if (rc) return rc;
  CT_REQUIRE(segments && counts, "bad arguments");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024 && n_queries >= 0),
             "hit_slots must be in [1, 1024]");
  if (batch->n_tiles == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  rc = fill<int32_t>(batch->counters, batch->n_queues, 0, st);
  if (rc) return rc;
  const BatchArgs a = args(batch);
  return dispatch(batch->max_n_vert, batch->stack,
                  EdgesL<false>{a, segments, counts, nullptr, 0, nullptr, nullptr, nullptr, hits,
                                hits ? hit_slots : 0, n_queries, (unsigned)batch->n_workgroups,
                                (unsigned)batch->block, st});
}

int ct_batch_edges_write_rec(const ct_batch* batch, const double* segments, int64_t n_queries,
                             const int64_t* offsets, const int32_t* hits, int32_t hit_slots, int64_t capacity,
                             int64_t* ii, int64_t* jj, double* xy, void* stream) {
  int rc = check_batch(batch);
  if (rc) return rc;
  CT_REQUIRE(segments && offsets && capacity >= 0, "bad arguments");
  CT_REQUIRE(capacity == 0 || (ii && jj && xy), "null output arrays");
  CT_REQUIRE(hits == nullptr || (hit_slots >= 1 && hit_slots <= 1024 && n_queries >= 0),
             "hit_slots must be in [1, 1024]");
  if (batch->n_tiles == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  rc = fill<int32_t>(batch->counters, batch->n_queues, 0, st);
  if (rc) return rc;
  const BatchArgs a = args(batch);
  return dispatch(batch->max_n_vert, batch->stack,
                  EdgesL<true>{a, segments, nullptr, offsets, capacity, ii, jj, xy, const_cast<int32_t*>(hits),
                               hits ? hit_slots : 0, n_queries, (unsigned)batch->n_workgroups,
                               (unsigned)batch->block, st});
}

int ct_batch_edges_count(const ct_batch* batch, const double* segments, int64_t* counts, void* stream) {
  return ct_batch_edges_count_rec(batch, segments, 0, counts, nullptr, 0, stream);
}

int ct_batch_edges_write(const ct_batch* batch, const double* segments, const int64_t* offsets,
                         int64_t capacity, int64_t* ii, int64_t* jj, double* xy, void* stream) {
  return ct_batch_edges_write_rec(batch, segments, 0, offsets, nullptr, 0, capacity, ii, jj, xy, stream);
}

int ct_batch_sort_intersections(const double* segments, int64_t n_segments, const int64_t* offsets,
                                const int64_t* ii, const int64_t* jj, const double* xy, int64_t capacity,
                                int32_t n_workgroups, int64_t* jj_out, double* xy_out, void* stream) {
  CT_REQUIRE(segments && offsets && n_segments >= 0 && capacity >= 0, "bad arguments");
  CT_REQUIRE(capacity == 0 || (ii && jj && xy && jj_out && xy_out), "null row arrays");
  CT_REQUIRE(n_workgroups >= 1 && n_workgroups <= 65536, "n_workgroups must be in [1, 65536]");
  if (capacity == 0 || n_segments == 0) return CT_OK;
  k_batch_sort<<<(unsigned)n_workgroups, kBlock, 0, as_stream(stream)>>>(segments, n_segments, offsets, ii, jj,
                                                                          xy, capacity, jj_out, xy_out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // extern "C"
