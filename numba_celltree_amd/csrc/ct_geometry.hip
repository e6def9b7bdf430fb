// ct_geometry.hip -- per-element geometry kernels that prepare a tree: counter-clockwise
// reordering, face / edge bounding boxes, and the bbox statistics (tree bbox, diagonals,
// default tolerance).  One lane per element; all FP64.
#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {

// geometry_utils.py:541-561 counter_clockwise.  a = v[len-2], b = v[len-1] are fixed
// once; the loop walks the *current* row, so after a flip it reads the reversed row and
// may flip again.  The row is kept in registers and each flip toggles a parity bit:
// current[i] = parity ? orig[len-1-i] : orig[i].
template <int MAXV>
__global__ void __launch_bounds__(kBlock) k_counter_clockwise(const double* __restrict__ vertices,
                                                              int32_t* __restrict__ faces,
                                                              int64_t n_faces, int nmax) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  int32_t* row = faces + f * nmax;
  int32_t idx[MAXV];
#pragma unroll
  for (int i = 0; i < MAXV; ++i) idx[i] = (i < nmax) ? row[i] : kFill;
  int len = nmax;
#pragma unroll
  for (int i = MAXV - 1; i >= 3; --i)
    if (i < nmax && idx[i] == kFill) len = i;

  auto vtx = [&](int i) { return ld_vertex(vertices, idx[i]); };
  P2 a = {0, 0}, b = {0, 0};
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    if (i == len - 2) a = vtx(i);
    if (i == len - 1) b = vtx(i);
  }
  bool parity = false;
  for (int i = 0; i < len; ++i) {
    const int src = parity ? (len - 1 - i) : i;
    int vi = idx[0];
#pragma unroll
    for (int j = 0; j < MAXV; ++j)
      if (j == src) vi = idx[j];
    const P2 c = ld_vertex(vertices, vi);
    const double product = cross(b.x - a.x, b.y - a.y, c.x - a.x, c.y - a.y);
    if (product == 0) {
      a = b;
      b = c;
    } else if (product < 0) {
      parity = !parity;
    } else {
      break;
    }
  }
  if (parity) {
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      if (i < len) {
        int v = idx[0];
#pragma unroll
        for (int j = 0; j < MAXV; ++j)
          if (j == len - 1 - i) v = idx[j];
        row[i] = v;
      }
    }
  }
}

// geometry_utils.py:421-456 bounding_box / build_face_bboxes: the first vertex seeds the
// box and the scan stops at the first fill value.
__global__ void __launch_bounds__(kBlock) k_face_bboxes(const int32_t* __restrict__ faces,
                                                        int64_t n_faces, int nmax,
                                                        const double* __restrict__ vertices,
                                                        double* __restrict__ out) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  const int32_t* row = faces + f * nmax;
  const P2 v0 = ld_vertex(vertices, row[0]);
  double xmin = v0.x, xmax = v0.x, ymin = v0.y, ymax = v0.y;
  for (int i = 1; i < nmax; ++i) {
    const int32_t k = row[i];
    if (k == kFill) break;
    const P2 v = ld_vertex(vertices, k);
    xmin = pmin(xmin, v.x);
    xmax = pmax(xmax, v.x);
    ymin = pmin(ymin, v.y);
    ymax = pmax(ymax, v.y);
  }
  reinterpret_cast<double4*>(out)[f] = make_double4(xmin, xmax, ymin, ymax);
}

// geometry_utils.py:459-487 edge_bounding_box / build_edge_bboxes (padded by tol).
__global__ void __launch_bounds__(kBlock) k_edge_bboxes(const int32_t* __restrict__ edges,
                                                        int64_t n, const double* __restrict__ vertices,
                                                        double tol, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const P2 p = ld_vertex(vertices, edges[2 * e]);
  const P2 q = ld_vertex(vertices, edges[2 * e + 1]);
  reinterpret_cast<double4*>(out)[e] =
      make_double4(pmin(p.x - tol, q.x - tol), pmax(p.x + tol, q.x + tol),
                   pmin(p.y - tol, q.y - tol), pmax(p.y + tol, q.y + tol));
}

// Reductions over all cells / vertices use a bounded grid (grid-stride loops) and one atomic
// per key per workgroup: one atomic per key per wavefront on the same five addresses
// serialises at the memory side (~0.9 ms for 1M cells).
constexpr unsigned kReduceGrid = 1024;

inline unsigned reduce_grid(int64_t n) {
  const unsigned g = grid_for(n);
  return g < kReduceGrid ? g : kReduceGrid;
}

// Max of K keys over the workgroup (wavefront shuffles, then LDS across the waves), then one
// atomicMax per key.
template <int K>
__device__ __forceinline__ void block_max_atomic(unsigned long long (&k)[K], unsigned long long* __restrict__ keys) {
  __shared__ unsigned long long part[kBlock / 64][K];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const unsigned long long t = __shfl_xor(k[j], d, 64);
      k[j] = t > k[j] ? t : k[j];
    }
  }
  const int w = threadIdx.x / 64;
  if (lane_id() == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j) part[w][j] = k[j];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    unsigned long long m = 0;
    for (int v = 0; v < (int)(blockDim.x / 64); ++v) m = part[v][threadIdx.x] > m ? part[v][threadIdx.x] : m;
    atomicMax(keys + threadIdx.x, m);
  }
}

// celltree_base.py:20-48: per-cell dx, dy, dxy = sqrt(dx**2 + dy**2) and the tree bbox /
// maximum diagonal as order-preserving keys (keys[0..4] = min x, max x, min y, max y,
// max dxy; min slots hold negated-order keys so one atomicMax serves all five).
__global__ void __launch_bounds__(kBlock) k_bbox_stats(const double* __restrict__ bb, int64_t n,
                                                       double* __restrict__ dist,
                                                       unsigned long long* __restrict__ keys) {
  unsigned long long k[5] = {0, 0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const Box b = ld_box(bb, i);
    const double dx = b.xmax - b.xmin;
    const double dy = b.ymax - b.ymin;
    const double dxy = sqrt(dx * dx + dy * dy);
    if (dist) {
      dist[3 * i] = dx;
      dist[3 * i + 1] = dy;
      dist[3 * i + 2] = dxy;
    }
    const unsigned long long c[5] = {~dkey(b.xmin), dkey(b.xmax), ~dkey(b.ymin), dkey(b.ymax), dkey(dxy)};
#pragma unroll
    for (int j = 0; j < 5; ++j) k[j] = c[j] > k[j] ? c[j] : k[j];
  }
  block_max_atomic<5>(k, keys);
}

__global__ void k_bbox_finish(const unsigned long long* __restrict__ keys, double* __restrict__ out,
                              int nk, const int* __restrict__ negate) {
  const int i = threadIdx.x;
  if (i < nk) out[i] = dval(negate[i] ? ~keys[i] : keys[i]);
}

// edge_celltree.py:60-64: extent of the vertices.
__global__ void __launch_bounds__(kBlock) k_vertex_extent(const double* __restrict__ v, int64_t n,
                                                          unsigned long long* __restrict__ keys) {
  unsigned long long k[4] = {0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const P2 p = ld_vertex(v, (int)i);
    const unsigned long long c[4] = {~dkey(p.x), dkey(p.x), ~dkey(p.y), dkey(p.y)};
#pragma unroll
    for (int j = 0; j < 4; ++j) k[j] = c[j] > k[j] ? c[j] : k[j];
  }
  block_max_atomic<4>(k, keys);
}

__global__ void k_extent_finish(const unsigned long long* __restrict__ keys, double* __restrict__ out) {
  const int i = threadIdx.x;
  if (i < 4) out[i] = dval((i % 2 == 0) ? ~keys[i] : keys[i]);
}

__global__ void k_stats_finish(const unsigned long long* __restrict__ keys, double* __restrict__ out) {
  const int i = threadIdx.x;
  if (i < 5) out[i] = dval((i == 0 || i == 2) ? ~keys[i] : keys[i]);
}

}  // namespace ct

using namespace ct;

extern "C" {

int ct_counter_clockwise(const double* vertices, int32_t* faces, int64_t n_faces,
                         int32_t n_max_vert, void* stream) {
  CT_REQUIRE(n_faces >= 0, "n_faces must be >= 0");
  CT_REQUIRE(n_max_vert >= 3 && n_max_vert <= CT_MAX_N_VERTEX, "n_max_vert must be in [3, 32]");
  if (n_faces == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  const unsigned g = grid_for(n_faces);
  if (n_max_vert <= 4) k_counter_clockwise<4><<<g, kBlock, 0, st>>>(vertices, faces, n_faces, n_max_vert);
  else if (n_max_vert <= 8) k_counter_clockwise<8><<<g, kBlock, 0, st>>>(vertices, faces, n_faces, n_max_vert);
  else if (n_max_vert <= 16) k_counter_clockwise<16><<<g, kBlock, 0, st>>>(vertices, faces, n_faces, n_max_vert);
  else k_counter_clockwise<32><<<g, kBlock, 0, st>>>(vertices, faces, n_faces, n_max_vert);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_build_face_bboxes(const int32_t* faces, int64_t n_faces, int32_t n_max_vert,
                         const double* vertices, double* out, void* stream) {
  CT_REQUIRE(n_faces >= 0 && n_max_vert >= 1, "bad sizes");
  if (n_faces == 0) return CT_OK;
  k_face_bboxes<<<grid_for(n_faces), kBlock, 0, as_stream(stream)>>>(faces, n_faces, n_max_vert,
                                                                      vertices, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_build_edge_bboxes(const int32_t* edges, int64_t n_edges, const double* vertices, double tol,
                         double* out, void* stream) {
  CT_REQUIRE(n_edges >= 0, "bad sizes");
  if (n_edges == 0) return CT_OK;
  k_edge_bboxes<<<grid_for(n_edges), kBlock, 0, as_stream(stream)>>>(edges, n_edges, vertices, tol,
                                                                      out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_bbox_stats(const double* bb_coords, int64_t n, double* distances, double* out5,
                  void* stream) {
  CT_REQUIRE(n > 0, "need at least one bounding box");
  hipStream_t st = as_stream(stream);
  // out5 doubles as the key scratch: 5 keys are reduced in place, then decoded.
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(out5);
  int rc = fill<unsigned long long>(keys, 5, 0ull, st);
  if (rc) return rc;
  k_bbox_stats<<<reduce_grid(n), kBlock, 0, st>>>(bb_coords, n, distances, keys);
  CT_LAUNCH_CHECK();
  k_stats_finish<<<1, 64, 0, st>>>(keys, out5);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_vertex_extent(const double* vertices, int64_t n, double* out4, void* stream) {
  CT_REQUIRE(n > 0, "need at least one vertex");
  hipStream_t st = as_stream(stream);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(out4);
  int rc = fill<unsigned long long>(keys, 4, 0ull, st);
  if (rc) return rc;
  k_vertex_extent<<<reduce_grid(n), kBlock, 0, st>>>(vertices, n, keys);
  CT_LAUNCH_CHECK();
  k_extent_finish<<<1, 64, 0, st>>>(keys, out4);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // extern "C"
