// ct_faces.hip -- pair-parallel FP64 polygon kernels: separating-axis test
// (separating_axis.py) and Sutherland-Hodgman clip area (sutherland_hodgman.py), one lane
// per (query face, tree face) candidate pair, both polygons in registers.
#include "ct_clip.h"
#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {
namespace {

constexpr double kFloatMax = 1.7976931348623157e308;

// separating_axis.py:17-39 extrema_projected / is_separating_axis.
template <int MA, int MB>
__device__ __forceinline__ bool separating(double nx, double ny, const P2 (&a)[MA], int la,
                                           const P2 (&b)[MB], int lb) {
  double mina = kFloatMax, maxa = -kFloatMax, minb = kFloatMax, maxb = -kFloatMax;
#pragma unroll
  for (int i = 0; i < MA; ++i) {
    if (i < la) {
      const double pr = dot(a[i].x, a[i].y, nx, ny);
      mina = pmin(mina, pr);
      maxa = pmax(maxa, pr);
    }
  }
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    if (i < lb) {
      const double pr = dot(b[i].x, b[i].y, nx, ny);
      minb = pmin(minb, pr);
      maxb = pmax(maxb, pr);
    }
  }
  return !(maxa > minb && maxb > mina);
}

// separating_axis.py:42-55 separating_axes(a, b): true when no edge normal of a separates.
template <int MA, int MB>
__device__ __forceinline__ bool no_separating_axis(const P2 (&a)[MA], int la, const P2 (&b)[MB], int lb) {
  P2 p = a[0];
#pragma unroll
  for (int i = 0; i < MA; ++i)
    if (i == la - 1) p = a[i];
#pragma unroll
  for (int i = 0; i < MA; ++i) {
    if (i < la) {
      const P2 q = a[i];
      const double nx = p.y - q.y, ny = q.x - p.x;
      p = q;
      if (!(nx == 0.0 && ny == 0.0) && separating<MA, MB>(nx, ny, a, la, b, lb)) return false;
    }
  }
  return true;
}

template <int MAXV>
__device__ __forceinline__ int load_face(const double* __restrict__ v, const int32_t* __restrict__ f, int nmax,
                                         int64_t k, P2 (&poly)[MAXV]) {
  return load_poly<MAXV>(f, nmax, k, v, poly);
}

// Fused pair stage of locate_faces / intersect_faces (celltree.py:186-269).
template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) k_face_pairs(const double* __restrict__ va,
                                                       const int32_t* __restrict__ fa, int na,
                                                       const double* __restrict__ vb,
                                                       const int32_t* __restrict__ fb, int nbm,
                                                       const int64_t* __restrict__ ii,
                                                       const int64_t* __restrict__ jj, int64_t n,
                                                       int with_area, double* __restrict__ area,
                                                       int64_t* __restrict__ keep) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  P2 a[MA], b[MB];
  const int la = load_face<MA>(va, fa, na, ii[k], a);
  const int lb = load_face<MB>(vb, fb, nbm, jj[k], b);
  const bool sat = no_separating_axis<MA, MB>(a, la, b, lb) && no_separating_axis<MB, MA>(b, lb, a, la);
  if (!with_area) {
    keep[k] = sat ? 1 : 0;
    return;
  }
  CT_CLIP_ROWS((clip_row_cap<MA, MB>()), row);
  double ar = 0.0;
  if (sat) ar = clip_area<MA, MB>(a, la, b, lb, row);
  area[k] = ar;
  keep[k] = (sat && ar > 0) ? 1 : 0;
}

// sutherland_hodgman.py:151-168 area_of_intersection (no SAT).
template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) k_area(const double* __restrict__ va, const int32_t* __restrict__ fa,
                                                 int na, const double* __restrict__ vb,
                                                 const int32_t* __restrict__ fb, int nbm,
                                                 const int64_t* __restrict__ ia,
                                                 const int64_t* __restrict__ ib, int64_t n,
                                                 double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  P2 a[MA], b[MB];
  const int la = load_face<MA>(va, fa, na, ia[k], a);
  const int lb = load_face<MB>(vb, fb, nbm, ib[k], b);
  CT_CLIP_ROWS((clip_row_cap<MA, MB>()), row);
  out[k] = clip_area<MA, MB>(a, la, b, lb, row);
}

// separating_axis.py:58-75 polygons_intersect.
template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) k_sat(const double* __restrict__ va, const int32_t* __restrict__ fa,
                                                int na, const double* __restrict__ vb,
                                                const int32_t* __restrict__ fb, int nbm,
                                                const int64_t* __restrict__ ia,
                                                const int64_t* __restrict__ ib, int64_t n,
                                                uint8_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  P2 a[MA], b[MB];
  const int la = load_face<MA>(va, fa, na, ia[k], a);
  const int lb = load_face<MB>(vb, fb, nbm, ib[k], b);
  out[k] = (no_separating_axis<MA, MB>(a, la, b, lb) && no_separating_axis<MB, MA>(b, lb, a, la)) ? 1 : 0;
}

// sutherland_hodgman.py:171-187 box_area_of_intersection (+ celltree.py:183 area > 0):
// subject = the box as a CCW quad (geometry_utils.py:513-524), clipper = the face.
template <int MB>
__global__ void __launch_bounds__(kBlock) k_box_pairs(const double* __restrict__ boxes,
                                                      const double* __restrict__ vb,
                                                      const int32_t* __restrict__ fb, int nbm,
                                                      const int64_t* __restrict__ ii,
                                                      const int64_t* __restrict__ jj, int64_t n,
                                                      double* __restrict__ area,
                                                      int64_t* __restrict__ keep) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const Box bx = ld_box(boxes, ii[k]);
  const P2 a[4] = {{bx.xmin, bx.ymin}, {bx.xmax, bx.ymin}, {bx.xmax, bx.ymax}, {bx.xmin, bx.ymax}};
  P2 b[MB];
  const int lb = load_face<MB>(vb, fb, nbm, jj[k], b);
  CT_CLIP_ROWS((clip_row_cap<4, MB>()), row);
  const double ar = clip_area<4, MB>(a, 4, b, lb, row);
  area[k] = ar;
  if (keep) keep[k] = ar > 0 ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_compact(const int64_t* __restrict__ ii,
                                                    const int64_t* __restrict__ jj,
                                                    const double* __restrict__ area,
                                                    const int64_t* __restrict__ keep,
                                                    const int64_t* __restrict__ offsets, int64_t n,
                                                    int64_t* __restrict__ io, int64_t* __restrict__ jo,
                                                    double* __restrict__ ao) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !keep[k]) return;
  const int64_t o = offsets[k];
  io[o] = ii[k];
  jo[o] = jj[k];
  if (area && ao) ao[o] = area[k];
}

inline int vclass(int nmax) { return nmax <= 4 ? 4 : nmax <= 8 ? 8 : nmax <= 16 ? 16 : 32; }

// Dispatches a pair kernel over the (subject, clipper) vertex-count classes.
template <template <int, int> class K, typename... Args>
int dispatch_pairs(int na, int nb, int64_t n, hipStream_t st, Args... args) {
  const unsigned g = grid_for(n);
  const int A = vclass(na), B = vclass(nb);
#define CT_CASE(X, Y)                                                     \
  if (A == X && B == Y) {                                                 \
    K<X, Y>::launch(g, st, args...);                                      \
    CT_LAUNCH_CHECK();                                                    \
    return CT_OK;                                                         \
  }
  CT_CASE(4, 4) CT_CASE(4, 8) CT_CASE(8, 4) CT_CASE(8, 8)
  CT_CASE(4, 16) CT_CASE(16, 4) CT_CASE(8, 16) CT_CASE(16, 8) CT_CASE(16, 16)
  CT_CASE(4, 32) CT_CASE(32, 4) CT_CASE(8, 32) CT_CASE(32, 8) CT_CASE(16, 32) CT_CASE(32, 16)
  CT_CASE(32, 32)
#undef CT_CASE
  set_error("unsupported vertex counts %d, %d", na, nb);
  return CT_ERR_ARG;
}

template <int X, int Y>
struct LFacePairs {
  template <typename... A>
  static void launch(unsigned g, hipStream_t st, A... a) {
    k_face_pairs<X, Y><<<g, kBlock, 0, st>>>(a...);
  }
};
template <int X, int Y>
struct LArea {
  template <typename... A>
  static void launch(unsigned g, hipStream_t st, A... a) {
    k_area<X, Y><<<g, kBlock, 0, st>>>(a...);
  }
};
template <int X, int Y>
struct LSat {
  template <typename... A>
  static void launch(unsigned g, hipStream_t st, A... a) {
    k_sat<X, Y><<<g, kBlock, 0, st>>>(a...);
  }
};

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_polygons_intersect(const double* vertices_a, const double* vertices_b, const int32_t* faces_a,
                          int32_t n_max_a, const int32_t* faces_b, int32_t n_max_b,
                          const int64_t* indices_a, const int64_t* indices_b, int64_t n, uint8_t* out,
                          void* stream) {
  CT_REQUIRE(n >= 0 && n_max_a >= 3 && n_max_b >= 3 && n_max_a <= 32 && n_max_b <= 32, "bad sizes");
  if (n == 0) return CT_OK;
  return dispatch_pairs<LSat>(n_max_a, n_max_b, n, as_stream(stream), vertices_a, faces_a, (int)n_max_a,
                              vertices_b, faces_b, (int)n_max_b, indices_a, indices_b, n, out);
}

int ct_area_of_intersection(const double* vertices_a, const double* vertices_b, const int32_t* faces_a,
                            int32_t n_max_a, const int32_t* faces_b, int32_t n_max_b,
                            const int64_t* indices_a, const int64_t* indices_b, int64_t n, double* out,
                            void* stream) {
  CT_REQUIRE(n >= 0 && n_max_a >= 3 && n_max_b >= 3 && n_max_a <= 32 && n_max_b <= 32, "bad sizes");
  if (n == 0) return CT_OK;
  return dispatch_pairs<LArea>(n_max_a, n_max_b, n, as_stream(stream), vertices_a, faces_a, (int)n_max_a,
                               vertices_b, faces_b, (int)n_max_b, indices_a, indices_b, n, out);
}

int ct_face_pairs(const double* vertices_a, const int32_t* faces_a, int32_t n_max_a,
                  const double* vertices_b, const int32_t* faces_b, int32_t n_max_b, const int64_t* ii,
                  const int64_t* jj, int64_t n, int32_t with_area, double* area, int64_t* keep,
                  void* stream) {
  CT_REQUIRE(n >= 0 && n_max_a >= 3 && n_max_b >= 3 && n_max_a <= 32 && n_max_b <= 32, "bad sizes");
  if (n == 0) return CT_OK;
  return dispatch_pairs<LFacePairs>(n_max_a, n_max_b, n, as_stream(stream), vertices_a, faces_a,
                                    (int)n_max_a, vertices_b, faces_b, (int)n_max_b, ii, jj, n,
                                    (int)with_area, area, keep);
}

int ct_box_area_of_intersection(const double* bbox_coords, const double* vertices, const int32_t* faces,
                                int32_t n_max_vert, const int64_t* indices_bbox,
                                const int64_t* indices_face, int64_t n, double* out, void* stream) {
  return ct_box_pairs(bbox_coords, vertices, faces, n_max_vert, indices_bbox, indices_face, n, out,
                      nullptr, stream);
}

int ct_box_pairs(const double* bbox_coords, const double* vertices, const int32_t* faces,
                 int32_t n_max_vert, const int64_t* ii, const int64_t* jj, int64_t n, double* area,
                 int64_t* keep, void* stream) {
  CT_REQUIRE(n >= 0 && n_max_vert >= 3 && n_max_vert <= 32, "bad sizes");
  if (n == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  const unsigned g = grid_for(n);
  switch (vclass(n_max_vert)) {
    case 4: k_box_pairs<4><<<g, kBlock, 0, st>>>(bbox_coords, vertices, faces, n_max_vert, ii, jj, n, area, keep); break;
    case 8: k_box_pairs<8><<<g, kBlock, 0, st>>>(bbox_coords, vertices, faces, n_max_vert, ii, jj, n, area, keep); break;
    case 16: k_box_pairs<16><<<g, kBlock, 0, st>>>(bbox_coords, vertices, faces, n_max_vert, ii, jj, n, area, keep); break;
    default: k_box_pairs<32><<<g, kBlock, 0, st>>>(bbox_coords, vertices, faces, n_max_vert, ii, jj, n, area, keep); break;
  }
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_compact_pairs(const int64_t* ii, const int64_t* jj, const double* area, const int64_t* keep,
                     const int64_t* offsets, int64_t n, int64_t* ii_out, int64_t* jj_out,
                     double* area_out, void* stream) {
  CT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return CT_OK;
  k_compact<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(ii, jj, area, keep, offsets, n, ii_out, jj_out,
                                                           area_out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // extern "C"
