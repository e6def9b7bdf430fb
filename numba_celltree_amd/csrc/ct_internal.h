// ct_internal.h -- shared internals of libcelltree_hip.so (gfx950).
//
// Every translation unit is compiled with -ffp-contract=off: each product and sum is
// a separately rounded IEEE operation, exactly as Numba compiles the reference, so the
// predicates below give bit-identical answers.  Association order follows the reference
// expression by expression (cited per function).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/celltree_hip.h"

namespace ct {

// ------------------------------------------------------------------ error state
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define CT_HIP(call)                                                       \
  do {                                                                     \
    hipError_t _e = (call);                                                \
    if (_e != hipSuccess) return ::ct::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define CT_LAUNCH_CHECK() CT_HIP(hipGetLastError())

#define CT_REQUIRE(cond, ...)          \
  do {                                 \
    if (!(cond)) {                     \
      ::ct::set_error(__VA_ARGS__);    \
      return CT_ERR_ARG;               \
    }                                  \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kBlock = 256;  // 4 wavefronts of 64 lanes

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// ------------------------------------------------------------------ constants
constexpr double kMinTolerance = 1e-15;     // constants.py:146
constexpr double kToleranceFactor = 1e-12;  // constants.py:147
constexpr int kFill = -1;                   // constants.py:135

// ------------------------------------------------------------------ FP64 keys
// Order-preserving map of a double to an unsigned 64-bit key: integer min/max
// atomics on keys compute exact double min/max (native 64-bit atomics on gfx950).
__device__ __forceinline__ unsigned long long dkey(double x) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double dval(unsigned long long k) {
  unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  __builtin_memcpy(&d, &u, sizeof(d));
  return d;
}

// ------------------------------------------------------------------ geometry
struct P2 {
  double x, y;
};

// Python's builtin min/max on floats keep the first argument on ties
// (Numba's do_minmax: select(b < a, b, a)).
__device__ __forceinline__ double pmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double pmax(double a, double b) { return (b > a) ? b : a; }

__device__ __forceinline__ P2 ld_vertex(const double* __restrict__ v, int i) {
  const double2 d = reinterpret_cast<const double2*>(v)[i];
  return {d.x, d.y};
}

// geometry_utils.py:58-59 cross_product(u, v) = u.x*v.y - u.y*v.x
__device__ __forceinline__ double cross(double ux, double uy, double vx, double vy) {
  return ux * vy - uy * vx;
}
// geometry_utils.py:63-64
__device__ __forceinline__ double dot(double ux, double uy, double vx, double vy) {
  return ux * vx + uy * vy;
}

// geometry_utils.py:150-166: (UxV*UxV) < ((tol*|U|^2)*tol)
__device__ __forceinline__ bool within_perp(double uxv, double ux, double uy, double tol) {
  return (uxv * uxv) < ((tol * (ux * ux + uy * uy)) * tol);
}

// geometry_utils.py:169-192
__device__ __forceinline__ bool in_bounds(P2 p, P2 a, P2 b) {
  const double xmin = pmin(a.x, b.x), xmax = pmax(a.x, b.x);
  const double ymin = pmin(a.y, b.y), ymax = pmax(a.y, b.y);
  const double dx = xmax - xmin, dy = ymax - ymin;
  const bool use_x = fabs(dx) >= fabs(dy);
  return (use_x && ((p.x >= xmin) && (p.x <= xmax))) ||
         (!use_x && ((p.y >= ymin) && (p.y <= ymax)));
}

// geometry_utils.py:195-223 point_in_polygon_or_on_edge.  Repeated vertices are
// skipped without advancing U / v0.  The crossing test divides only when the edge
// straddles p.y (short-circuit), as the reference does.
template <int MAXV>
__device__ __forceinline__ bool point_in_poly_or_edge(P2 p, const P2 (&poly)[MAXV], int len,
                                                      double tol) {
  P2 v0 = poly[0];
#pragma unroll
  for (int i = 0; i < MAXV; ++i)
    if (i == len - 1) v0 = poly[i];
  double ux = v0.x - p.x, uy = v0.y - p.y;
  bool c = false;
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    if (i < len) {
      const P2 v1 = poly[i];
      if (!(v1.x == v0.x && v1.y == v0.y)) {
        const double vx = v1.x - p.x, vy = v1.y - p.y;
        const double A = cross(ux, uy, vx, vy);
        const double wx = v1.x - v0.x, wy = v1.y - v0.y;
        if (within_perp(A, wx, wy, tol) && in_bounds(p, v0, v1)) return true;
        if ((v0.y > p.y) != (v1.y > p.y)) {
          if (p.x < ((v1.x - v0.x) * (p.y - v0.y) / (v1.y - v0.y) + v0.x)) c = !c;
        }
        v0 = v1;
        ux = vx;
        uy = vy;
      }
    }
  }
  return c;
}

// point_in_poly_or_edge on a row of exactly MAXV vertices whose unused tail repeats the last
// vertex (the packed leaf rows).  Same answer as the reference's loop, with no branches on
// the common path:
//  - a repeated vertex (v1 == v0) gives A = fl(ux*uy) - fl(uy*ux) = 0 and W = 0, so the
//    on-edge test fails, and v0.y == v1.y, so the edge does not straddle p.y; moving v0 to
//    an equal point changes nothing (signed zeros compare equal), so the reference's skip
//    is not needed, and the padded tail edges contribute nothing either;
//  - the reference returns True at the first on-edge hit and the parity otherwise, which
//    is (any on-edge) || parity;
//  - in_bounds only matters once within_perp holds (rare), so it sits behind that branch;
//  - the crossing abscissa is computed as the reference writes it, with a correctly
//    rounded division, and only used where the edge straddles p.y.
template <int MAXV>
__device__ __forceinline__ bool point_in_cell(P2 p, const P2 (&poly)[MAXV], double tol) {
  P2 v0 = poly[MAXV - 1];
  double ux = v0.x - p.x, uy = v0.y - p.y;
  bool c = false, on = false;
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    const P2 v1 = poly[i];
    const double vx = v1.x - p.x, vy = v1.y - p.y;
    const double A = cross(ux, uy, vx, vy);
    const double wx = v1.x - v0.x, wy = v1.y - v0.y;
    if (within_perp(A, wx, wy, tol)) on = on || in_bounds(p, v0, v1);
    const bool straddle = (v0.y > p.y) != (v1.y > p.y);
    const double xi = wx * (p.y - v0.y) / wy + v0.x;
    c = c != (straddle && (p.x < xi));
    v0 = v1;
    ux = vx;
    uy = vy;
  }
  return on || c;
}

// geometry_utils.py:226-238
__device__ __forceinline__ bool point_on_edge(P2 p, P2 v0, P2 v1, double tol) {
  if (v1.x == v0.x && v1.y == v0.y) return false;
  const double ux = v0.x - p.x, uy = v0.y - p.y;
  const double vx = v1.x - p.x, vy = v1.y - p.y;
  const double wx = v1.x - v0.x, wy = v1.y - v0.y;
  const double A = cross(ux, uy, vx, vy);
  return within_perp(A, wx, wy, tol) && in_bounds(p, v0, v1);
}

struct Box {
  double xmin, xmax, ymin, ymax;
};

__device__ __forceinline__ Box ld_box(const double* __restrict__ bb, int64_t i) {
  const double4 d = reinterpret_cast<const double4*>(bb)[i];
  return {d.x, d.y, d.z, d.w};
}

// geometry_utils.py:292-300 (strict)
__device__ __forceinline__ bool boxes_intersect(const Box& a, const Box& b) {
  return a.xmin < b.xmax && b.xmin < a.xmax && a.ymin < b.ymax && b.ymin < a.ymax;
}
// geometry_utils.py:303-315
__device__ __forceinline__ bool box_contained(const Box& a, const Box& b) {
  return a.xmin >= b.xmin && a.xmax <= b.xmax && a.ymin >= b.ymin && a.ymax <= b.ymax;
}

// geometry_utils.py:72-79 polygon_length: the first fill value at position >= 3.
__device__ __forceinline__ int poly_len(const int32_t* __restrict__ face, int nmax) {
  for (int i = 3; i < nmax; ++i)
    if (face[i] == kFill) return i;
  return nmax;
}

// Loads face `f` (row of width nmax) into registers; returns its length.
template <int MAXV>
__device__ __forceinline__ int load_poly(const int32_t* __restrict__ faces, int nmax, int64_t f,
                                         const double* __restrict__ vertices, P2 (&poly)[MAXV]) {
  const int32_t* row = faces + f * nmax;
  int idx[MAXV];
#pragma unroll
  for (int i = 0; i < MAXV; ++i) idx[i] = (i < nmax) ? row[i] : kFill;
  int len = nmax < MAXV ? nmax : MAXV;
#pragma unroll
  for (int i = MAXV - 1; i >= 3; --i)
    if (i < nmax && idx[i] == kFill) len = i;
#pragma unroll
  for (int i = 0; i < MAXV; ++i) poly[i] = (i < len) ? ld_vertex(vertices, idx[i]) : P2{0.0, 0.0};
  return len;
}

// geometry_utils.py:82-95 polygon_area: fan from vertex 0, abs per triangle.
template <int CAP>
__device__ __forceinline__ double polygon_area(const P2 (&poly)[CAP], int len) {
  const P2 a = poly[0];
  double ux = poly[1].x - a.x, uy = poly[1].y - a.y;
  double area = 0.0;
#pragma unroll
  for (int i = 2; i < CAP; ++i) {
    if (i < len) {
      const double vx = a.x - poly[i].x, vy = a.y - poly[i].y;
      area += fabs(cross(ux, uy, vx, vy));
      ux = vx;
      uy = vy;
    }
  }
  return 0.5 * area;
}

// ---------------------------------------------------------- Cohen-Sutherland
// cohen_sutherland.py:18-33
__device__ __forceinline__ int cs_code(P2 a, const Box& b) {
  int p = 0;
  if (a.x < b.xmin) p |= 1;
  else if (a.x > b.xmax) p |= 2;
  if (a.y < b.ymin) p |= 4;
  else if (a.y > b.ymax) p |= 8;
  return p;
}

// cohen_sutherland.py:36-101.  The reference's loop is unbounded; in exact arithmetic
// it ends within four clips.  The device loop stops after 64 (a clip that has not
// converged by then reports no intersection instead of spinning).
__device__ __forceinline__ bool cohen_sutherland(P2 a, P2 b, const Box& box, P2* c, P2* d) {
  double dx = b.x - a.x, dy = b.y - a.y;
  if (dx == 0.0 && dy == 0.0) return false;
  int k1 = cs_code(a, box), k2 = cs_code(b, box);
  for (int iter = 0; (k1 | k2) != 0; ++iter) {
    if ((k1 & k2) != 0 || iter >= 64) return false;
    const int opt = k1 ? k1 : k2;
    double x, y;
    if (opt & 8) {
      x = a.x + dx * (box.ymax - a.y) / dy;
      y = box.ymax;
    } else if (opt & 4) {
      x = a.x + dx * (box.ymin - a.y) / dy;
      y = box.ymin;
    } else if (opt & 2) {
      y = a.y + dy * (box.xmax - a.x) / dx;
      x = box.xmax;
    } else {
      y = a.y + dy * (box.xmin - a.x) / dx;
      x = box.xmin;
    }
    if (opt == k1) {
      a = {x, y};
      k1 = cs_code(a, box);
    } else if (opt == k2) {
      b = {x, y};
      k2 = cs_code(b, box);
    }
    dx = b.x - a.x;
    dy = b.y - a.y;
    if (dx == 0.0 && dy == 0.0) return false;
  }
  *c = a;
  *d = b;
  return true;
}

// ---------------------------------------------------------- Cyrus-Beck (Skala)
// cyrus_beck.py:35-52 (returns "entering", t by pointer)
__device__ __forceinline__ bool cb_t(P2 a, double sx, double sy, P2 v0, P2 v1, double* t) {
  const double six = v0.x - a.x, siy = v0.y - a.y;
  const double nx = -(v1.y - v0.y), ny = (v1.x - v0.x);
  const double n_si = dot(nx, ny, six, siy);
  const double k = dot(nx, ny, sx, sy);
  *t = n_si / k;
  return n_si > 0;
}

// cyrus_beck.py:85-100
__device__ __forceinline__ bool cb_aligned(double ux, double uy, double vx, double vy) {
  if ((ux == 0 && uy == 0) || (vx == 0 && vy == 0)) return true;
  if (ux != 0 && vx != 0) return (ux > 0) == (vx > 0);
  if (uy != 0 && vy != 0) return (uy > 0) == (vy > 0);
  return false;
}

// cyrus_beck.py:103-139 (collinear segment and polygon edge)
__device__ __forceinline__ bool cb_collinear(P2 a, P2 b, P2 v0, P2 v1, P2* c, P2* d) {
  const double bx = b.x - a.x, by = b.y - a.y;
  double v0x = v0.x - a.x, v0y = v0.y - a.y;
  double v1x = v1.x - a.x, v1y = v1.y - a.y;
  if (!cb_aligned(bx, by, v1x - v0x, v1y - v0y)) {
    const P2 t = v0;
    v0 = v1;
    v1 = t;
    double tx = v0x, ty = v0y;
    v0x = v1x;
    v0y = v1y;
    v1x = tx;
    v1y = ty;
  }
  const double nx = -by, ny = bx;
  double ta = 0.0;
  double tb = cross(nx, ny, bx, by);
  double t0 = cross(nx, ny, v0x, v0y);
  double t1 = cross(nx, ny, v1x, v1y);
  // cyrus_beck.py:75-82 overlap (on copies; the originals decide p0/p1 below)
  double oa = ta, ob = tb, o0 = t0, o1 = t1;
  if (oa > ob) { const double s = oa; oa = ob; ob = s; }
  if (o0 > o1) { const double s = o0; o0 = o1; o1 = s; }
  const double x = pmin(ob, o1) - pmax(oa, o0);
  const double ov = (x > 0) ? x : 0.0;
  if (!(ov > 0.0)) return false;
  *c = (t0 < ta) ? v0 : a;
  *d = (t1 > tb) ? v1 : b;
  return true;
}

// cyrus_beck.py:142-241 cyrus_beck_line_polygon_clip (polygon CCW).
template <int MAXV>
__device__ __forceinline__ bool cyrus_beck(P2 a, P2 b, const P2 (&poly)[MAXV], int len, double tol,
                                           P2* c, P2* d) {
  const double sx = b.x - a.x, sy = b.y - a.y;
  if (sx == 0 && sy == 0) return false;
  const bool a_in = point_in_poly_or_edge<MAXV>(a, poly, len, tol);
  const bool b_in = point_in_poly_or_edge<MAXV>(b, poly, len, tol);
  if (a_in && b_in) {
    *c = a;
    *d = b;
    return true;
  }
  int i0 = -1, i1 = -1, k = 0;
  double ksi = cross(poly[0].x - a.x, poly[0].y - a.y, sx, sy);
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    if (i < len && k < 2) {
      const P2 v0 = poly[i];
      P2 v1 = poly[0];
#pragma unroll
      for (int j = 0; j < MAXV; ++j)
        if (j == i + 1 && j < len) v1 = poly[j];
      const double eta = cross(v1.x - a.x, v1.y - a.y, sx, sy);
      if ((ksi < 0.0) != (eta < 0.0)) {
        if (k == 0) i0 = i;
        else i1 = i;
        k += 1;
      } else if ((ksi == 0.0) && (eta == 0.0)) {
        return cb_collinear(a, b, v0, v1, c, d);
      }
      ksi = eta;
    }
  }
  if (k == 0) return false;
  // cyrus_beck.py:55-72 intersections; with k == 1, i1 == -1 reads the last vertex
  // (Python's poly[-1]) and its successor poly[0].
  const int j0 = i0, j1 = (i1 < 0) ? len - 1 : i1;
  const int j0n = (j0 + 1 == len) ? 0 : j0 + 1;
  const int j1n = (j1 + 1 == len) ? 0 : j1 + 1;
  P2 v0 = poly[0], v01 = poly[0], v1 = poly[0], v11 = poly[0];
#pragma unroll
  for (int j = 0; j < MAXV; ++j) {
    if (j == j0) v0 = poly[j];
    if (j == j0n) v01 = poly[j];
    if (j == j1) v1 = poly[j];
    if (j == j1n) v11 = poly[j];
  }
  double ta, tb;
  cb_t(a, sx, sy, v0, v01, &ta);
  const bool enters1 = cb_t(a, sx, sy, v1, v11, &tb);
  double t0 = enters1 ? tb : ta;
  double t1 = enters1 ? ta : tb;
  if (t0 == t1) {
    if (a_in && t1 != 0.0) t0 = 0.0;
    else if (b_in && t0 != 1.0) t1 = 1.0;
    else return false;
  }
  if (t1 < t0) {
    const double s = t0;
    t0 = t1;
    t1 = s;
  }
  const bool valid0 = t0 >= 0 && t0 < 1;
  const bool valid1 = t1 > 0 && t1 <= 1;
  if (valid0 && valid1) {
    *c = {a.x + t0 * sx, a.y + t0 * sy};
    *d = {a.x + t1 * sx, a.y + t1 * sy};
    return true;
  } else if (valid0) {
    *c = {a.x + t0 * sx, a.y + t0 * sy};
    *d = b;
    return true;
  } else if (valid1) {
    *c = a;
    *d = {a.x + t1 * sx, a.y + t1 * sy};
    return true;
  }
  return false;
}

// ---------------------------------------------------------- segment x segment
// geometry_utils.py:318-329
__device__ __forceinline__ bool left_of(P2 a, P2 p, double ux, double uy) {
  return ux * (a.y - p.y) > uy * (a.x - p.x);
}
__device__ __forceinline__ bool has_overlap(double a, double b, double p, double q, double tol) {
  return ((pmin(a, b) - pmax(p, q)) < tol) && ((pmin(p, q) - pmax(a, b)) < tol);
}

// geometry_utils.py:377-418 lines_intersect (a->b against p->q); out = point or NaN.
__device__ __forceinline__ bool lines_intersect(P2 a, P2 b, P2 p, P2 q, P2* out) {
  const double nan = __builtin_nan("");
  *out = {nan, nan};
  const double vx = b.x - a.x, vy = b.y - a.y;
  const double ux = q.x - p.x, uy = q.y - p.y;
  if ((ux == 0 && uy == 0) || (vx == 0 && vy == 0)) return false;
  if ((ux == 0) && (vx == 0) && a.x != p.x) return false;
  if ((uy == 0) && (vy == 0) && a.y != p.y) return false;
  const double tol = pmax(kMinTolerance, kToleranceFactor * pmax(fabs(ux), fabs(uy)));
  if ((!has_overlap(a.x, b.x, p.x, q.x, tol)) || (!has_overlap(a.y, b.y, p.y, q.y, tol)))
    return false;
  if ((left_of(a, p, ux, uy) != left_of(b, p, ux, uy)) &&
      (left_of(p, a, vx, vy) != left_of(q, a, vx, vy))) {
    // geometry_utils.py:332-345 intersection_location_point
    const double denom = cross(vx, vy, ux, uy);
    if (!within_perp(denom, vx, vy, tol)) {
      const double t = cross(p.x - a.x, p.y - a.y, ux, uy) / denom;
      *out = {a.x + t * vx, a.y + t * vy};
    }
    return true;
  }
  const double rx = p.x - a.x, ry = p.y - a.y;
  const double sx = q.x - a.x, sy = q.y - a.y;
  if (within_perp(cross(vx, vy, rx, ry), vx, vy, tol) &&
      within_perp(cross(vx, vy, sx, sy), vx, vy, tol)) {
    // geometry_utils.py:348-374 midpoint_collinear_lines
    P2 a2 = a, b2 = b, p2 = p, q2 = q;
    if (a2.x > b2.x || (a2.x == b2.x && a2.y > b2.y)) { const P2 t = a2; a2 = b2; b2 = t; }
    if (p2.x > q2.x || (p2.x == q2.x && p2.y > q2.y)) { const P2 t = p2; p2 = q2; q2 = t; }
    const double ox0 = pmax(a2.x, p2.x), oy0 = pmax(a2.y, p2.y);
    const double ox1 = pmin(b2.x, q2.x), oy1 = pmin(b2.y, q2.y);
    if (!(ox0 > ox1 || oy0 > oy1)) *out = {0.5 * (ox0 + ox1), 0.5 * (oy0 + oy1)};
    return true;
  }
  return false;
}

}  // namespace ct
