// refcpu.cpp -- CPU oracle for the cell-tree hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a line-by-line restatement, in C++17,
// of the arithmetic of numba_celltree v0.4.2 (pure Python compiled by Numba with
// IEEE semantics and no FMA contraction).  It exists so that tests, the smoke
// check and the CPU-baseline leg of bench.py can check the HIP library against
// the reference algorithm.  Nothing in the product package links, loads or calls
// it.  Citations are "file:line" relative to the numba_celltree source tree.
//
// Build: oracle/build_oracle.py  (g++ -O3 -fopenmp -ffp-contract=off ...).
// -ffp-contract=off is essential: gcc contracts a*b+c to FMA by default, which
// would change the rounding of every cross product and break bit-exactness.
//
// Conventions kept from the reference:
//   * indices are int64 (np.intp, constants.py:27), fill value -1 (constants.py:135);
//   * Python's builtin min/max on two floats: min(a,b) = (b < a) ? b : a,
//     max(a,b) = (b > a) ? b : a (Numba's do_minmax keeps the first argument on ties);
//   * every expression keeps the reference's association order.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double MIN_TOLERANCE = 1e-15;     // constants.py:146
constexpr double TOLERANCE_FACTOR = 1e-12;  // constants.py:147
constexpr double FLOAT_MAX = std::numeric_limits<double>::max();   // constants.py:150
constexpr double FLOAT_MIN = -std::numeric_limits<double>::max();  // np.finfo.min, constants.py:149
constexpr int64_t INT_MAX64 = std::numeric_limits<int64_t>::max(); // constants.py:151
constexpr int64_t FILL_VALUE = -1;          // constants.py:135
constexpr int MAX_N_VERTEX = 32;            // constants.py:134
constexpr int CLIP_MAX_N_VERTEX = 2 * MAX_N_VERTEX;  // utils.py:99
const double NaN = std::numeric_limits<double>::quiet_NaN();

struct Point { double x, y; };
struct Vector { double x, y; };
struct Box { double xmin, xmax, ymin, ymax; };

inline double pymin(double a, double b) { return (b < a) ? b : a; }
inline double pymax(double a, double b) { return (b > a) ? b : a; }
inline bool peq(Point a, Point b) { return a.x == b.x && a.y == b.y; }

// geometry_utils.py:24-25
inline Vector to_vector(Point a, Point b) { return {b.x - a.x, b.y - a.y}; }
// geometry_utils.py:53-54
inline Point to_point(double t, Point a, Vector V) { return {a.x + t * V.x, a.y + t * V.y}; }
// geometry_utils.py:58-59
inline double cross_product(Vector u, Vector v) { return u.x * v.y - u.y * v.x; }
// geometry_utils.py:63-64
inline double dot_product(Vector u, Vector v) { return u.x * v.x + u.y * v.y; }
// geometry_utils.py:68-69
inline double length_squared(Vector v) { return v.x * v.x + v.y * v.y; }

// geometry_utils.py:72-79: a face has at least 3 vertices; the first fill value
// at position >= 3 ends it.
inline int polygon_length(const int64_t* face, int n) {
  for (int i = 3; i < n; ++i)
    if (face[i] == FILL_VALUE) return i;
  return n;
}

// FP64 operation count of the executed path, for the roofline of the clip kernels: every
// add, subtract, multiply and divide of separating_axes, polygon_polygon_clip_area and
// polygon_area (compares, abs, min / max and sign flips are not counted).  Functions take
// COUNT = false by default, which compiles the counting out; ref_face_pair_flops runs them
// with COUNT = true, one pair at a time per thread.
thread_local int64_t t_flops = 0;
template <bool COUNT>
inline void flops(int64_t k) {
  if constexpr (COUNT) t_flops += k;
}

// geometry_utils.py:82-95 (fan triangulation, abs per triangle)
template <bool COUNT = false>
double polygon_area(const Point* poly, int length) {
  double area = 0.0;
  Point a = poly[0];
  Point b = poly[1];
  Vector U = to_vector(a, b);
  flops<COUNT>(2);
  for (int i = 2; i < length; ++i) {
    Point c = poly[i];
    Vector V = to_vector(c, a);
    area += std::fabs(cross_product(U, V));
    flops<COUNT>(2 + 3 + 1);  // V, the cross product, the sum
    b = c;
    U = V;
  }
  (void)b;
  flops<COUNT>(1);
  return 0.5 * area;
}

// geometry_utils.py:150-166
inline bool within_perpendicular_distance(double UxV, Vector U, double tol) {
  return (UxV * UxV) < ((tol * length_squared(U)) * tol);
}

// geometry_utils.py:169-192
inline bool in_bounds(Point p, Point a, Point b) {
  double xmin = pymin(a.x, b.x);
  double xmax = pymax(a.x, b.x);
  double ymin = pymin(a.y, b.y);
  double ymax = pymax(a.y, b.y);
  double dx = xmax - xmin;
  double dy = ymax - ymin;
  bool use_x_bound = std::fabs(dx) >= std::fabs(dy);
  return (use_x_bound && ((p.x >= xmin) && (p.x <= xmax))) ||
         (!use_x_bound && ((p.y >= ymin) && (p.y <= ymax)));
}

// geometry_utils.py:195-223
bool point_in_polygon_or_on_edge(Point p, const Point* poly, int length, double tol) {
  Point v0 = poly[length - 1];
  Vector U = to_vector(p, v0);
  bool c = false;
  for (int i = 0; i < length; ++i) {
    Point v1 = poly[i];
    if (peq(v1, v0)) continue;
    Vector V = to_vector(p, v1);
    double A = cross_product(U, V);
    Vector W = to_vector(v0, v1);
    if (within_perpendicular_distance(A, W, tol) && in_bounds(p, v0, v1)) return true;
    if (((v0.y > p.y) != (v1.y > p.y)) &&
        p.x < ((v1.x - v0.x) * (p.y - v0.y) / (v1.y - v0.y) + v0.x))
      c = !c;
    v0 = v1;
    U = V;
  }
  return c;
}

// geometry_utils.py:226-238
bool point_on_edge(Point p, Point v0, Point v1, double tol) {
  if (peq(v1, v0)) return false;
  Vector U = to_vector(p, v0);
  Vector V = to_vector(p, v1);
  Vector W = to_vector(v0, v1);
  double A = cross_product(U, V);
  return within_perpendicular_distance(A, W, tol) && in_bounds(p, v0, v1);
}

// geometry_utils.py:292-300 (strict)
inline bool boxes_intersect(Box a, Box b) {
  return a.xmin < b.xmax && b.xmin < a.xmax && a.ymin < b.ymax && b.ymin < a.ymax;
}

// geometry_utils.py:303-315
inline bool box_contained(Box a, Box b) {
  return a.xmin >= b.xmin && a.xmax <= b.xmax && a.ymin >= b.ymin && a.ymax <= b.ymax;
}

// geometry_utils.py:318-322
inline bool left_of(Point a, Point p, Vector U) {
  return U.x * (a.y - p.y) > U.y * (a.x - p.x);
}

// geometry_utils.py:325-329
inline bool has_overlap(double a, double b, double p, double q, double tol) {
  return ((pymin(a, b) - pymax(p, q)) < tol) && ((pymin(p, q) - pymax(a, b)) < tol);
}

// geometry_utils.py:332-345
inline Point intersection_location_point(Vector V, Vector U, Point a, Point p, double tol) {
  double denom = cross_product(V, U);
  if (within_perpendicular_distance(denom, V, tol)) return {NaN, NaN};
  Vector R = to_vector(a, p);
  double t = cross_product(R, U) / denom;
  return {a.x + t * V.x, a.y + t * V.y};
}

// geometry_utils.py:348-374
inline Point midpoint_collinear_lines(Point a, Point b, Point p, Point q) {
  if (a.x > b.x || (a.x == b.x && a.y > b.y)) std::swap(a, b);
  if (p.x > q.x || (p.x == q.x && p.y > q.y)) std::swap(p, q);
  double sx = pymax(a.x, p.x);
  double sy = pymax(a.y, p.y);
  double ex = pymin(b.x, q.x);
  double ey = pymin(b.y, q.y);
  if (sx > ex || sy > ey) return {NaN, NaN};
  return {0.5 * (sx + ex), 0.5 * (sy + ey)};
}

// geometry_utils.py:377-418
bool lines_intersect(Point a, Point b, Point p, Point q, Point* out) {
  Vector V = to_vector(a, b);
  Vector U = to_vector(p, q);
  *out = {NaN, NaN};
  if ((U.x == 0 && U.y == 0) || (V.x == 0 && V.y == 0)) return false;
  if ((U.x == 0) && (V.x == 0) && a.x != p.x) return false;
  if ((U.y == 0) && (V.y == 0) && a.y != p.y) return false;
  double tol = pymax(MIN_TOLERANCE, TOLERANCE_FACTOR * pymax(std::fabs(U.x), std::fabs(U.y)));
  if ((!has_overlap(a.x, b.x, p.x, q.x, tol)) || (!has_overlap(a.y, b.y, p.y, q.y, tol)))
    return false;
  if ((left_of(a, p, U) != left_of(b, p, U)) && (left_of(p, a, V) != left_of(q, a, V))) {
    *out = intersection_location_point(V, U, a, p, tol);
    return true;
  }
  Vector R = to_vector(a, p);
  Vector S = to_vector(a, q);
  if (within_perpendicular_distance(cross_product(V, R), V, tol) &&
      within_perpendicular_distance(cross_product(V, S), V, tol)) {
    *out = midpoint_collinear_lines(a, b, p, q);
    return true;
  }
  return false;
}

// geometry_utils.py:421-440
inline Box bounding_box(const int64_t* polygon, int max_n_verts, const double* vertices) {
  const double* fv = vertices + 2 * polygon[0];
  double xmin = fv[0], xmax = fv[0], ymin = fv[1], ymax = fv[1];
  for (int i = 1; i < max_n_verts; ++i) {
    int64_t index = polygon[i];
    if (index == FILL_VALUE) break;
    double x = vertices[2 * index], y = vertices[2 * index + 1];
    xmin = pymin(xmin, x);
    xmax = pymax(xmax, x);
    ymin = pymin(ymin, y);
    ymax = pymax(ymax, y);
  }
  return {xmin, xmax, ymin, ymax};
}

// geometry_utils.py:459-470
inline Box edge_bounding_box(const int64_t* edge, const double* vertices, double tol) {
  double x0 = vertices[2 * edge[0]], y0 = vertices[2 * edge[0] + 1];
  double x1 = vertices[2 * edge[1]], y1 = vertices[2 * edge[1] + 1];
  return {pymin(x0 - tol, x1 - tol), pymax(x0 + tol, x1 + tol),
          pymin(y0 - tol, y1 - tol), pymax(y0 + tol, y1 + tol)};
}

// geometry_utils.py:501-510 (copy_vertices_into / copy_vertices)
inline int copy_vertices(const double* vertices, const int64_t* face, int n_max, Point* out) {
  int length = polygon_length(face, n_max);
  for (int i = 0; i < length; ++i) out[i] = {vertices[2 * face[i]], vertices[2 * face[i] + 1]};
  return length;
}

// geometry_utils.py:532-538
inline void flip(int64_t* face, int length) {
  int end = length - 1;
  for (int i = 0; i < length / 2; ++i) std::swap(face[i], face[end - i]);
}

// geometry_utils.py:541-561: the first non-zero cross product decides; after a
// flip the loop continues with the *original* a, b (no break).
void counter_clockwise_row(const double* vertices, int64_t* face, int n_max) {
  int length = polygon_length(face, n_max);
  Point a = {vertices[2 * face[length - 2]], vertices[2 * face[length - 2] + 1]};
  Point b = {vertices[2 * face[length - 1]], vertices[2 * face[length - 1] + 1]};
  for (int i = 0; i < length; ++i) {
    Point c = {vertices[2 * face[i]], vertices[2 * face[i] + 1]};
    Vector u = to_vector(a, b);
    Vector v = to_vector(a, c);
    double product = cross_product(u, v);
    if (product == 0) {
      a = b;
      b = c;
    } else if (product < 0) {
      flip(face, length);
    } else {
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// algorithms/cohen_sutherland.py
// ---------------------------------------------------------------------------
constexpr int CS_INSIDE = 0, CS_LEFT = 1, CS_RIGHT = 2, CS_LOWER = 4, CS_UPPER = 8;  // :12

// cohen_sutherland.py:18-33
inline int get_clip(Point a, Box box) {
  int p = CS_INSIDE;
  if (a.x < box.xmin) p |= CS_LEFT;
  else if (a.x > box.xmax) p |= CS_RIGHT;
  if (a.y < box.ymin) p |= CS_LOWER;
  else if (a.y > box.ymax) p |= CS_UPPER;
  return p;
}

// cohen_sutherland.py:36-101
bool cohen_sutherland_line_box_clip(Point a, Point b, Box box, Point* c, Point* d) {
  *c = {NaN, NaN};
  *d = {NaN, NaN};
  double dx = b.x - a.x;
  double dy = b.y - a.y;
  if (dx == 0.0 && dy == 0.0) return false;
  int k1 = get_clip(a, box);
  int k2 = get_clip(b, box);
  while ((k1 | k2) != CS_INSIDE) {
    if ((k1 & k2) != 0) return false;
    int opt = k1 ? k1 : k2;
    double x, y;
    if (opt & CS_UPPER) {
      x = a.x + dx * (box.ymax - a.y) / dy;
      y = box.ymax;
    } else if (opt & CS_LOWER) {
      x = a.x + dx * (box.ymin - a.y) / dy;
      y = box.ymin;
    } else if (opt & CS_RIGHT) {
      y = a.y + dy * (box.xmax - a.x) / dx;
      x = box.xmax;
    } else {  // opt & CS_LEFT; the "Undefined clipping state" branch is unreachable
      y = a.y + dy * (box.xmin - a.x) / dx;
      x = box.xmin;
    }
    if (opt == k1) {
      a = {x, y};
      k1 = get_clip(a, box);
    } else if (opt == k2) {
      b = {x, y};
      k2 = get_clip(b, box);
    }
    dx = b.x - a.x;
    dy = b.y - a.y;
    if (dx == 0.0 && dy == 0.0) return false;
  }
  *c = a;
  *d = b;
  return true;
}

// ---------------------------------------------------------------------------
// algorithms/cyrus_beck.py
// ---------------------------------------------------------------------------

// cyrus_beck.py:35-52
inline bool cb_compute_intersection(Point a, Vector s, Point v0, Point v1, double* t) {
  Vector si = to_vector(a, v0);
  Vector n = {-(v1.y - v0.y), (v1.x - v0.x)};
  double n_si = dot_product(n, si);
  double k = dot_product(n, s);
  *t = n_si / k;
  return n_si > 0;
}

// cyrus_beck.py:55-72.  With k == 1, i1 == -1 and poly[-1] is the last vertex.
inline void cb_intersections(Point a, Vector s, const Point* poly, int length, int i0, int i1,
                             double* t0, double* t1) {
  int j0 = (i0 < 0) ? i0 + length : i0;
  int j1 = (i1 < 0) ? i1 + length : i1;
  Point v0 = poly[j0];
  Point v01 = poly[(((i0 + 1) % length) + length) % length];
  Point v1 = poly[j1];
  Point v11 = poly[(((i1 + 1) % length) + length) % length];
  double ta, tb;
  cb_compute_intersection(a, s, v0, v01, &ta);
  bool enters1 = cb_compute_intersection(a, s, v1, v11, &tb);
  if (enters1) {
    *t0 = tb;
    *t1 = ta;
  } else {
    *t0 = ta;
    *t1 = tb;
  }
}

// cyrus_beck.py:75-82
inline bool cb_overlap(double ta, double tb, double t0, double t1) {
  if (ta > tb) std::swap(ta, tb);
  if (t0 > t1) std::swap(t0, t1);
  double x = pymin(tb, t1) - pymax(ta, t0);
  double vector_overlap = (x > 0) ? x : 0.0;
  return vector_overlap > 0.0;
}

// cyrus_beck.py:85-100
inline bool cb_aligned(Vector U, Vector V) {
  if ((U.x == 0 && U.y == 0) || (V.x == 0 && V.y == 0)) return true;
  if (U.x != 0 && V.x != 0) return (U.x > 0) == (V.x > 0);
  if (U.y != 0 && V.y != 0) return (U.y > 0) == (V.y > 0);
  return false;
}

// cyrus_beck.py:103-139
bool cb_collinear_case(Point a, Point b, Point v0, Point v1, Point* c, Point* d) {
  Point _b = {b.x - a.x, b.y - a.y};
  Point _v0 = {v0.x - a.x, v0.y - a.y};
  Point _v1 = {v1.x - a.x, v1.y - a.y};
  Vector U = {_b.x, _b.y};
  Vector V = to_vector(_v0, _v1);
  if (!cb_aligned(U, V)) {
    std::swap(v0, v1);
    std::swap(_v0, _v1);
  }
  Vector n = {-_b.y, _b.x};
  double ta = 0.0;
  double tb = cross_product(n, {_b.x, _b.y});
  double t0 = cross_product(n, {_v0.x, _v0.y});
  double t1 = cross_product(n, {_v1.x, _v1.y});
  if (!cb_overlap(ta, tb, t0, t1)) {
    *c = {NaN, NaN};
    *d = {NaN, NaN};
    return false;
  }
  *c = (t0 < ta) ? v0 : a;
  *d = (t1 > tb) ? v1 : b;
  return true;
}

// cyrus_beck.py:142-241
bool cyrus_beck_line_polygon_clip(Point a, Point b, const Point* poly, int length, double tol,
                                  Point* c, Point* d) {
  *c = {NaN, NaN};
  *d = {NaN, NaN};
  Vector s = to_vector(a, b);
  if (s.x == 0 && s.y == 0) return false;
  bool a_inside = point_in_polygon_or_on_edge(a, poly, length, tol);
  bool b_inside = point_in_polygon_or_on_edge(b, poly, length, tol);
  if (a_inside && b_inside) {
    *c = a;
    *d = b;
    return true;
  }
  int i0 = -1, i1 = -1, i = 0, k = 0;
  Point v = poly[i];
  double ksi = cross_product(to_vector(a, v), s);
  while (i < length && k < 2) {
    Point v0 = poly[i];
    Point v1 = poly[(i + 1) % length];
    double eta = cross_product(to_vector(a, v1), s);
    if ((ksi < 0.0) ^ (eta < 0.0)) {
      if (k == 0) i0 = i;
      else i1 = i;
      k += 1;
    } else if ((ksi == 0.0) && (eta == 0.0)) {
      return cb_collinear_case(a, b, v0, v1, c, d);
    }
    ksi = eta;
    i += 1;
  }
  if (k == 0) return false;
  double t0, t1;
  cb_intersections(a, s, poly, length, i0, i1, &t0, &t1);
  if (t0 == t1) {
    if (a_inside && t1 != 0.0) t0 = 0.0;
    else if (b_inside && t0 != 1.0) t1 = 1.0;
    else return false;
  }
  if (t1 < t0) std::swap(t0, t1);
  bool valid0 = t0 >= 0 && t0 < 1;
  bool valid1 = t1 > 0 && t1 <= 1;
  if (valid0 && valid1) {
    *c = to_point(t0, a, s);
    *d = to_point(t1, a, s);
    return true;
  } else if (valid0) {
    *c = to_point(t0, a, s);
    *d = b;
    return true;
  } else if (valid1) {
    *c = a;
    *d = to_point(t1, a, s);
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// algorithms/separating_axis.py
// ---------------------------------------------------------------------------

// separating_axis.py:17-27
template <bool COUNT = false>
inline void extrema_projected(Vector norm, const Point* poly, int length, double* mn, double* mx) {
  flops<COUNT>(3 * (int64_t)length);  // one dot product per vertex
  double min_proj = FLOAT_MAX, max_proj = FLOAT_MIN;
  for (int i = 0; i < length; ++i) {
    double proj = dot_product({poly[i].x, poly[i].y}, norm);
    min_proj = pymin(min_proj, proj);
    max_proj = pymax(max_proj, proj);
  }
  *mn = min_proj;
  *mx = max_proj;
}

// separating_axis.py:30-39
template <bool COUNT = false>
inline bool is_separating_axis(Vector norm, const Point* a, const Point* b, int la, int lb) {
  double mina, maxa, minb, maxb;
  extrema_projected<COUNT>(norm, a, la, &mina, &maxa);
  extrema_projected<COUNT>(norm, b, lb, &minb, &maxb);
  return !(maxa > minb && maxb > mina);
}

// separating_axis.py:42-55
template <bool COUNT = false>
bool separating_axes(const Point* a, int la, const Point* b, int lb) {
  Point p = a[la - 1];
  for (int i = 0; i < la; ++i) {
    Point q = a[i];
    Vector norm = {p.y - q.y, q.x - p.x};
    flops<COUNT>(2);
    p = q;
    if (norm.x == 0.0 && norm.y == 0.0) continue;
    if (is_separating_axis<COUNT>(norm, a, b, la, lb)) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// algorithms/sutherland_hodgman.py
// ---------------------------------------------------------------------------

// sutherland_hodgman.py:56-60 (strict)
template <bool COUNT = false>
inline bool sh_inside(Point p, Point r, Vector U) {
  flops<COUNT>(4);
  return U.x * (p.y - r.y) > U.y * (p.x - r.x);
}

// sutherland_hodgman.py:63-74
template <bool COUNT = false>
inline bool sh_intersection(Point a, Vector V, Point r, Vector N, Point* out) {
  Vector W = {r.x - a.x, r.y - a.y};
  double nw = dot_product(N, W);
  double nv = dot_product(N, V);
  flops<COUNT>(2 + 3 + 3);
  if (nv != 0) {
    flops<COUNT>(1 + 4);  // t, then the point
    double t = nw / nv;
    *out = {a.x + t * V.x, a.y + t * V.y};
    return true;
  }
  *out = {NaN, NaN};
  return false;
}

// sutherland_hodgman.py:84-148.  subject = polygon, clipper = clipper.
template <bool COUNT = false>
double polygon_polygon_clip_area(const Point* polygon, int n_polygon, const Point* clipper,
                                 int n_clip) {
  Point subject[CLIP_MAX_N_VERTEX];
  Point output[CLIP_MAX_N_VERTEX];
  int n_output = n_polygon;
  for (int i = 0; i < n_polygon; ++i) output[i] = polygon[i];
  Point r = clipper[n_clip - 1];
  for (int i = 0; i < n_clip; ++i) {
    Point s = clipper[i];
    Vector U = {s.x - r.x, s.y - r.y};
    flops<COUNT>(2);
    if (U.x == 0 && U.y == 0) continue;
    Vector N = {-U.y, U.x};
    int length = n_output;
    for (int j = 0; j < length; ++j) subject[j] = output[j];
    n_output = 0;
    Point a = subject[length - 1];
    bool a_inside = sh_inside<COUNT>(a, r, U);
    for (int j = 0; j < length; ++j) {
      Point b = subject[j];
      Vector V = {b.x - a.x, b.y - a.y};
      flops<COUNT>(2);
      if (V.x == 0 && V.y == 0) continue;
      bool b_inside = sh_inside<COUNT>(b, r, U);
      if (b_inside) {
        if (!a_inside) {
          Point pt;
          if (sh_intersection<COUNT>(a, V, r, N, &pt) && n_output < CLIP_MAX_N_VERTEX)
            output[n_output++] = pt;
        }
        if (n_output < CLIP_MAX_N_VERTEX) output[n_output++] = b;
      } else if (a_inside) {
        Point pt;
        if (sh_intersection<COUNT>(a, V, r, N, &pt)) {
          if (n_output < CLIP_MAX_N_VERTEX) output[n_output++] = pt;
        } else {
          b_inside = true;
          if (n_output < CLIP_MAX_N_VERTEX) output[n_output++] = b;
        }
      }
      a = b;
      a_inside = b_inside;
    }
    if (n_output < 3) return 0.0;
    r = s;
  }
  return polygon_area<COUNT>(output, n_output);
}

// geometry_utils.py:513-524
inline void copy_box_vertices(Box box, Point* a) {
  a[0] = {box.xmin, box.ymin};
  a[1] = {box.xmax, box.ymin};
  a[2] = {box.xmax, box.ymax};
  a[3] = {box.xmin, box.ymax};
}

// ---------------------------------------------------------------------------
// algorithms/barycentric_*.py
// ---------------------------------------------------------------------------

// barycentric_triangle.py:28-43
inline void bary_triangle(Point ta, Point tb, Point tc, Point p, double* w) {
  Vector ab = to_vector(ta, tb);
  Vector ac = to_vector(ta, tc);
  Vector ap = to_vector(ta, p);
  double Aa = std::fabs(cross_product(ab, ap));
  double Ac = std::fabs(cross_product(ac, ap));
  double A = std::fabs(cross_product(ab, ac));
  double inv_denom = 1.0 / A;
  double ww = inv_denom * Aa;
  double v = inv_denom * Ac;
  double u = 1.0 - v - ww;
  w[0] = u;
  w[1] = v;
  w[2] = ww;
}

// barycentric_wachspress.py:26-35 (weights has n_w columns)
inline void interp_edge_case(Point a, Vector U, Point p, double* weights, int n_w, int i, int j) {
  for (int k = 0; k < n_w; ++k) weights[k] = 0.0;
  Vector V = to_vector(a, p);
  double w = std::sqrt(dot_product(V, V)) / std::sqrt(dot_product(U, U));
  weights[i] = 1.0 - w;
  weights[j] = w;
}

// barycentric_wachspress.py:38-85
void wachspress_weights(const Point* polygon, int n, Point p, double* weights, int n_w,
                        double tol) {
  double w_sum = 0.0;
  Point a = polygon[n - 1];
  Point b = polygon[0];
  Vector U = to_vector(a, b);
  Vector V = to_vector(a, p);
  double Ai = std::fabs(cross_product(U, V));
  if (within_perpendicular_distance(Ai, U, tol)) {
    interp_edge_case(a, U, p, weights, n_w, n - 1, 0);
    return;
  }
  for (int i = 0; i < n; ++i) {
    int i_next = (i + 1) % n;
    Point c = polygon[i_next];
    Vector W = to_vector(a, c);
    double Ci = std::fabs(cross_product(U, W));
    U = to_vector(b, c);
    V = to_vector(b, p);
    double Aj = std::fabs(cross_product(U, V));
    if (within_perpendicular_distance(Aj, U, tol)) {
      interp_edge_case(b, U, p, weights, n_w, i, i_next);
      return;
    }
    double w = 2 * Ci / (Ai * Aj);
    weights[i] = w;
    w_sum += w;
    a = b;
    b = c;
    Ai = Aj;
  }
  for (int i = 0; i < n; ++i) weights[i] /= w_sum;
}

// ---------------------------------------------------------------------------
// creation.py -- Garth-Joy bucket-split tree, serial DFS with explicit stack
// ---------------------------------------------------------------------------
#pragma pack(push, 1)
struct RefNode {  // NodeDType, constants.py:92-106 (41 bytes, packed)
  int64_t child;
  double Lmax;
  double Rmin;
  int64_t ptr;
  int64_t size;
  uint8_t dim;
};
#pragma pack(pop)
static_assert(sizeof(RefNode) == 41, "NodeDType is 41 bytes");

struct Bucket {  // BucketDType, constants.py:109-122
  double Max, Min, Rmin, Lmax;
  int64_t index, size;
};

// creation.py:27-29
inline RefNode create_node(int64_t ptr, int64_t size, bool dim) {
  return RefNode{-1, -1.0, -1.0, ptr, size, (uint8_t)dim};
}

// creation.py:44-51
inline bool centroid_test(const Bucket& bucket, const double* box, int dim) {
  double centroid = box[2 * dim] + 0.5 * (box[2 * dim + 1] - box[2 * dim]);
  return (centroid >= bucket.Min) && (centroid < bucket.Max);
}

// creation.py:54-112
int64_t stable_partition(int64_t* bb_indices, const double* bb, int64_t begin, int64_t end,
                         const Bucket& bucket, int dim, std::vector<int64_t>& temp) {
  temp.resize(end - begin);
  int64_t count_true = 0;
  int64_t count_false = 0;  // counts from the back: temp[n-1-count_false]
  int64_t n = end - begin;
  for (int64_t k = begin; k < end; ++k) {
    int64_t i = bb_indices[k];
    if (centroid_test(bucket, bb + 4 * i, dim)) temp[count_true++] = i;
    else temp[n - 1 - (count_false++)] = i;
  }
  for (int64_t k = 0; k < count_true; ++k) bb_indices[begin + k] = temp[k];
  int64_t start_second = begin + count_true;
  for (int64_t k = 0; k < count_false; ++k) bb_indices[start_second + k] = temp[n - 1 - k];
  return start_second;
}

// creation.py:115-151.  Returns false when a cell's centroid falls in no bucket
// (the reference would index past the bucket list and raise).
bool sort_bbox_indices(int64_t* bb_indices, const double* bb, std::vector<Bucket>& buckets,
                       const RefNode& node, int dim, std::vector<int64_t>& temp) {
  int64_t current = node.ptr;
  int64_t end = node.ptr + node.size;
  buckets[0].index = node.ptr;
  size_t i = 1;
  while (current != end) {
    if (i - 1 >= buckets.size()) return false;
    Bucket bucket = buckets[i - 1];
    current = stable_partition(bb_indices, bb, current, end, bucket, dim, temp);
    int64_t start = bucket.index;
    buckets[i - 1].size = current - start;
    if (i < buckets.size()) buckets[i].index = buckets[i - 1].index + buckets[i - 1].size;
    i += 1;
  }
  return true;
}

// creation.py:153-171
inline void get_bounds(int64_t index, int64_t size, const double* bb, const int64_t* bb_indices,
                       int dim, double* Rmin_out, double* Lmax_out) {
  double Rmin = FLOAT_MAX, Lmax = FLOAT_MIN;
  for (int64_t i = index; i < index + size; ++i) {
    int64_t d = bb_indices[i];
    double value = bb[4 * d + 2 * dim];
    if (value < Rmin) Rmin = value;
    value = bb[4 * d + 2 * dim + 1];
    if (value > Lmax) Lmax = value;
  }
  *Rmin_out = Rmin;
  *Lmax_out = Lmax;
}

// creation.py:174-213
inline int64_t split_plane(const std::vector<Bucket>& buckets, const RefNode& root,
                           double range_Lmax, double range_Rmin, double bucket_length,
                           double* Lmax_out, double* Rmin_out) {
  double plane_min_cost = FLOAT_MAX;
  int64_t plane = INT_MAX64;
  int64_t bbs_in_left = 0, bbs_in_right = 0;
  for (size_t i = 1; i < buckets.size(); ++i) {
    const Bucket& current_bucket = buckets[i - 1];
    const Bucket& next_bucket = buckets[i];
    bbs_in_left += current_bucket.size;
    bbs_in_right = root.size - bbs_in_left;
    double left_volume = (current_bucket.Lmax - range_Rmin) / bucket_length;
    double right_volume = (range_Lmax - next_bucket.Rmin) / bucket_length;
    double plane_cost = left_volume * (double)bbs_in_left + right_volume * (double)bbs_in_right;
    if (plane_cost < plane_min_cost) {
      plane_min_cost = plane_cost;
      plane = (int64_t)i;
    }
  }
  double Lmax = FLOAT_MIN, Rmin = FLOAT_MAX;
  if (plane == INT_MAX64) return plane;
  for (int64_t i = 0; i < plane; ++i) {
    double v = buckets[i].Lmax;
    if (v > Lmax) Lmax = v;
  }
  for (size_t i = plane; i < buckets.size(); ++i) {
    double v = buckets[i].Rmin;
    if (v < Rmin) Rmin = v;
  }
  *Lmax_out = Lmax;
  *Rmin_out = Rmin;
  return plane;
}

// creation.py:216-230
int64_t pessimistic_n_nodes(int64_t n_elements) {
  int64_t n_nodes = n_elements;
  int64_t nodes = (n_elements + 1) / 2;  // ceil(n/2)
  while (nodes > 1) {
    n_nodes += nodes;
    nodes = (nodes + 1) / 2;
  }
  return n_nodes + 1;
}

// creation.py:233-381.  Returns the node count, or -1 when the bucketing
// degenerates (a centroid outside every bucket, or no finite split cost).
int64_t build(RefNode* nodes, int64_t node_index, int64_t* bb_indices, const double* bb,
              int n_buckets, int cells_per_leaf) {
  std::vector<std::pair<int64_t, int64_t>> stack;  // (root_index, dim flag)
  stack.reserve(64);
  stack.push_back({0, 0});
  std::vector<Bucket> buckets;
  std::vector<int64_t> temp;
  while (!stack.empty()) {
    auto top = stack.back();
    stack.pop_back();
    int64_t root_index = top.first;
    int64_t dim = top.second;
    int64_t dim_flag = dim;
    if (dim < 0) dim += 2;
    RefNode root = nodes[root_index];
    if (root.size <= cells_per_leaf) continue;

    double range_Rmin, range_Lmax;
    get_bounds(root.ptr, root.size, bb, bb_indices, (int)dim, &range_Rmin, &range_Lmax);
    double bucket_length = (range_Lmax - range_Rmin) / (double)n_buckets;

    buckets.clear();
    for (int i = 0; i < n_buckets; ++i)
      buckets.push_back(Bucket{(double)(i + 1) * bucket_length + range_Rmin,
                               (double)i * bucket_length + range_Rmin, -1.0, -1.0, -1, 0});
    if (!sort_bbox_indices(bb_indices, bb, buckets, root, (int)dim, temp)) return -1;

    for (int i = 0; i < n_buckets; ++i) {
      double Rmin, Lmax;
      get_bounds(buckets[i].index, buckets[i].size, bb, bb_indices, (int)dim, &Rmin, &Lmax);
      buckets[i].Rmin = Rmin;
      buckets[i].Lmax = Lmax;
    }

    if ((cells_per_leaf == 1) && (root.size == 2)) {  // creation.py:312-320
      nodes[root_index].Lmax = range_Lmax;
      nodes[root_index].Rmin = range_Rmin;
      RefNode left_child = create_node(root.ptr, 1, !dim);
      RefNode right_child = create_node(root.ptr + 1, 1, !dim);
      nodes[root_index].child = node_index;
      nodes[node_index++] = left_child;
      nodes[node_index++] = right_child;
      continue;
    }

    while (buckets[0].size == 0) {  // creation.py:322-325
      buckets[1].Min = buckets[0].Min;
      buckets.erase(buckets.begin());
    }
    size_t i = 1;
    while (i < buckets.size()) {  // creation.py:327-341
      if (buckets[i].size == 0) {
        buckets[i - 1].Max = buckets[i].Max;
        buckets.erase(buckets.begin() + i);
      } else {
        i += 1;
      }
    }

    bool needs_continue = false;  // creation.py:343-358
    for (const Bucket& bucket : buckets) {
      if (bucket.size == root.size) {
        needs_continue = true;
        if (dim_flag >= 0) {
          dim_flag = (int64_t)(!dim) - 2;
          nodes[root_index].dim = (uint8_t)(!root.dim);
          stack.push_back({root_index, dim_flag});
        } else {
          nodes[root_index].Lmax = -1;
          nodes[root_index].Rmin = -1;
        }
        break;
      }
    }
    if (needs_continue) continue;

    double Lmax, Rmin;
    int64_t plane = split_plane(buckets, root, range_Lmax, range_Rmin, bucket_length, &Lmax, &Rmin);
    if (plane == INT_MAX64) return -1;
    int64_t right_index = buckets[plane].index;
    int64_t right_size = root.ptr + root.size - right_index;
    int64_t left_index = root.ptr;
    int64_t left_size = root.size - right_size;
    nodes[root_index].Lmax = Lmax;
    nodes[root_index].Rmin = Rmin;
    RefNode left_child = create_node(left_index, left_size, !dim);
    RefNode right_child = create_node(right_index, right_size, !dim);
    nodes[root_index].child = node_index;
    int64_t child_ind = node_index;
    nodes[node_index++] = left_child;
    nodes[node_index++] = right_child;
    stack.push_back({child_ind + 1, (int64_t)right_child.dim});
    stack.push_back({child_ind, (int64_t)left_child.dim});
  }
  return node_index;
}

// ---------------------------------------------------------------------------
// query.py
// ---------------------------------------------------------------------------
struct Tree {  // CellTreeData, constants.py:82-89
  const int64_t* elements;
  int64_t n_elements;
  int64_t n_max_vert;
  const double* vertices;
  const RefNode* nodes;
  int64_t n_nodes;
  const int64_t* bb_indices;
  const double* bb_coords;
  double bbox[4];
};

inline Box as_box(const double* a) { return {a[0], a[1], a[2], a[3]}; }
inline Point vtx(const Tree& t, int64_t i) { return {t.vertices[2 * i], t.vertices[2 * i + 1]}; }

// query.py:62-107 (first hit; near side popped first)
int64_t locate_point(Point point, const Tree& tree, double tol, int64_t* visits) {
  std::vector<int64_t> stack;
  stack.reserve(64);
  stack.push_back(0);
  Point poly[MAX_N_VERTEX];
  while (!stack.empty()) {
    int64_t node_index = stack.back();
    stack.pop_back();
    const RefNode& node = tree.nodes[node_index];
    if (visits) ++visits[0];
    if (node.child == -1) {
      for (int64_t i = node.ptr; i < node.ptr + node.size; ++i) {
        int64_t bbox_index = tree.bb_indices[i];
        const int64_t* face = tree.elements + bbox_index * tree.n_max_vert;
        int len = copy_vertices(tree.vertices, face, (int)tree.n_max_vert, poly);
        if (visits) ++visits[1];
        if (point_in_polygon_or_on_edge(point, poly, len, tol)) return bbox_index;
      }
      continue;
    }
    double pd = node.dim ? point.y : point.x;
    bool left = pd <= node.Lmax;
    bool right = pd >= node.Rmin;
    int64_t left_child = node.child, right_child = left_child + 1;
    if (left && right) {
      if ((node.Lmax - pd) < (pd - node.Rmin)) {
        stack.push_back(left_child);
        stack.push_back(right_child);
      } else {
        stack.push_back(right_child);
        stack.push_back(left_child);
      }
    } else if (left) {
      stack.push_back(left_child);
    } else if (right) {
      stack.push_back(right_child);
    }
  }
  return -1;
}

// query.py:120-164
int64_t locate_point_on_edge(Point point, const Tree& tree, double tol) {
  std::vector<int64_t> stack;
  stack.reserve(64);
  stack.push_back(0);
  while (!stack.empty()) {
    int64_t node_index = stack.back();
    stack.pop_back();
    const RefNode& node = tree.nodes[node_index];
    if (node.child == -1) {
      for (int64_t i = node.ptr; i < node.ptr + node.size; ++i) {
        int64_t bbox_index = tree.bb_indices[i];
        const int64_t* edge = tree.elements + bbox_index * tree.n_max_vert;
        if (point_on_edge(point, vtx(tree, edge[0]), vtx(tree, edge[1]), tol)) return bbox_index;
      }
      continue;
    }
    double pd = node.dim ? point.y : point.x;
    bool left = pd <= node.Lmax;
    bool right = pd >= node.Rmin;
    int64_t left_child = node.child, right_child = left_child + 1;
    if (left && right) {
      if ((node.Lmax - pd) < (pd - node.Rmin)) {
        stack.push_back(left_child);
        stack.push_back(right_child);
      } else {
        stack.push_back(right_child);
        stack.push_back(left_child);
      }
    } else if (left) {
      stack.push_back(left_child);
    } else if (right) {
      stack.push_back(right_child);
    }
  }
  return -1;
}

// query.py:177-244 (every strict bbox overlap; right child popped first)
void locate_box(Box box, const Tree& tree, int64_t index, std::vector<int64_t>& ii,
                std::vector<int64_t>& jj) {
  Box tree_bbox = as_box(tree.bbox);
  if (!boxes_intersect(box, tree_bbox)) return;
  std::vector<int64_t> stack;
  stack.reserve(64);
  stack.push_back(0);
  while (!stack.empty()) {
    int64_t node_index = stack.back();
    stack.pop_back();
    const RefNode& node = tree.nodes[node_index];
    if (node.child == -1) {
      for (int64_t i = node.ptr; i < node.ptr + node.size; ++i) {
        int64_t bbox_index = tree.bb_indices[i];
        Box leaf_box = as_box(tree.bb_coords + 4 * bbox_index);
        if (boxes_intersect(box, leaf_box)) {
          ii.push_back(index);
          jj.push_back(bbox_index);
        }
      }
    } else {
      double bmin = node.dim ? box.ymin : box.xmin;
      double bmax = node.dim ? box.ymax : box.xmax;
      bool left = bmin <= node.Lmax;
      bool right = bmax >= node.Rmin;
      int64_t left_child = node.child, right_child = left_child + 1;
      if (left && right) {
        stack.push_back(left_child);
        stack.push_back(right_child);
      } else if (left) {
        stack.push_back(left_child);
      } else if (right) {
        stack.push_back(right_child);
      }
    }
  }
}

// query.py:292-306
inline bool compute_edge_edge_intersect(const Tree& tree, int64_t bbox_index, Point a, Point b,
                                        Point* c, Point* d) {
  const int64_t* e = tree.elements + bbox_index * tree.n_max_vert;
  Point p = vtx(tree, e[0]), q = vtx(tree, e[1]);
  Point x;
  bool intersects = lines_intersect(a, b, p, q, &x);
  *c = x;
  *d = x;
  return intersects;
}

// query.py:309-328
inline bool compute_edge_face_intersect(const Tree& tree, int64_t bbox_index, Point a, Point b,
                                        Point* c, Point* d) {
  Box box = as_box(tree.bb_coords + 4 * bbox_index);
  bool intersects = cohen_sutherland_line_box_clip(a, b, box, c, d);
  if (intersects) {
    Point poly[MAX_N_VERTEX];
    int len = copy_vertices(tree.vertices, tree.elements + bbox_index * tree.n_max_vert,
                            (int)tree.n_max_vert, poly);
    double tol = pymax(MIN_TOLERANCE,
                       TOLERANCE_FACTOR * pymax(box.xmax - box.xmin, box.ymax - box.ymin));
    intersects = cyrus_beck_line_polygon_clip(a, b, poly, len, tol, c, d);
  }
  return intersects;
}

// query.py:357-455 (slab pruning per node; right child popped first)
void locate_edge(Point a, Point b, const Tree& tree, int64_t index, bool edge_face,
                 std::vector<int64_t>& ii, std::vector<int64_t>& jj, std::vector<double>& xy) {
  Box tree_bbox = as_box(tree.bbox);
  Point c0, d0;
  if (!cohen_sutherland_line_box_clip(a, b, tree_bbox, &c0, &d0)) return;
  Vector V = to_vector(a, b);
  std::vector<int64_t> stack;
  stack.reserve(64);
  stack.push_back(0);
  while (!stack.empty()) {
    int64_t node_index = stack.back();
    stack.pop_back();
    const RefNode& node = tree.nodes[node_index];
    if (node.child == -1) {
      for (int64_t i = node.ptr; i < node.ptr + node.size; ++i) {
        int64_t bbox_index = tree.bb_indices[i];
        Point c, d;
        bool intersects = edge_face ? compute_edge_face_intersect(tree, bbox_index, a, b, &c, &d)
                                    : compute_edge_edge_intersect(tree, bbox_index, a, b, &c, &d);
        if (intersects) {
          ii.push_back(index);
          jj.push_back(bbox_index);
          xy.push_back(c.x);
          xy.push_back(c.y);
          xy.push_back(d.x);
          xy.push_back(d.y);
        }
      }
      continue;
    }
    int nd = node.dim ? 1 : 0;
    double dx = nd ? V.y : V.x;
    double an = nd ? a.y : a.x;
    double bn = nd ? b.y : b.x;
    double dx_left, dx_right;
    if (dx > 0.0) {
      dx_left = node.Lmax - an;
      dx_right = node.Rmin - bn;
    } else {
      dx_left = node.Lmax - bn;
      dx_right = node.Rmin - an;
    }
    bool left = dx_left >= 0.0;
    bool right = dx_right <= 0.0;
    if (dx > 0.0) {
      if (left) left = (dx_left / dx) >= 0.0;
      if (right) right = (dx_right / dx) <= 1.0;
    } else if (dx < 0.0) {
      if (left) left = (1.0 - (dx_left / dx)) >= 0.0;
      if (right) right = (1.0 - (dx_right / dx)) <= 1.0;
    }
    int64_t left_child = node.child, right_child = left_child + 1;
    if (left && right) {
      stack.push_back(left_child);
      stack.push_back(right_child);
    } else if (left) {
      stack.push_back(left_child);
    } else if (right) {
      stack.push_back(right_child);
    }
  }
}

int n_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // namespace

// ===========================================================================
// C ABI of the oracle (loaded by oracle/refcpu.py with ctypes)
// ===========================================================================
// ------------------------------------------------------------ overlap-area adjoint
// No reference counterpart (the reference computes no derivatives): the derivative of
// polygon_polygon_clip_area (sutherland_hodgman.py:84-148) of two convex counter-clockwise
// polygons a (subject) and b (clipper) with respect to their vertices, restated from the
// mathematics for the roofline and the checks of the GPU regrid backward.  The overlap P is
// bounded by the pieces of the input edges that lie in the other polygon; moving the ends of
// edge v0 -> v1 (d = v1 - v0) moves the point at parameter u by (1 - u) and u times the
// displacement, and the area changes by the outward normal velocity integrated over the piece
// [alpha, beta] of the edge inside the other polygon:
//   dA/dv0 = m ((beta - alpha) - (beta^2 - alpha^2) / 2),   dA/dv1 = m (beta^2 - alpha^2) / 2,
// with m = (d.y, -d.x).  A piece is an edge clipped by the other polygon's half-planes
// (inside: u x (p - r) > 0, the clip's test): subject edges by open half-planes, clipper edges
// by closed ones, where a clipper edge on a subject edge line counts when both run the same
// way (the clip keeps such a boundary piece as the clipper's).  Bounds are fractions
// (numerator / positive denominator) compared by cross-multiplication, each end divided once.
// Operations are counted as in the forward (flops<>), the cross-multiplications included.
struct PieceRange {
  double n0 = 0.0, d0 = 1.0, n1 = 1.0, d1 = 1.0;
  bool empty = false;
};

template <bool COUNT>
void clip_piece(Point v0, double dx, double dy, Point r, Point s, bool closed, PieceRange& e) {
  const double ux = s.x - r.x, uy = s.y - r.y;
  const double c0 = ux * (v0.y - r.y) - uy * (v0.x - r.x);
  const double c1 = ux * dy - uy * dx;
  flops<COUNT>(2 + 5 + 3);
  if (c1 == 0.0) {
    bool in = c0 > 0.0;
    if (closed && c0 == 0.0) {
      in = (ux * dx + uy * dy) > 0.0;
      flops<COUNT>(3);
    }
    e.empty = e.empty || !in;
  } else if (c1 > 0.0) {  // t >= -c0 / c1
    const double num = -c0;
    flops<COUNT>(2);
    if (num * e.d0 > e.n0 * c1) {
      e.n0 = num;
      e.d0 = c1;
    }
  } else {  // t <= c0 / -c1
    const double den = -c1;
    flops<COUNT>(2);
    if (c0 * e.d1 < e.n1 * den) {
      e.n1 = c0;
      e.d1 = den;
    }
  }
}

// The two end contributions of edge v0 -> v1's piece inside polygon c (lc vertices), scaled
// by gA; false when the piece is empty.  clipper_edges: v0 -> v1 is an edge of the clipper
// and c the subject (closed half-planes).
template <bool COUNT>
bool piece_gradient(Point v0, Point v1, const Point* c, int lc, bool clipper_edges, double gA, Vector* g0,
                    Vector* g1) {
  const double dx = v1.x - v0.x, dy = v1.y - v0.y;
  flops<COUNT>(2);
  PieceRange e;
  for (int j = 0; j < lc; ++j) {
    // subject edges meet the clipper's edges c[j - 1] -> c[j]; clipper edges the subject's
    // c[j] -> c[j + 1] (the GPU kernel's enumeration, so that equal bounds resolve alike)
    const Point r = clipper_edges ? c[j] : c[(j + lc - 1) % lc];
    const Point s = clipper_edges ? c[(j + 1) % lc] : c[j];
    clip_piece<COUNT>(v0, dx, dy, r, s, clipper_edges, e);
  }
  const double alpha = e.n0 / e.d0, beta = e.n1 / e.d1;
  flops<COUNT>(2);
  if (e.empty || !(beta > alpha)) return false;
  const double w1 = 0.5 * (beta * beta - alpha * alpha);
  const double w0 = (beta - alpha) - w1;
  *g0 = {gA * w0 * dy, -gA * w0 * dx};
  *g1 = {gA * w1 * dy, -gA * w1 * dx};
  flops<COUNT>(4 + 2 + 4 + 4);
  return true;
}

// Pair gradient: ga[i] = gA dA/da[i], gb[j] = gA dA/db[j] (la, lb vertices), summed edge by
// edge (4 adds per contributing edge).
template <bool COUNT>
void pair_area_gradient(const Point* a, int la, const Point* b, int lb, double gA, Vector* ga, Vector* gb) {
  for (int i = 0; i < la; ++i) ga[i] = {0.0, 0.0};
  for (int j = 0; j < lb; ++j) gb[j] = {0.0, 0.0};
  Vector g0, g1;
  for (int i = 0; i < la; ++i) {
    const int i1 = (i + 1) % la;
    if (piece_gradient<COUNT>(a[i], a[i1], b, lb, false, gA, &g0, &g1)) {
      ga[i] = {ga[i].x + g0.x, ga[i].y + g0.y};
      ga[i1] = {ga[i1].x + g1.x, ga[i1].y + g1.y};
      flops<COUNT>(4);
    }
  }
  for (int j = 0; j < lb; ++j) {
    const int j0 = (j + lb - 1) % lb;
    if (piece_gradient<COUNT>(b[j0], b[j], a, la, true, gA, &g0, &g1)) {
      gb[j0] = {gb[j0].x + g0.x, gb[j0].y + g0.y};
      gb[j] = {gb[j].x + g1.x, gb[j].y + g1.y};
      flops<COUNT>(4);
    }
  }
}

extern "C" {

struct ref_tree_t {
  const int64_t* elements;
  int64_t n_elements;
  int64_t n_max_vert;
  const double* vertices;
  const void* nodes;  // RefNode[n_nodes], NodeDType layout
  int64_t n_nodes;
  const int64_t* bb_indices;
  const double* bb_coords;
  double bbox[4];
};

struct ref_pairs_t {  // variable-length outputs (malloc'ed; free with ref_pairs_free)
  int64_t n;
  int64_t* ii;
  int64_t* jj;
  double* xy;  // (n, 2, 2) or null
};

static Tree to_tree(const ref_tree_t* t) {
  Tree tree;
  tree.elements = t->elements;
  tree.n_elements = t->n_elements;
  tree.n_max_vert = t->n_max_vert;
  tree.vertices = t->vertices;
  tree.nodes = (const RefNode*)t->nodes;
  tree.n_nodes = t->n_nodes;
  tree.bb_indices = t->bb_indices;
  tree.bb_coords = t->bb_coords;
  for (int i = 0; i < 4; ++i) tree.bbox[i] = t->bbox[i];
  return tree;
}

int ref_num_threads(void) { return n_threads(); }

int64_t ref_pessimistic_n_nodes(int64_t n) { return pessimistic_n_nodes(n); }

// creation.py:384-413.  nodes_out must hold pessimistic_n_nodes(n) records.
int64_t ref_initialize(const double* bb_coords, int64_t n, int n_buckets, int cells_per_leaf,
                       void* nodes_out, int64_t* bb_indices_out) {
  RefNode* nodes = (RefNode*)nodes_out;
  for (int64_t i = 0; i < n; ++i) bb_indices_out[i] = i;
  nodes[0] = create_node(0, n, false);
  return build(nodes, 1, bb_indices_out, bb_coords, n_buckets, cells_per_leaf);
}

void ref_counter_clockwise(const double* vertices, int64_t* faces, int64_t n_faces, int n_max) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_faces; ++i) counter_clockwise_row(vertices, faces + i * n_max, n_max);
}

void ref_build_face_bboxes(const int64_t* faces, int64_t n, int n_max, const double* vertices,
                           double* out) {
  for (int64_t i = 0; i < n; ++i) {
    Box b = bounding_box(faces + i * n_max, n_max, vertices);
    out[4 * i] = b.xmin;
    out[4 * i + 1] = b.xmax;
    out[4 * i + 2] = b.ymin;
    out[4 * i + 3] = b.ymax;
  }
}

void ref_build_edge_bboxes(const int64_t* edges, int64_t n, const double* vertices, double tol,
                           double* out) {
  for (int64_t i = 0; i < n; ++i) {
    Box b = edge_bounding_box(edges + 2 * i, vertices, tol);
    out[4 * i] = b.xmin;
    out[4 * i + 1] = b.xmax;
    out[4 * i + 2] = b.ymin;
    out[4 * i + 3] = b.ymax;
  }
}

// query.py:110-117
void ref_locate_points(const ref_tree_t* t, const double* points, int64_t n, double tol,
                       int64_t* out) {
  Tree tree = to_tree(t);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i)
    out[i] = locate_point({points[2 * i], points[2 * i + 1]}, tree, tol, nullptr);
}

// Instrumented single-thread variant: visits[0] += node visits, visits[1] += cell tests.
void ref_locate_points_counted(const ref_tree_t* t, const double* points, int64_t n, double tol,
                               int64_t* out, int64_t* visits) {
  Tree tree = to_tree(t);
  for (int64_t i = 0; i < n; ++i)
    out[i] = locate_point({points[2 * i], points[2 * i + 1]}, tree, tol, visits);
}

// query.py:167-174
void ref_locate_points_on_edge(const ref_tree_t* t, const double* points, int64_t n, double tol,
                               int64_t* out) {
  Tree tree = to_tree(t);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i)
    out[i] = locate_point_on_edge({points[2 * i], points[2 * i + 1]}, tree, tol);
}

static void pairs_from_chunks(std::vector<std::vector<int64_t>>& cii,
                              std::vector<std::vector<int64_t>>& cjj,
                              std::vector<std::vector<double>>* cxy, ref_pairs_t* out) {
  int64_t total = 0;
  for (auto& v : cii) total += (int64_t)v.size();
  out->n = total;
  out->ii = (int64_t*)std::malloc(sizeof(int64_t) * (total ? total : 1));
  out->jj = (int64_t*)std::malloc(sizeof(int64_t) * (total ? total : 1));
  out->xy = cxy ? (double*)std::malloc(sizeof(double) * 4 * (total ? total : 1)) : nullptr;
  int64_t start = 0;
  for (size_t c = 0; c < cii.size(); ++c) {  // query.py:46-59, chunk order = query order
    std::copy(cii[c].begin(), cii[c].end(), out->ii + start);
    std::copy(cjj[c].begin(), cjj[c].end(), out->jj + start);
    if (cxy) std::copy((*cxy)[c].begin(), (*cxy)[c].end(), out->xy + 4 * start);
    start += (int64_t)cii[c].size();
  }
}

// query.py:278-289 (one contiguous chunk per thread, concatenated in order)
void ref_locate_boxes(const ref_tree_t* t, const double* boxes, int64_t n, ref_pairs_t* out) {
  Tree tree = to_tree(t);
  int nc = n_threads();
  std::vector<std::vector<int64_t>> cii(nc), cjj(nc);
#pragma omp parallel for schedule(static, 1)
  for (int c = 0; c < nc; ++c) {
    int64_t lo = n * c / nc, hi = n * (c + 1) / nc;
    for (int64_t i = lo; i < hi; ++i) locate_box(as_box(boxes + 4 * i), tree, i, cii[c], cjj[c]);
  }
  pairs_from_chunks(cii, cjj, nullptr, out);
}

// query.py:500-544 (edge_face != 0: locate_edge_faces, else locate_edge_edges)
void ref_locate_edges(const ref_tree_t* t, const double* edges, int64_t n, int edge_face,
                      ref_pairs_t* out) {
  Tree tree = to_tree(t);
  int nc = n_threads();
  std::vector<std::vector<int64_t>> cii(nc), cjj(nc);
  std::vector<std::vector<double>> cxy(nc);
#pragma omp parallel for schedule(static, 1)
  for (int c = 0; c < nc; ++c) {
    int64_t lo = n * c / nc, hi = n * (c + 1) / nc;
    for (int64_t i = lo; i < hi; ++i) {
      Point a = {edges[4 * i], edges[4 * i + 1]};
      Point b = {edges[4 * i + 2], edges[4 * i + 3]};
      locate_edge(a, b, tree, i, edge_face != 0, cii[c], cjj[c], cxy[c]);
    }
  }
  pairs_from_chunks(cii, cjj, &cxy, out);
}

void ref_pairs_free(ref_pairs_t* p) {
  std::free(p->ii);
  std::free(p->jj);
  std::free(p->xy);
  p->ii = p->jj = nullptr;
  p->xy = nullptr;
  p->n = 0;
}

// separating_axis.py:58-75
void ref_polygons_intersect(const double* va, const double* vb, const int64_t* fa, int na_max,
                            const int64_t* fb, int nb_max, const int64_t* ia, const int64_t* ib,
                            int64_t n, uint8_t* out) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) {
    Point a[MAX_N_VERTEX], b[MAX_N_VERTEX];
    int la = copy_vertices(va, fa + ia[i] * na_max, na_max, a);
    int lb = copy_vertices(vb, fb + ib[i] * nb_max, nb_max, b);
    out[i] = separating_axes(a, la, b, lb) && separating_axes(b, lb, a, la);
  }
}

// sutherland_hodgman.py:151-168 (subject = face a, clipper = face b)
void ref_area_of_intersection(const double* va, const double* vb, const int64_t* fa, int na_max,
                              const int64_t* fb, int nb_max, const int64_t* ia,
                              const int64_t* ib, int64_t n, double* out) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) {
    Point a[MAX_N_VERTEX], b[MAX_N_VERTEX];
    int la = copy_vertices(va, fa + ia[i] * na_max, na_max, a);
    int lb = copy_vertices(vb, fb + ib[i] * nb_max, nb_max, b);
    out[i] = polygon_polygon_clip_area(a, la, b, lb);
  }
}

// Operation counts of intersect_faces' pair stage (celltree.py:248-269) for every candidate
// pair: sat[i] = FP64 ops of polygons_intersect (separating_axis.py:58-75, both directions,
// short-circuit as `and`), passed[i] = its result, clip[i] = ops of area_of_intersection
// (sutherland_hodgman.py:151-168, clip + polygon_area) for pairs that passed, else 0, and
// area[i] = that area (0 where not computed).  Counts are of the executed path (see
// flops<>).
void ref_face_pair_flops(const double* va, const double* vb, const int64_t* fa, int na_max,
                         const int64_t* fb, int nb_max, const int64_t* ia, const int64_t* ib,
                         int64_t n, int64_t* sat, uint8_t* passed, int64_t* clip, double* area) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) {
    Point a[MAX_N_VERTEX], b[MAX_N_VERTEX];
    int la = copy_vertices(va, fa + ia[i] * na_max, na_max, a);
    int lb = copy_vertices(vb, fb + ib[i] * nb_max, nb_max, b);
    t_flops = 0;
    const bool ok = separating_axes<true>(a, la, b, lb) && separating_axes<true>(b, lb, a, la);
    sat[i] = t_flops;
    passed[i] = ok;
    t_flops = 0;
    area[i] = ok ? polygon_polygon_clip_area<true>(a, la, b, lb) : 0.0;
    clip[i] = t_flops;
  }
}

// Per pair i (subject a = faces_a[ia[i]], clipper b = faces_b[ib[i]], both convex and
// counter-clockwise): the overlap-area adjoint scaled by gA[i] (see pair_area_gradient) in
// ga[i * na_max + slot] / gb[i * nb_max + slot] (x, y per vertex slot; unused slots 0), and
// its FP64 operation count in ops[i] (0 where gA[i] == 0: no gradient is formed).
void ref_pair_area_gradient(const double* va, const double* vb, const int64_t* fa, int na_max,
                            const int64_t* fb, int nb_max, const int64_t* ia, const int64_t* ib, int64_t n,
                            const double* gA, double* ga, double* gb, int64_t* ops) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) {
    Point a[MAX_N_VERTEX], b[MAX_N_VERTEX];
    Vector da[MAX_N_VERTEX], db[MAX_N_VERTEX];
    const int la = copy_vertices(va, fa + ia[i] * na_max, na_max, a);
    const int lb = copy_vertices(vb, fb + ib[i] * nb_max, nb_max, b);
    for (int k = 0; k < na_max; ++k) da[k] = {0.0, 0.0};
    for (int k = 0; k < nb_max; ++k) db[k] = {0.0, 0.0};
    t_flops = 0;
    if (gA[i] != 0.0) pair_area_gradient<true>(a, la, b, lb, gA[i], da, db);
    ops[i] = t_flops;
    for (int k = 0; k < na_max; ++k) {
      ga[2 * (i * na_max + k)] = da[k].x;
      ga[2 * (i * na_max + k) + 1] = da[k].y;
    }
    for (int k = 0; k < nb_max; ++k) {
      gb[2 * (i * nb_max + k)] = db[k].x;
      gb[2 * (i * nb_max + k) + 1] = db[k].y;
    }
  }
}

// sutherland_hodgman.py:171-187 (subject = box, clipper = face)
void ref_box_area_of_intersection(const double* bbox_coords, const double* vertices,
                                  const int64_t* faces, int n_max, const int64_t* ibox,
                                  const int64_t* iface, int64_t n, double* out) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) {
    Point a[4], b[MAX_N_VERTEX];
    copy_box_vertices(as_box(bbox_coords + 4 * ibox[i]), a);
    int lb = copy_vertices(vertices, faces + iface[i] * n_max, n_max, b);
    out[i] = polygon_polygon_clip_area(a, 4, b, lb);
  }
}

// barycentric_triangle.py:46-64
void ref_barycentric_triangle_weights(const double* points, const int64_t* face_indices,
                                      int64_t n, const int64_t* faces, int n_max,
                                      const double* vertices, double tol, double* out) {
  (void)tol;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double* w = out + 3 * i;
    w[0] = w[1] = w[2] = 0.0;
    int64_t f = face_indices[i];
    if (f == -1) continue;
    const int64_t* face = faces + f * n_max;
    Point a = {vertices[2 * face[0]], vertices[2 * face[0] + 1]};
    Point b = {vertices[2 * face[1]], vertices[2 * face[1] + 1]};
    Point c = {vertices[2 * face[2]], vertices[2 * face[2] + 1]};
    bary_triangle(a, b, c, {points[2 * i], points[2 * i + 1]}, w);
  }
}

// barycentric_wachspress.py:88-107
void ref_barycentric_wachspress_weights(const double* points, const int64_t* face_indices,
                                        int64_t n, const int64_t* faces, int n_max,
                                        const double* vertices, double tol, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double* w = out + n_max * i;
    for (int k = 0; k < n_max; ++k) w[k] = 0.0;
    int64_t f = face_indices[i];
    if (f == -1) continue;
    Point poly[MAX_N_VERTEX];
    int len = copy_vertices(vertices, faces + f * n_max, n_max, poly);
    wachspress_weights(poly, len, {points[2 * i], points[2 * i + 1]}, w, n_max, tol);
  }
}

// query.py:568-622
void ref_collect_node_bounds(const ref_tree_t* t, double* node_bounds) {
  Tree tree = to_tree(t);
  for (int k = 0; k < 4; ++k) node_bounds[k] = tree.bbox[k];
  if (tree.n_nodes < 3) return;  // a single leaf: only the root's bounds exist
  struct Item { int64_t node, parent, side; };
  std::vector<Item> stack = {{2, 0, 0}, {1, 0, 1}};  // right (side 0) then left (side 1)
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    const RefNode& parent = tree.nodes[it.parent];
    const double* bbox = node_bounds + 4 * it.parent;
    int dim = parent.dim ? 1 : 0;
    double* nb = node_bounds + 4 * it.node;
    for (int k = 0; k < 4; ++k) nb[k] = bbox[k];
    double bound = it.side ? parent.Lmax : parent.Rmin;
    nb[dim * 2 + it.side] = bound;
    const RefNode& node = tree.nodes[it.node];
    if (node.child == -1) continue;
    stack.push_back({node.child + 1, it.node, 0});
    stack.push_back({node.child, it.node, 1});
  }
}

// query.py:625-664
void ref_validate_node_bounds(const ref_tree_t* t, const double* node_bounds, uint8_t* validity) {
  Tree tree = to_tree(t);
  for (int64_t i = 0; i < tree.n_nodes; ++i) validity[i] = 0;
  std::vector<int64_t> stack = {0};
  while (!stack.empty()) {
    int64_t node_index = stack.back();
    stack.pop_back();
    Box bbox = as_box(node_bounds + 4 * node_index);
    const RefNode& node = tree.nodes[node_index];
    if (node.child == -1) {
      bool valid = true;
      for (int64_t i = node.ptr; i < node.ptr + node.size; ++i) {
        Box leaf_box = as_box(tree.bb_coords + 4 * tree.bb_indices[i]);
        valid = valid && box_contained(leaf_box, bbox);
      }
      validity[node_index] = valid;
      continue;
    }
    int64_t l = node.child, r = l + 1;
    validity[node_index] = box_contained(as_box(node_bounds + 4 * l), bbox) &&
                           box_contained(as_box(node_bounds + 4 * r), bbox);
    stack.push_back(r);
    stack.push_back(l);
  }
}

// ---- single primitives, for the unit tests of the algorithms ----

double ref_polygon_area(const double* poly, int n) {
  return polygon_area((const Point*)poly, n);
}

double ref_polygon_polygon_clip_area(const double* subject, int ns, const double* clipper, int nc) {
  return polygon_polygon_clip_area((const Point*)subject, ns, (const Point*)clipper, nc);
}

int ref_sh_intersection(const double* a, const double* V, const double* r, const double* N,
                        double* out) {
  Point p;
  bool ok = sh_intersection({a[0], a[1]}, {V[0], V[1]}, {r[0], r[1]}, {N[0], N[1]}, &p);
  out[0] = p.x;
  out[1] = p.y;
  return ok;
}

int ref_separating_axes(const double* a, int na, const double* b, int nb) {
  return separating_axes((const Point*)a, na, (const Point*)b, nb);
}

int ref_point_in_polygon_or_on_edge(const double* p, const double* poly, int n, double tol) {
  return point_in_polygon_or_on_edge({p[0], p[1]}, (const Point*)poly, n, tol);
}

int ref_point_on_edge(const double* p, const double* edge, double tol) {
  return point_on_edge({p[0], p[1]}, {edge[0], edge[1]}, {edge[2], edge[3]}, tol);
}

int ref_cohen_sutherland(const double* a, const double* b, const double* box, double* cd) {
  Point c, d;
  bool ok = cohen_sutherland_line_box_clip({a[0], a[1]}, {b[0], b[1]}, as_box(box), &c, &d);
  cd[0] = c.x; cd[1] = c.y; cd[2] = d.x; cd[3] = d.y;
  return ok;
}

int ref_cyrus_beck(const double* a, const double* b, const double* poly, int n, double tol,
                   double* cd) {
  Point c, d;
  bool ok = cyrus_beck_line_polygon_clip({a[0], a[1]}, {b[0], b[1]}, (const Point*)poly, n, tol,
                                         &c, &d);
  cd[0] = c.x; cd[1] = c.y; cd[2] = d.x; cd[3] = d.y;
  return ok;
}

int ref_lines_intersect(const double* a, const double* b, const double* p, const double* q,
                        double* xy) {
  Point x;
  bool ok = lines_intersect({a[0], a[1]}, {b[0], b[1]}, {p[0], p[1]}, {q[0], q[1]}, &x);
  xy[0] = x.x;
  xy[1] = x.y;
  return ok;
}

void ref_bary_triangle(const double* tri, const double* p, double* w) {
  bary_triangle({tri[0], tri[1]}, {tri[2], tri[3]}, {tri[4], tri[5]}, {p[0], p[1]}, w);
}

void ref_wachspress(const double* poly, int n, const double* p, double* w, int n_w, double tol) {
  for (int k = 0; k < n_w; ++k) w[k] = 0.0;
  wachspress_weights((const Point*)poly, n, {p[0], p[1]}, w, n_w, tol);
}

void ref_interp_edge_case(const double* a, const double* U, const double* p, double* w, int n_w,
                          int i, int j) {
  interp_edge_case({a[0], a[1]}, {U[0], U[1]}, {p[0], p[1]}, w, n_w, i, j);
}

}  // extern "C"
