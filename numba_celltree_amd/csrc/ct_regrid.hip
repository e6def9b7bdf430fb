// ct_regrid.hip -- conservative area-weighted regridding and its adjoint (no reference
// counterpart; the forward is the pattern of examples/spatial_indexing_2d_grids.py:268-323).
//
//   y[t, f] = sum_{k: t_k = t} A_k * F[s_k, f] / area_t
//
// A_k is the Sutherland-Hodgman overlap area of pair k (subject = target polygon t_k,
// clipper = source polygon s_k; sutherland_hodgman.py:84-148) and area_t the fan area of
// the target polygon (geometry_utils.py:82-95).  The pair topology (which pairs overlap) is
// held fixed; the adjoint is the analytic derivative of the overlap areas and target areas
// with respect to F, the source vertices and the target vertices.
//
// Pair lists are grouped by target (ascending t, as intersect_faces returns them): the
// forward is a deterministic segmented sum per target.  The transpose uses a by-source
// permutation of the pairs (also deterministic).  Vertex gradients are reduced without
// atomics, in a fixed order (per-pair corner rows, per-face corner sums, per-vertex CSR
// gathers), so they are identical from run to run.
//
// Input polygons live in registers: loops run to the compile-time capacity with a length
// guard.  Runtime-indexed data goes to per-lane rows in LDS (the clip's output polygon,
// ct_clip.h, whose slots past the tenth are private memory that only polygons of more than 10
// vertices reach; the adjoint's target-corner sums) or through select chains.
#include <vector>

#include "ct_clip.h"
#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {
namespace {

// arr[n] += v for a runtime n, as one predicated add per slot.
template <int CAP>
__device__ __forceinline__ void add_at(P2 (&arr)[CAP], int n, P2 v) {
#pragma unroll
  for (int k = 0; k < CAP; ++k) arr[k] = sel(k == n, P2{arr[k].x + v.x, arr[k].y + v.y}, arr[k]);
}

// geometry_utils.py:82-95 polygon_area (fan from vertex 0, abs per triangle)
template <int CAP>
__device__ __forceinline__ double fan_area(const P2 (&p)[CAP], int len) {
  return polygon_area<CAP>(p, len);
}

// Adjoint of fan_area: g[i] += d(area)/d(p[i]) * garea.
template <int CAP>
__device__ __forceinline__ void fan_area_adjoint(const P2 (&p)[CAP], int len, double garea, P2 (&g)[CAP]) {
#pragma unroll
  for (int i = 2; i < CAP; ++i) {
    if (i < len) {
      // term = |U x V|, U = P_{i-1} - P0 (P1 - P0 for i = 2), V = P0 - P_i
      const int ip = i - 1;
      double ux, uy;
      if (i == 2) {
        ux = p[1].x - p[0].x;
        uy = p[1].y - p[0].y;
      } else {
        ux = p[0].x - p[ip].x;
        uy = p[0].y - p[ip].y;
      }
      const double vx = p[0].x - p[i].x, vy = p[0].y - p[i].y;
      const double c = ux * vy - uy * vx;
      const double gc = 0.5 * garea * ((c > 0) ? 1.0 : (c < 0 ? -1.0 : 0.0));
      const double gux = gc * vy, guy = -gc * vx;
      const double gvx = -gc * uy, gvy = gc * ux;
      if (i == 2) {  // U = P1 - P0
        g[1].x += gux;
        g[1].y += guy;
        g[0].x -= gux;
        g[0].y -= guy;
      } else {  // U = P0 - P_{i-1}
        g[0].x += gux;
        g[0].y += guy;
        g[ip].x -= gux;
        g[ip].y -= guy;
      }
      g[0].x += gvx;  // V = P0 - P_i
      g[0].y += gvy;
      g[i].x -= gvx;
      g[i].y -= gvy;
    }
  }
}


template <int MAXV>
__device__ __forceinline__ int load_face_idx(const int32_t* __restrict__ f, int nmax, int64_t k,
                                             const double* __restrict__ v, P2 (&poly)[MAXV], int (&vid)[MAXV]) {
  const int32_t* row = f + k * nmax;
  int raw[MAXV];
#pragma unroll
  for (int i = 0; i < MAXV; ++i) raw[i] = (i < nmax) ? row[i] : kFill;
  int len = nmax < MAXV ? nmax : MAXV;
#pragma unroll
  for (int i = MAXV - 1; i >= 3; --i)
    if (i < nmax && raw[i] == kFill) len = i;
#pragma unroll
  for (int i = 0; i < MAXV; ++i) {
    vid[i] = (i < len) ? raw[i] : 0;
    poly[i] = (i < len) ? ld_vertex(v, vid[i]) : P2{0.0, 0.0};
  }
  return len;
}

template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) k_pair_areas(const double* __restrict__ tv, const int32_t* __restrict__ tf,
                                                       int ntm, const double* __restrict__ sv,
                                                       const int32_t* __restrict__ sf, int nsm,
                                                       const int64_t* __restrict__ tt,
                                                       const int64_t* __restrict__ ss, int64_t n,
                                                       double* __restrict__ area) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  P2 a[MA], b[MB];
  int ai[MA], bi[MB];
  const int la = load_face_idx<MA>(tf, ntm, tt[k], tv, a, ai);
  const int lb = load_face_idx<MB>(sf, nsm, ss[k], sv, b, bi);
  CT_CLIP_ROWS((clip_row_cap<MA, MB>()), row);
  area[k] = clip_area<MA, MB>(a, la, b, lb, row);
}

// Per pair: gA = sum_f gy[t,f] F[s,f] / area_t, times the derivative of the overlap area.
//
// With both polygons convex and counter-clockwise, the clip computes their intersection P,
// whose area is a smooth function of the input vertices away from degenerate contacts; its
// derivative lives on the boundary of P.  That boundary is made of the pieces of the input
// edges that lie in the other polygon: every input edge v0 -> v1 (d = v1 - v0) contributes
// the parameter range [alpha, beta] of it that lies in the other polygon, and moving v0 (v1)
// moves the point at parameter u by (1 - u) (u) times the displacement.  The area changes by
// the outward normal velocity integrated over that range, so with m = (d.y, -d.x) (outward
// normal times |d|):
//   dA/dv0 += m * ((beta - alpha) - (beta^2 - alpha^2) / 2),   dA/dv1 += m * (beta^2 - alpha^2) / 2.
// The ranges come from clipping each edge of one polygon against the half-planes of the
// other (Cyrus-Beck), with every index static: no clipped polygon, no runtime-indexed
// registers.  Where an edge of the target (the clip's subject) lies on an edge line of the
// source (the clipper), the Sutherland-Hodgman pass counts that boundary piece as the
// source's edge (its inside test is strict, sutherland_hodgman.py:57-60): target edges are
// clipped against open half-planes, source edges against closed ones, and a source edge on
// a target edge line counts only when both run the same way (overlapping interiors; when
// they run opposite ways the cells only touch and the clip has no polygon).
// Checked against finite differences, torch.autograd.gradcheck and an autograd restatement
// of the clip.

// Parameter range [n0 / d0, n1 / d1] of a segment (d0, d1 > 0), kept as fractions so that
// clipping compares by cross-multiplication and divides once per end at the close.
struct EdgeRange {
  double n0 = 0.0, d0 = 1.0, n1 = 1.0, d1 = 1.0;
  bool empty = false;
};

// Shrinks the range to the part of v0 + t d inside the line through r with direction u
// (inside: u x (p - r) > 0, sh_in's test).  CLOSED: on the line counts as inside when the
// segment runs the same way as u.  The bound of a crossing line is -c0 / c1.
template <bool CLOSED>
__device__ __forceinline__ void clip_range(P2 v0, double dx, double dy, P2 r, double ux, double uy,
                                           EdgeRange& e) {
  const double c0 = ux * (v0.y - r.y) - uy * (v0.x - r.x);
  const double c1 = ux * dy - uy * dx;
  if (c1 == 0.0) {
    bool in = c0 > 0.0;
    if (CLOSED && c0 == 0.0) in = (ux * dx + uy * dy) > 0.0;
    e.empty = e.empty || !in;
  } else if (c1 > 0.0) {  // t >= -c0 / c1
    const double n = -c0;
    if (n * e.d0 > e.n0 * c1) {
      e.n0 = n;
      e.d0 = c1;
    }
  } else {  // t <= c0 / -c1
    const double d = -c1;
    if (c0 * e.d1 < e.n1 * d) {
      e.n1 = c0;
      e.d1 = d;
    }
  }
}

// The two vertex contributions of the range [alpha, beta] of edge d (times gA).
__device__ __forceinline__ void edge_range_gradient(double alpha, double beta, double dx, double dy, double gA,
                                                    P2& g0, P2& g1) {
  const double w1 = 0.5 * (beta * beta - alpha * alpha);
  const double w0 = (beta - alpha) - w1;
  g0 = {gA * w0 * dy, -gA * w0 * dx};
  g1 = {gA * w1 * dy, -gA * w1 * dx};
}

// A lane's target-corner accumulators in LDS (slot i at base[i * kBlock], i < n = the face
// width), so the 2 * MA registers they would take stay free for occupancy.  The gradient's
// wrap-around edge adds 0.0 to slot la, which is past the row for a face of full width: such
// adds (to a slot no face vertex uses) are skipped.
struct CornerRow {
  double2* base;
  int n;
  __device__ __forceinline__ void add(int i, double x, double y) const {
    if (i >= n) return;
    double2 v = base[i * kBlock];
    v.x += x;
    v.y += y;
    base[i * kBlock] = v;
  }
};

template <int MA, int MB>
__device__ __forceinline__ void overlap_area_gradient(const P2 (&a)[MA], int la, const P2 (&b)[MB], int lb,
                                                      double gA, const CornerRow& ga, P2 (&gb)[MB]) {
  const P2 b_last = get(b, lb - 1);
  // target edges a[i] -> a[i + 1 mod la] against the source's open half-planes
#pragma unroll
  for (int i = 0; i < MA; ++i) {
    if (i < la) {
      const P2 v0 = a[i];
      const P2 v1 = (i + 1 < la) ? a[(i + 1 < MA) ? i + 1 : 0] : a[0];
      const double dx = v1.x - v0.x, dy = v1.y - v0.y;
      EdgeRange e;
#pragma unroll
      for (int j = 0; j < MB; ++j) {
        if (j < lb) {
          const P2 r = (j == 0) ? b_last : b[(j > 0) ? j - 1 : 0];
          clip_range<false>(v0, dx, dy, r, b[j].x - r.x, b[j].y - r.y, e);
        }
      }
      const double t0 = e.n0 / e.d0, t1 = e.n1 / e.d1;
      if (!e.empty && t1 > t0) {
        P2 g0, g1;
        edge_range_gradient(t0, t1, dx, dy, gA, g0, g1);
        ga.add(i, g0.x, g0.y);
        const bool wrap = !(i + 1 < la);
        if (i + 1 < MA) ga.add((i + 1 < MA) ? i + 1 : 0, wrap ? 0.0 : g1.x, wrap ? 0.0 : g1.y);
        ga.add(0, wrap ? g1.x : 0.0, wrap ? g1.y : 0.0);
      }
    }
  }
  // source edges b[j - 1 mod lb] -> b[j] against the target's closed half-planes
#pragma unroll
  for (int j = 0; j < MB; ++j) {
    if (j < lb) {
      const P2 v0 = (j == 0) ? b_last : b[(j > 0) ? j - 1 : 0];
      const P2 v1 = b[j];
      const double dx = v1.x - v0.x, dy = v1.y - v0.y;
      EdgeRange e;
#pragma unroll
      for (int i = 0; i < MA; ++i) {
        if (i < la) {
          const P2 r = a[i];
          const P2 s = (i + 1 < la) ? a[(i + 1 < MA) ? i + 1 : 0] : a[0];
          clip_range<true>(v0, dx, dy, r, s.x - r.x, s.y - r.y, e);
        }
      }
      const double t0 = e.n0 / e.d0, t1 = e.n1 / e.d1;
      if (!e.empty && t1 > t0) {
        P2 g0, g1;
        edge_range_gradient(t0, t1, dx, dy, gA, g0, g1);
        // g0 goes to vertex j - 1 mod lb, g1 to vertex j
        if (j > 0) {
          gb[(j > 0) ? j - 1 : 0].x += g0.x;
          gb[(j > 0) ? j - 1 : 0].y += g0.y;
        } else {
          add_at(gb, lb - 1, g0);
        }
        gb[j].x += g1.x;
        gb[j].y += g1.y;
      }
    }
  }
}

// Vertex gradients without atomics, in a fixed summation order (deterministic):
//   k_pair_backward         pair k -> GB[k] (its source's corners) and its target's corners: a
//                           target whose pairs all lie in the workgroup's block of kBlock
//                           consecutive pairs is summed there (TC[t], from the rows in LDS, in
//                           pair order, plus the target-area term); the pairs of the others
//                           write GA[k]
//   k_target_corners        target t -> TC[t] = sum of GA over t's pairs (pair order), plus the
//                           target-area term -sum_f gy y / area_t through the fan area, for the
//                           targets k_pair_backward did not finish
//   k_source_corners        source s -> SC[s] = sum of GB over s's pairs (by-source order)
//   k_vertex_gather         vertex v -> sum of the corner rows (TC or SC) of v, in corner order
//                           (both meshes in one launch)
// Corner rows have one P2 per face slot (stride n_max of that mesh).

// True when target pairs [begin, end) all fall in one block of kBlock consecutive pairs, the
// pairs of one k_pair_backward workgroup.
__host__ __device__ __forceinline__ bool whole_in_block(int64_t begin, int64_t end) {
  return end > begin && begin / kBlock == (end - 1) / kBlock;
}

// TC[t] slot `slot` = (sum over the target's pairs of `row(k)`, in pair order) plus the
// target-area term of that slot, as k_target_corners sums it.
template <int MAXV, typename Row>
__device__ __forceinline__ P2 target_corner(const double* __restrict__ v, const int32_t* __restrict__ f,
                                            int nmax, int64_t t, int slot, int64_t begin, int64_t end,
                                            Row row, const double* __restrict__ y,
                                            const double* __restrict__ gy, const double* __restrict__ tarea,
                                            int nf) {
  double gx = 0.0, gyv = 0.0;
  int64_t k = begin;
  for (; k + 4 <= end; k += 4) {  // four rows in flight, summed in pair order
    const double2 d0 = row(k), d1 = row(k + 1), d2 = row(k + 2), d3 = row(k + 3);
    gx += d0.x;
    gyv += d0.y;
    gx += d1.x;
    gyv += d1.y;
    gx += d2.x;
    gyv += d2.y;
    gx += d3.x;
    gyv += d3.y;
  }
  for (; k < end; ++k) {
    const double2 d = row(k);
    gx += d.x;
    gyv += d.y;
  }
  double s = 0.0;
  const double* gyt = gy + t * nf;
  const double* yt = y + t * nf;
  int j = 0;
  for (; j + 4 <= nf; j += 4) {  // four fields in flight, summed in field order
    const double g0 = gyt[j], g1 = gyt[j + 1], g2 = gyt[j + 2], g3 = gyt[j + 3];
    const double y0 = yt[j], y1 = yt[j + 1], y2 = yt[j + 2], y3 = yt[j + 3];
    s += g0 * y0;
    s += g1 * y1;
    s += g2 * y2;
    s += g3 * y3;
  }
  for (; j < nf; ++j) s += gyt[j] * yt[j];
  const double garea = -s / tarea[t];
  if (garea != 0.0) {
    P2 p[MAXV];
    int id[MAXV];
    const int len = load_face_idx<MAXV>(f, nmax, t, v, p, id);
    P2 g[MAXV];
#pragma unroll
    for (int i = 0; i < MAXV; ++i) g[i] = {0.0, 0.0};
    fan_area_adjoint<MAXV>(p, len, garea, g);
    const P2 mine = get(g, slot);
    gx += mine.x;
    gyv += mine.y;
  }
  return P2{gx, gyv};
}

// One lane per pair k = blockIdx.x * kBlock + lane.
template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) k_pair_backward(
    const double* __restrict__ tv, const int32_t* __restrict__ tf, int ntm, const double* __restrict__ sv,
    const int32_t* __restrict__ sf, int nsm, const int64_t* __restrict__ tt, const int64_t* __restrict__ ss,
    int64_t n, const double* __restrict__ F, const double* __restrict__ gy, const double* __restrict__ tarea,
    int nf, const int64_t* __restrict__ toff, const double* __restrict__ y,
    double* __restrict__ GA, double* __restrict__ GB, double* __restrict__ TC) {
  extern __shared__ double2 rows[];  // ntm slots per lane: slot i of lane l at rows[i * kBlock + l]
  __shared__ int starts[kBlock];     // lanes where a target finished here begins
  __shared__ int n_starts;
  const int lane = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * kBlock;
  const int64_t k = k0 + lane;
  if (lane == 0) n_starts = 0;
  __syncthreads();
  if (k < n) {
    const int64_t t = tt[k], s = ss[k];
    double gA = 0.0;
    const double* gyt = gy + t * nf;
    const double* Fs = F + s * nf;
    int f = 0;
    for (; f + 4 <= nf; f += 4) {  // four fields in flight, summed in field order
      const double y0 = gyt[f], y1 = gyt[f + 1], y2 = gyt[f + 2], y3 = gyt[f + 3];
      const double f0 = Fs[f], f1 = Fs[f + 1], f2 = Fs[f + 2], f3 = Fs[f + 3];
      gA += y0 * f0;
      gA += y1 * f1;
      gA += y2 * f2;
      gA += y3 * f3;
    }
    for (; f < nf; ++f) gA += gyt[f] * Fs[f];
    gA /= tarea[t];
    const CornerRow ga{rows + lane, ntm};
    for (int i = 0; i < ntm; ++i) rows[i * kBlock + lane] = make_double2(0.0, 0.0);
    P2 gb[MB];
#pragma unroll
    for (int i = 0; i < MB; ++i) gb[i] = {0.0, 0.0};
    if (gA != 0.0) {
      P2 a[MA], b[MB];
      int ai[MA], bi[MB];
      const int la = load_face_idx<MA>(tf, ntm, t, tv, a, ai);
      const int lb = load_face_idx<MB>(sf, nsm, s, sv, b, bi);
      overlap_area_gradient<MA, MB>(a, la, b, lb, gA, ga, gb);
    }
    double2* rb = reinterpret_cast<double2*>(GB) + k * nsm;
#pragma unroll
    for (int i = 0; i < MB; ++i)
      if (i < nsm) rb[i] = make_double2(gb[i].x, gb[i].y);
    const int64_t begin = toff[t], end = toff[t + 1];
    if (whole_in_block(begin, end)) {
      if (begin == k) starts[atomicAdd(&n_starts, 1)] = lane;
    } else {
      double2* ra = reinterpret_cast<double2*>(GA) + k * ntm;
      for (int i = 0; i < ntm; ++i) ra[i] = rows[i * kBlock + lane];
    }
  }
  __syncthreads();
  // The corners of the targets finished here, one lane per (target, slot); the summation
  // order of a target does not depend on which lane takes it.
  const int work = n_starts * ntm;
  for (int w = lane; w < work; w += kBlock) {
    const int q = w / ntm;
    const int slot = w - q * ntm;
    const int first = starts[q];
    const int64_t t = tt[k0 + first];
    const int64_t end = toff[t + 1];
    const P2 c = target_corner<MA>(tv, tf, ntm, t, slot, k0 + first, end,
                                   [&](int64_t kk) { return rows[slot * kBlock + (int)(kk - k0)]; }, y, gy,
                                   tarea, nf);
    reinterpret_cast<double2*>(TC)[t * ntm + slot] = make_double2(c.x, c.y);
  }
}

// Target corners, one lane per corner (t, i), so the lanes of a target read each pair row
// GA[k] together: the pairs of target t are toff[t] .. toff[t + 1] (intersect_faces order).
// The target-area term is the fan-area adjoint, evaluated by every lane of the target and kept
// for its own slot.  Targets are rest[c / nmax] when the topology lists the targets the pair
// kernel leaves (ct_regrid_rest_targets); otherwise c / nmax, skipping those whose pairs share
// one block of k_pair_backward (finished there).
template <int MAXV>
__global__ void __launch_bounds__(kBlock) k_target_corners(const double* __restrict__ v, const int32_t* __restrict__ f,
                                                           int nmax, int64_t n, const int64_t* __restrict__ rest,
                                                           const int64_t* __restrict__ toff,
                                                           const double* __restrict__ GA,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ gy,
                                                           const double* __restrict__ tarea, int nf,
                                                           double* __restrict__ TC) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n * nmax) return;
  const int64_t q = c / nmax;
  const int slot = (int)(c - q * nmax);
  const int64_t t = rest ? rest[q] : q;
  const int64_t begin = toff[t], end = toff[t + 1];
  if (!rest && whole_in_block(begin, end)) return;
  const double2* ga = reinterpret_cast<const double2*>(GA) + slot;
  const P2 r = target_corner<MAXV>(v, f, nmax, t, slot, begin, end,
                                   [&](int64_t k) { return ga[k * nmax]; }, y, gy, tarea, nf);
  reinterpret_cast<double2*>(TC)[t * nmax + slot] = make_double2(r.x, r.y);
}

// The segmented gathers below (a source's pairs, a vertex's corners) read their list in
// batches of kGatherBatch: first the batch's indices, then the rows they address, then the
// adds, in list order, of the positions that exist.  A position past the list's end reads
// the list's last entry again, so every load is unconditional and in range (the loop runs
// only for a non-empty list).  A batch costs two dependent load latencies whatever its fill;
// a loop over the entries one at a time costs two per entry, and most lists here are 2-6
// entries long.
constexpr int kGatherBatch = 4;

__device__ __forceinline__ int64_t clamp_pos(int64_t m, int64_t end) { return m < end ? m : end - 1; }

// Source corners, one lane per corner (s, i) = c / nmax, c % nmax, so the lanes of a source
// read each pair row GB[k] together: the pairs of source s are perm[soff[s] .. soff[s + 1]).
// (One lane per source with the corner sums in registers takes twice as long: a quarter of
// the lanes to hide the same chains of dependent loads.)
__global__ void __launch_bounds__(kBlock) k_source_corners(int nmax, int64_t n, const int64_t* __restrict__ soff,
                                                           const int64_t* __restrict__ perm,
                                                           const double* __restrict__ GB, double* __restrict__ SC) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n * nmax) return;
  const int64_t s = c / nmax;
  const int slot = (int)(c - s * nmax);
  double gx = 0.0, gyv = 0.0;
  const double2* gb = reinterpret_cast<const double2*>(GB) + slot;
  const int64_t end = soff[s + 1];
  for (int64_t m = soff[s]; m < end; m += kGatherBatch) {  // summed in perm order
    int64_t k[kGatherBatch];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j) k[j] = perm[clamp_pos(m + j, end)];
    double2 d[kGatherBatch];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j) d[j] = gb[k[j] * nmax];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j)
      if (m + j < end) {
        gx += d[j].x;
        gyv += d[j].y;
      }
  }
  reinterpret_cast<double2*>(SC)[c] = make_double2(gx, gyv);
}

// Per-vertex gather of one mesh: g_v[v] = sum of corner rows corners[voff[v] .. voff[v + 1])
// (row index = face * nmax + slot), summed in corner order.
struct VertexGather {
  int64_t n_vertices;
  const int64_t* voff;
  const int64_t* corners;
  const double* C;
  double* g_v;
};

__device__ __forceinline__ void gather_vertex(const VertexGather& m, int64_t v) {
  double gx = 0.0, gy = 0.0;
  const double2* rows = reinterpret_cast<const double2*>(m.C);
  const int64_t end = m.voff[v + 1];
  for (int64_t c = m.voff[v]; c < end; c += kGatherBatch) {  // summed in corner order
    int64_t idx[kGatherBatch];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j) idx[j] = m.corners[clamp_pos(c + j, end)];
    double2 d[kGatherBatch];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j) d[j] = rows[idx[j]];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j)
      if (c + j < end) {
        gx += d[j].x;
        gy += d[j].y;
      }
  }
  reinterpret_cast<double2*>(m.g_v)[v] = make_double2(gx, gy);
}

// Both meshes' vertex gathers in one launch (lanes [0, a.n) for the first, the rest for the
// second): two short latency-bound passes share the machine instead of running one after the
// other.
__global__ void __launch_bounds__(kBlock) k_vertex_gather(VertexGather a, VertexGather b) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < a.n_vertices) gather_vertex(a, v);
  else if (v - a.n_vertices < b.n_vertices) gather_vertex(b, v - a.n_vertices);
}

template <int MAXV>
__global__ void __launch_bounds__(kBlock) k_polygon_areas(const double* __restrict__ v, const int32_t* __restrict__ f,
                                                          int nmax, int64_t n, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  P2 p[MAXV];
  int id[MAXV];
  const int len = load_face_idx<MAXV>(f, nmax, k, v, p, id);
  out[k] = fan_area<MAXV>(p, len);
}


// y[t, f] = sum over the target's pairs (in pair order) of area_k F[s_k, f], / area_t.
__global__ void __launch_bounds__(kBlock) k_apply(const int64_t* __restrict__ toff, const int64_t* __restrict__ ss,
                                                  const double* __restrict__ area, const double* __restrict__ tarea,
                                                  const double* __restrict__ F, int nf, int64_t n_t,
                                                  double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_t * nf) return;
  const int64_t t = i / nf;
  const int f = (int)(i - t * nf);
  double acc = 0.0;
  int64_t k = toff[t];
  const int64_t end = toff[t + 1];
  for (; k + 4 <= end; k += 4) {  // four pairs in flight, summed in pair order
    const int64_t s0 = ss[k], s1 = ss[k + 1], s2 = ss[k + 2], s3 = ss[k + 3];
    const double a0 = area[k], a1 = area[k + 1], a2 = area[k + 2], a3 = area[k + 3];
    const double f0 = F[s0 * nf + f], f1 = F[s1 * nf + f], f2 = F[s2 * nf + f], f3 = F[s3 * nf + f];
    acc += a0 * f0;
    acc += a1 * f1;
    acc += a2 * f2;
    acc += a3 * f3;
  }
  for (; k < end; ++k) acc += area[k] * F[ss[k] * nf + f];
  y[i] = acc / tarea[t];
}

// k_apply_t with nf / FPL lanes per source, each holding FPL consecutive fields in registers:
// a pair's weight and target are loaded by the source's lanes together (one request) and its
// gy row is read as nf / FPL contiguous pieces, instead of by nf lanes that each repeat the
// index and weight loads, or by one lane whose 16 accumulators leave few wavefronts to hide
// the gathers (16 fields, 1M sources: 163 -> 137 us).  Same sums, same order, so the same bits.
template <int FPL>
__global__ void __launch_bounds__(kBlock) k_apply_t_rows(const int64_t* __restrict__ soff,
                                                         const int64_t* __restrict__ perm,
                                                         const int64_t* __restrict__ src_tgt,
                                                         const double* __restrict__ area,
                                                         const double* __restrict__ tarea,
                                                         const double* __restrict__ gy, int nf, int64_t n_s,
                                                         double* __restrict__ gF) {
  const int parts = nf / FPL;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_s * parts) return;
  const int64_t s = i / parts;
  const int f0 = (int)(i - s * parts) * FPL;
  double acc[FPL];
#pragma unroll
  for (int f = 0; f < FPL; ++f) acc[f] = 0.0;
  const int64_t end = soff[s + 1];
  for (int64_t m = soff[s]; m < end; m += kGatherBatch) {  // summed in perm order
    int64_t k[kGatherBatch], t[kGatherBatch];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j) {
      const int64_t mj = clamp_pos(m + j, end);
      k[j] = perm[mj];
      t[j] = src_tgt[mj];
    }
    double w[kGatherBatch], g[kGatherBatch][FPL];
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j) {
      w[j] = area[k[j]] / tarea[t[j]];
      const double* row = gy + t[j] * nf + f0;
#pragma unroll
      for (int f = 0; f < FPL; ++f) g[j][f] = row[f];
    }
#pragma unroll
    for (int j = 0; j < kGatherBatch; ++j)
      if (m + j < end) {
#pragma unroll
        for (int f = 0; f < FPL; ++f) acc[f] += w[j] * g[j][f];
      }
  }
  double* out = gF + s * nf + f0;
#pragma unroll
  for (int f = 0; f < FPL; ++f) out[f] = acc[f];
}

// gF[s, f] = sum over the source's pairs (perm order) of area_k / area_t gy[t_k, f].
// src_tgt[m] = tt[perm[m]] (the pairs' targets in by-source order) makes the pair index and
// its target two independent loads instead of a chain perm -> tt -> gy.
__global__ void __launch_bounds__(kBlock) k_apply_t(const int64_t* __restrict__ soff, const int64_t* __restrict__ perm,
                                                    const int64_t* __restrict__ src_tgt, const double* __restrict__ area,
                                                    const double* __restrict__ tarea, const double* __restrict__ gy,
                                                    int nf, int64_t n_s, double* __restrict__ gF) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_s * nf) return;
  const int64_t s = i / nf;
  const int f = (int)(i - s * nf);
  double acc = 0.0;
  int64_t m = soff[s];
  const int64_t end = soff[s + 1];
  for (; m + 4 <= end; m += 4) {  // four pairs in flight, summed in perm order
    const int64_t k0 = perm[m], k1 = perm[m + 1], k2 = perm[m + 2], k3 = perm[m + 3];
    const int64_t t0 = src_tgt[m], t1 = src_tgt[m + 1], t2 = src_tgt[m + 2], t3 = src_tgt[m + 3];
    const double w0 = area[k0] / tarea[t0], w1 = area[k1] / tarea[t1], w2 = area[k2] / tarea[t2],
                 w3 = area[k3] / tarea[t3];
    const double g0 = gy[t0 * nf + f], g1 = gy[t1 * nf + f], g2 = gy[t2 * nf + f], g3 = gy[t3 * nf + f];
    acc += w0 * g0;
    acc += w1 * g1;
    acc += w2 * g2;
    acc += w3 * g3;
  }
  for (; m < end; ++m) {
    const int64_t k = perm[m];
    const int64_t t = src_tgt[m];
    acc += area[k] / tarea[t] * gy[t * nf + f];
  }
  gF[i] = acc;
}

// Adjoint of barycentric_triangle.py:28-43 compute_weights (face held fixed):
//   ab = b - a, ac = c - a, ap = p - a; Aa = |ab x ap|, Ac = |ac x ap|, A = |ab x ac|
//   w = Aa / A (as (1/A) * Aa), v = Ac / A, u = 1 - v - w.
__global__ void __launch_bounds__(kBlock) k_bary_tri_backward(const double* __restrict__ pts,
                                                              const int64_t* __restrict__ face, int64_t n,
                                                              const int32_t* __restrict__ faces, int nmax,
                                                              const double* __restrict__ vtx,
                                                              const double* __restrict__ gw,
                                                              double* __restrict__ g_pts,
                                                              double* __restrict__ g_vtx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t f = face[i];
  if (f < 0) {
    if (g_pts) {
      g_pts[2 * i] = 0.0;
      g_pts[2 * i + 1] = 0.0;
    }
    return;
  }
  const int32_t* row = faces + f * nmax;
  const int ia = row[0], ib = row[1], ic = row[2];
  const P2 a = ld_vertex(vtx, ia), b = ld_vertex(vtx, ib), c = ld_vertex(vtx, ic);
  const P2 p = {pts[2 * i], pts[2 * i + 1]};
  const double abx = b.x - a.x, aby = b.y - a.y;
  const double acx = c.x - a.x, acy = c.y - a.y;
  const double apx = p.x - a.x, apy = p.y - a.y;
  const double cAa = abx * apy - aby * apx;
  const double cAc = acx * apy - acy * apx;
  const double cA = abx * acy - aby * acx;
  const double Aa = fabs(cAa), Ac = fabs(cAc), A = fabs(cA);
  const double inv = 1.0 / A;
  const double gu = gw[3 * i], gv = gw[3 * i + 1], gww = gw[3 * i + 2];
  const double gv_t = gv - gu, gw_t = gww - gu;  // u = 1 - v - w
  const double g_inv = gw_t * Aa + gv_t * Ac;
  const double gAa = gw_t * inv, gAc = gv_t * inv;
  const double gA = -g_inv * inv * inv;
  auto sgn = [](double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); };
  const double g1 = gAa * sgn(cAa), g2 = gAc * sgn(cAc), g3 = gA * sgn(cA);
  // cross(u, v) = u.x v.y - u.y v.x: d/du = (v.y, -v.x), d/dv = (-u.y, u.x)
  double gabx = g1 * apy + g3 * acy, gaby = -g1 * apx - g3 * acx;
  double gacx = g2 * apy - g3 * aby, gacy = -g2 * apx + g3 * abx;
  double gapx = -g1 * aby - g2 * acy, gapy = g1 * abx + g2 * acx;
  if (g_pts) {
    g_pts[2 * i] = gapx;
    g_pts[2 * i + 1] = gapy;
  }
  if (g_vtx) {
    atomicAdd(g_vtx + 2 * ib, gabx);
    atomicAdd(g_vtx + 2 * ib + 1, gaby);
    atomicAdd(g_vtx + 2 * ic, gacx);
    atomicAdd(g_vtx + 2 * ic + 1, gacy);
    atomicAdd(g_vtx + 2 * ia, -(gabx + gacx + gapx));
    atomicAdd(g_vtx + 2 * ia + 1, -(gaby + gacy + gapy));
  }
}

// Adjoint of barycentric_wachspress.py:38-85 compute_weights (face and branch held fixed).
// With E_k = cross(v_{k+1} - v_k, p - v_k) (edge k, v_n = v_0) and
// X_i = cross(v_i - v_{i-1}, v_{i+1} - v_{i-1}):
//   interior: w_i = 2 |X_i| / (|E_{i-1}| |E_i|), lambda_i = w_i / sum_j w_j;
//   edge case (p within tol of edge k, the first such edge in the forward's order):
//   lambda_k = 1 - t, lambda_{k+1} = t, t = |p - v_k| / |v_{k+1} - v_k|.
// The forward quantities are recomputed with the forward's own expressions, so the branch
// (interior or which edge) is the one the forward took.  Vertex gradients are added with
// atomics (a vertex is shared by many faces).
template <int MAXV>
__global__ void __launch_bounds__(kBlock) k_bary_wachspress_backward(
    const double* __restrict__ pts, const int64_t* __restrict__ face, int64_t n, const int32_t* __restrict__ faces,
    int nmax, const double* __restrict__ vtx, double tol, const double* __restrict__ gw, double* __restrict__ g_pts,
    double* __restrict__ g_vtx) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int64_t f = face[q];
  if (f < 0) {
    if (g_pts) {
      g_pts[2 * q] = 0.0;
      g_pts[2 * q + 1] = 0.0;
    }
    return;
  }
  P2 poly[MAXV];
  int vid[MAXV];
  const int len = load_face_idx<MAXV>(faces, nmax, f, vtx, poly, vid);
  const P2 p = {pts[2 * q], pts[2 * q + 1]};
  const double* g = gw + q * nmax;
  P2 last = poly[0];
#pragma unroll
  for (int k = 0; k < MAXV; ++k)
    if (k == len - 1) last = poly[k];
  P2 gv[MAXV];
#pragma unroll
  for (int k = 0; k < MAXV; ++k) gv[k] = P2{0.0, 0.0};
  P2 gv_last = {0.0, 0.0};  // added to vertex len - 1 at the end
  double gpx = 0.0, gpy = 0.0;
  // Edge-case adjoint on edge (a -> b): t = |V| / |U|, V = p - a, U = b - a.
  auto edge_adjoint = [&](P2 a, P2 b, double gt, P2* ga, P2* gb) {
    const double vx = p.x - a.x, vy = p.y - a.y, ux = b.x - a.x, uy = b.y - a.y;
    const double nv = sqrt(dot(vx, vy, vx, vy)), nu = sqrt(dot(ux, uy, ux, uy));
    double tvx = 0.0, tvy = 0.0;  // dt/dV (0 at V = 0, where |V| has no derivative)
    if (nv > 0.0) {
      tvx = vx / (nv * nu);
      tvy = vy / (nv * nu);
    }
    const double s = nv / (nu * nu * nu);
    const double tux = -s * ux, tuy = -s * uy;  // dt/dU
    gpx += gt * tvx;
    gpy += gt * tvy;
    ga->x += -gt * (tvx + tux);
    ga->y += -gt * (tvy + tuy);
    gb->x += gt * tux;
    gb->y += gt * tuy;
  };
  // Forward, in the forward's order: the edge case on edge len - 1 first, then per vertex i
  // X_i and E_i, stopping at the first edge within tolerance.
  double E[MAXV], X[MAXV];
  int edge = -1;
  double ux = poly[0].x - last.x, uy = poly[0].y - last.y;
  const double e_last = cross(ux, uy, p.x - last.x, p.y - last.y);
  if (within_perp(fabs(e_last), ux, uy, tol)) edge = len - 1;
  if (edge < 0) {
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      if (i < len && edge < 0) {
        const P2 a = (i == 0) ? last : poly[i > 0 ? i - 1 : 0];
        const P2 b = poly[i];
        const P2 c = (i + 1 == len) ? poly[0] : poly[i + 1 < MAXV ? i + 1 : 0];
        X[i] = cross(ux, uy, c.x - a.x, c.y - a.y);
        ux = c.x - b.x;
        uy = c.y - b.y;
        E[i] = cross(ux, uy, p.x - b.x, p.y - b.y);
        if (within_perp(fabs(E[i]), ux, uy, tol)) edge = i;
      }
    }
  }
  if (edge >= 0) {
    // lambda_edge = 1 - t, lambda_next = t
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      if (i == edge) {
        const int j = (i + 1 == len) ? 0 : i + 1;
        const double gt = g[j] - g[i];
        const P2 a = poly[i];
        const P2 b = (i + 1 == len) ? poly[0] : poly[i + 1 < MAXV ? i + 1 : 0];
        P2 ga = {0.0, 0.0}, gb = {0.0, 0.0};
        edge_adjoint(a, b, gt, &ga, &gb);
        gv[i].x += ga.x;
        gv[i].y += ga.y;
        if (i + 1 == len) {
          gv[0].x += gb.x;
          gv[0].y += gb.y;
        } else if (i + 1 < MAXV) {
          gv[i + 1].x += gb.x;
          gv[i + 1].y += gb.y;
        }
      }
    }
  } else {
    // interior: w_i = 2 |X_i| / (|E_{i-1}| |E_i|), E_{-1} = E_{len-1} = e_last
    double w[MAXV], S = 0.0;
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      if (i < len) {
        const double ep = (i == 0) ? e_last : E[i > 0 ? i - 1 : 0];
        w[i] = 2 * fabs(X[i]) / (fabs(ep) * fabs(E[i]));
        S += w[i];
      }
    }
    double dotgl = 0.0;
#pragma unroll
    for (int i = 0; i < MAXV; ++i)
      if (i < len) dotgl += g[i] * (w[i] / S);
    double gE[MAXV];
#pragma unroll
    for (int i = 0; i < MAXV; ++i) gE[i] = 0.0;
    double gE_last = 0.0;  // extra gradient on E_{len-1} from vertex 0
    auto sgn = [](double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); };
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      if (i < len) {
        const double gwi = (g[i] - dotgl) / S;
        const double ep = (i == 0) ? e_last : E[i > 0 ? i - 1 : 0];
        const double Ap = fabs(ep), Ai = fabs(E[i]);
        const double gX = gwi * 2.0 / (Ap * Ai) * sgn(X[i]);
        const double gAp = -gwi * w[i] / Ap, gAi = -gwi * w[i] / Ai;
        gE[i] += gAi * sgn(E[i]);
        if (i == 0) gE_last += gAp * sgn(ep);
        else gE[i > 0 ? i - 1 : 0] += gAp * sgn(ep);
        // X_i = cross(P, Q), P = v_i - v_{i-1}, Q = v_{i+1} - v_{i-1}
        const P2 a = (i == 0) ? last : poly[i > 0 ? i - 1 : 0];
        const P2 b = poly[i];
        const P2 c = (i + 1 == len) ? poly[0] : poly[i + 1 < MAXV ? i + 1 : 0];
        const double Px = b.x - a.x, Py = b.y - a.y, Qx = c.x - a.x, Qy = c.y - a.y;
        const double gPx = gX * Qy, gPy = -gX * Qx, gQx = -gX * Py, gQy = gX * Px;
        gv[i].x += gPx;
        gv[i].y += gPy;
        if (i + 1 == len) {
          gv[0].x += gQx;
          gv[0].y += gQy;
        } else if (i + 1 < MAXV) {
          gv[i + 1].x += gQx;
          gv[i + 1].y += gQy;
        }
        if (i == 0) {
          gv_last.x -= gPx + gQx;
          gv_last.y -= gPy + gQy;
        } else {
          gv[i > 0 ? i - 1 : 0].x -= gPx + gQx;
          gv[i > 0 ? i - 1 : 0].y -= gPy + gQy;
        }
      }
    }
    // E_k = cross(U, V), U = v_{k+1} - v_k, V = p - v_k
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
      if (k < len) {
        const double ge = gE[k] + (k == len - 1 ? gE_last : 0.0);
        const P2 a = poly[k];
        const P2 b = (k + 1 == len) ? poly[0] : poly[k + 1 < MAXV ? k + 1 : 0];
        const double Ux = b.x - a.x, Uy = b.y - a.y, Vx = p.x - a.x, Vy = p.y - a.y;
        const double gUx = ge * Vy, gUy = -ge * Vx, gVx = -ge * Uy, gVy = ge * Ux;
        gpx += gVx;
        gpy += gVy;
        gv[k].x -= gUx + gVx;
        gv[k].y -= gUy + gVy;
        if (k + 1 == len) {
          gv[0].x += gUx;
          gv[0].y += gUy;
        } else if (k + 1 < MAXV) {
          gv[k + 1].x += gUx;
          gv[k + 1].y += gUy;
        }
      }
    }
  }
  if (g_pts) {
    g_pts[2 * q] = gpx;
    g_pts[2 * q + 1] = gpy;
  }
  if (g_vtx) {
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
      if (k < len) {
        double gx = gv[k].x, gy = gv[k].y;
        if (k == len - 1) {
          gx += gv_last.x;
          gy += gv_last.y;
        }
        atomicAdd(g_vtx + 2 * vid[k], gx);
        atomicAdd(g_vtx + 2 * vid[k] + 1, gy);
      }
    }
  }
}

inline int vclass(int nmax) { return nmax <= 4 ? 4 : nmax <= 8 ? 8 : 16; }

#define CT_PAIR_DISPATCH(K, A, B, ...)                                                       \
  do {                                                                                       \
    const int _a = vclass(A), _b = vclass(B);                                                \
    if (_a == 4 && _b == 4) K<4, 4><<<g, kBlock, 0, st>>>(__VA_ARGS__);                      \
    else if (_a == 8 && _b == 4) K<8, 4><<<g, kBlock, 0, st>>>(__VA_ARGS__);                 \
    else if (_a == 4 && _b == 8) K<4, 8><<<g, kBlock, 0, st>>>(__VA_ARGS__);                 \
    else if (_a == 8 && _b == 8) K<8, 8><<<g, kBlock, 0, st>>>(__VA_ARGS__);                 \
    else {                                                                                   \
      set_error("regrid supports faces of at most 8 vertices");                              \
      return CT_ERR_ARG;                                                                     \
    }                                                                                        \
  } while (0)

// gF by the k_apply_t_rows form that fits n_fields.
int launch_apply_t(const int64_t* so, const int64_t* pm, const int64_t* tg, const double* area,
                   const double* tgt_area, const double* gy, int32_t n_fields, int64_t n_src, double* gF,
                   hipStream_t st) {
  if (n_fields % 4 == 0)
    k_apply_t_rows<4><<<grid_for(n_src * (n_fields / 4)), kBlock, 0, st>>>(so, pm, tg, area, tgt_area, gy, n_fields,
                                                                          n_src, gF);
  else if (n_fields % 2 == 0)
    k_apply_t_rows<2><<<grid_for(n_src * (n_fields / 2)), kBlock, 0, st>>>(so, pm, tg, area, tgt_area, gy, n_fields,
                                                                          n_src, gF);
  else
    k_apply_t<<<grid_for(n_src * n_fields), kBlock, 0, st>>>(so, pm, tg, area, tgt_area, gy, n_fields, n_src, gF);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_regrid_pair_areas(const double* tgt_vertices, const int32_t* tgt_faces, int32_t n_max_t,
                         const double* src_vertices, const int32_t* src_faces, int32_t n_max_s,
                         const int64_t* tt, const int64_t* ss, int64_t n_pairs, double* area, void* stream) {
  CT_REQUIRE(n_pairs >= 0 && n_max_t >= 3 && n_max_s >= 3, "bad sizes");
  if (n_pairs == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  const unsigned g = grid_for(n_pairs);
  CT_PAIR_DISPATCH(k_pair_areas, n_max_t, n_max_s, tgt_vertices, tgt_faces, n_max_t, src_vertices, src_faces,
                   n_max_s, tt, ss, n_pairs, area);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_polygon_areas(const double* vertices, const int32_t* faces, int32_t n_max_vert, int64_t n, double* out,
                     void* stream) {
  CT_REQUIRE(n >= 0 && n_max_vert >= 3 && n_max_vert <= 8, "bad sizes");
  if (n == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  if (n_max_vert <= 4) k_polygon_areas<4><<<grid_for(n), kBlock, 0, st>>>(vertices, faces, n_max_vert, n, out);
  else k_polygon_areas<8><<<grid_for(n), kBlock, 0, st>>>(vertices, faces, n_max_vert, n, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_regrid_apply(const int64_t* tgt_offsets, const int64_t* ss, const double* area, const double* tgt_area,
                    const double* F, int32_t n_fields, int64_t n_tgt, double* y, void* stream) {
  CT_REQUIRE(n_tgt >= 0 && n_fields >= 1, "bad sizes");
  if (n_tgt == 0) return CT_OK;
  k_apply<<<grid_for(n_tgt * n_fields), kBlock, 0, as_stream(stream)>>>(tgt_offsets, ss, area, tgt_area, F,
                                                                        n_fields, n_tgt, y);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_regrid_apply_transpose(const int64_t* src_offsets, const int64_t* perm, const int64_t* src_tgt,
                              const double* area, const double* tgt_area, const double* gy, int32_t n_fields,
                              int64_t n_src, double* gF, void* stream) {
  CT_REQUIRE(n_src >= 0 && n_fields >= 1, "bad sizes");
  if (n_src == 0) return CT_OK;
  return launch_apply_t(src_offsets, perm, src_tgt, area, tgt_area, gy, n_fields, n_src, gF, as_stream(stream));
}

int ct_barycentric_triangle_backward(const double* points, const int64_t* face_indices, int64_t n,
                                     const int32_t* faces, int32_t n_max_vert, const double* vertices,
                                     const double* g_weights, double* g_points, double* g_vertices,
                                     void* stream) {
  CT_REQUIRE(n >= 0 && n_max_vert >= 3, "bad sizes");
  if (n == 0) return CT_OK;
  k_bary_tri_backward<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(points, face_indices, n, faces, n_max_vert,
                                                                     vertices, g_weights, g_points, g_vertices);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_barycentric_wachspress_backward(const double* points, const int64_t* face_indices, int64_t n,
                                       const int32_t* faces, int32_t n_max_vert, const double* vertices, double tol,
                                       const double* g_weights, double* g_points, double* g_vertices, void* stream) {
  CT_REQUIRE(n >= 0 && n_max_vert >= 3 && n_max_vert <= 16, "n_max_vert must be in [3, 16]");
  if (n == 0) return CT_OK;
  const unsigned g = grid_for(n);
  hipStream_t st = as_stream(stream);
  if (n_max_vert <= 4)
    k_bary_wachspress_backward<4><<<g, kBlock, 0, st>>>(points, face_indices, n, faces, n_max_vert, vertices, tol,
                                                        g_weights, g_points, g_vertices);
  else if (n_max_vert <= 8)
    k_bary_wachspress_backward<8><<<g, kBlock, 0, st>>>(points, face_indices, n, faces, n_max_vert, vertices, tol,
                                                        g_weights, g_points, g_vertices);
  else
    k_bary_wachspress_backward<16><<<g, kBlock, 0, st>>>(points, face_indices, n, faces, n_max_vert, vertices, tol,
                                                         g_weights, g_points, g_vertices);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int ct_regrid_rest_targets(const ct_regrid_topology* topo, int64_t* rest_targets, int64_t* n_out, void* stream) {
  CT_REQUIRE(topo && topo->n_tgt >= 0 && n_out, "bad arguments");
  CT_REQUIRE(topo->n_tgt == 0 || (rest_targets && topo->tgt_offsets), "null arrays");
  *n_out = 0;
  if (topo->n_tgt == 0) return CT_OK;
  hipStream_t st = as_stream(stream);
  std::vector<int64_t> off(topo->n_tgt + 1), rest;
  CT_HIP(hipMemcpyAsync(off.data(), topo->tgt_offsets, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost, st));
  CT_HIP(hipStreamSynchronize(st));
  for (int64_t t = 0; t < topo->n_tgt; ++t)
    if (!whole_in_block(off[t], off[t + 1])) rest.push_back(t);
  if (!rest.empty()) {
    CT_HIP(hipMemcpyAsync(rest_targets, rest.data(), sizeof(int64_t) * rest.size(), hipMemcpyHostToDevice, st));
    CT_HIP(hipStreamSynchronize(st));
  }
  *n_out = (int64_t)rest.size();
  return CT_OK;
}

size_t ct_regrid_backward_workspace_bytes(const ct_regrid_topology* topo) {
  if (!topo) return 0;
  auto a = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t p2 = 2 * sizeof(double);
  return a(p2 * topo->n_pairs * topo->n_max_t) + a(p2 * topo->n_pairs * topo->n_max_s) +
         a(p2 * topo->n_tgt * topo->n_max_t) + a(p2 * topo->n_src * topo->n_max_s) + 256;
}

int ct_regrid_vertex_backward(const ct_regrid_topology* topo, const double* tgt_vertices,
                              const double* src_vertices, const double* F, const double* y, const double* gy,
                              const double* tgt_area, int32_t n_fields, void* workspace, size_t workspace_bytes,
                              double* g_tgt_vertices, double* g_src_vertices, void* stream) {
  CT_REQUIRE(topo && topo->n_pairs >= 0 && topo->n_tgt >= 0 && topo->n_src >= 0 && n_fields >= 1, "bad sizes");
  const int n_max_t = topo->n_max_t, n_max_s = topo->n_max_s;
  CT_REQUIRE(n_max_t >= 3 && n_max_t <= 8 && n_max_s >= 3 && n_max_s <= 8,
             "regrid supports faces of 3..8 vertices");
  CT_REQUIRE(g_tgt_vertices && g_src_vertices, "null gradient arrays");
  const size_t need = ct_regrid_backward_workspace_bytes(topo);
  if (workspace_bytes < need) {
    set_error("regrid backward workspace too small: %zu < %zu", workspace_bytes, need);
    return CT_ERR_WORKSPACE;
  }
  auto a = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t p2 = 2 * sizeof(double);
  char* w = static_cast<char*>(workspace);
  double* GA = reinterpret_cast<double*>(w);
  w += a(p2 * topo->n_pairs * n_max_t);
  double* GB = reinterpret_cast<double*>(w);
  w += a(p2 * topo->n_pairs * n_max_s);
  double* TC = reinterpret_cast<double*>(w);
  w += a(p2 * topo->n_tgt * n_max_t);
  double* SC = reinterpret_cast<double*>(w);
  hipStream_t st = as_stream(stream);
  const int64_t n_pairs = topo->n_pairs, n_tgt = topo->n_tgt, n_src = topo->n_src;
  if (n_pairs > 0) {
    const unsigned g = grid_for(n_pairs);
    const size_t lds = sizeof(double2) * kBlock * n_max_t;  // the per-lane corner rows
#define CT_PAIR_BACKWARD(X, Y)                                                                                  \
  k_pair_backward<X, Y><<<g, kBlock, lds, st>>>(tgt_vertices, topo->tgt_faces, n_max_t, src_vertices,           \
                                                topo->src_faces, n_max_s, topo->tt, topo->ss, n_pairs, F, gy,   \
                                                tgt_area, n_fields, topo->tgt_offsets, y, GA, GB, TC)
    const int ca = vclass(n_max_t), cb = vclass(n_max_s);
    if (ca == 4 && cb == 4) CT_PAIR_BACKWARD(4, 4);
    else if (ca == 8 && cb == 4) CT_PAIR_BACKWARD(8, 4);
    else if (ca == 4 && cb == 8) CT_PAIR_BACKWARD(4, 8);
    else CT_PAIR_BACKWARD(8, 8);
#undef CT_PAIR_BACKWARD
    CT_LAUNCH_CHECK();
  }
  const int64_t* rest = topo->rest_targets;
  const int64_t n_rt = rest ? topo->n_rest_targets : n_tgt;  // targets the corner pass visits
  if (n_rt > 0) {
    const unsigned g = grid_for(n_rt * n_max_t);
    if (n_max_t <= 4)
      k_target_corners<4><<<g, kBlock, 0, st>>>(tgt_vertices, topo->tgt_faces, n_max_t, n_rt, rest, topo->tgt_offsets,
                                                GA, y, gy, tgt_area, n_fields, TC);
    else
      k_target_corners<8><<<g, kBlock, 0, st>>>(tgt_vertices, topo->tgt_faces, n_max_t, n_rt, rest, topo->tgt_offsets,
                                                GA, y, gy, tgt_area, n_fields, TC);
    CT_LAUNCH_CHECK();
  }
  if (n_src > 0) {
    k_source_corners<<<grid_for(n_src * n_max_s), kBlock, 0, st>>>(n_max_s, n_src, topo->src_offsets,
                                                                    topo->perm_by_src, GB, SC);
    CT_LAUNCH_CHECK();
  }
  const int64_t n_gather = topo->n_tgt_vertices + topo->n_src_vertices;
  if (n_gather > 0) {
    const VertexGather tg{topo->n_tgt_vertices, topo->tgt_vertex_offsets, topo->tgt_vertex_corners, TC,
                          g_tgt_vertices};
    const VertexGather sg{topo->n_src_vertices, topo->src_vertex_offsets, topo->src_vertex_corners, SC,
                          g_src_vertices};
    k_vertex_gather<<<grid_for(n_gather), kBlock, 0, st>>>(tg, sg);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

}  // extern "C"
