// ct_regrid.hip -- conservative area-weighted regridding and its adjoint (no reference
// counterpart; the forward is the pattern of examples/spatial_indexing_2d_grids.py:268-323).
//
//   y[t, f] = sum_{k: t_k = t} A_k * F[s_k, f] / area_t
//
// A_k is the Sutherland-Hodgman overlap area of pair k (subject = target polygon t_k,
// clipper = source polygon s_k; sutherland_hodgman.py:84-148) and area_t the fan area of
// the target polygon (geometry_utils.py:82-95).  The pair topology (which pairs, and the
// path each clip takes) is held fixed; the adjoint differentiates the arithmetic of that
// path with respect to F, the source vertices and the target vertices.
//
// Pair lists are grouped by target (ascending t, as intersect_faces returns them): the
// forward is a deterministic segmented sum per target.  The transpose uses a by-source
// permutation of the pairs (also deterministic).  Vertex gradients are scattered with FP64
// atomics (summation order may vary between runs; each contribution is exact).
#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {
namespace {

// -------------------------------------------------------------------- forward clip
// One Sutherland-Hodgman stage (clip edge r -> s) on polygon `in` of length `len`,
// recording for every output vertex its provenance: src[o] = j for a copy of in[j], or
// src[o] = -(1 + j) for the intersection of input edge (aprev[o] -> j) with the clip line.
template <int CAP>
__device__ __forceinline__ int sh_stage(const P2 (&in)[CAP], int len, P2 r, P2 s, P2 (&out)[CAP],
                                        int (&src)[CAP], int (&aprev)[CAP]) {
  const double ux = s.x - r.x, uy = s.y - r.y;
  const double nx = -uy, ny = ux;
  int n_out = 0;
  int ai = len - 1;
  P2 a = in[0];
  for (int j = 0; j < CAP; ++j)
    if (j == len - 1) a = in[j];
  bool a_in = ux * (a.y - r.y) > uy * (a.x - r.x);
  for (int j = 0; j < len; ++j) {
    const P2 b = in[j];
    const double vx = b.x - a.x, vy = b.y - a.y;
    if (vx == 0 && vy == 0) continue;
    bool b_in = ux * (b.y - r.y) > uy * (b.x - r.x);
    const double nw = nx * (r.x - a.x) + ny * (r.y - a.y);
    const double nv = nx * vx + ny * vy;
    if (b_in) {
      if (!a_in && nv != 0 && n_out < CAP) {
        const double t = nw / nv;
        out[n_out] = {a.x + t * vx, a.y + t * vy};
        src[n_out] = -(1 + j);
        aprev[n_out] = ai;
        ++n_out;
      }
      if (n_out < CAP) {
        out[n_out] = b;
        src[n_out] = j;
        aprev[n_out] = ai;
        ++n_out;
      }
    } else if (a_in) {
      if (nv != 0) {
        if (n_out < CAP) {
          const double t = nw / nv;
          out[n_out] = {a.x + t * vx, a.y + t * vy};
          src[n_out] = -(1 + j);
          aprev[n_out] = ai;
          ++n_out;
        }
      } else {
        b_in = true;
        if (n_out < CAP) {
          out[n_out] = b;
          src[n_out] = j;
          aprev[n_out] = ai;
          ++n_out;
        }
      }
    }
    a = b;
    ai = j;
    a_in = b_in;
  }
  return n_out;
}

// Runs stages 0 .. upto-1 of the clip (skipping zero-length clip edges, as the reference
// does) and returns the polygon entering stage `upto`; *dead = true if a stage left fewer
// than 3 vertices (area 0, zero gradient).
template <int MA, int MB, int CAP>
__device__ __forceinline__ int sh_run(const P2 (&subj)[MA], int ls, const P2 (&clip)[MB], int lc, int upto,
                                      P2 (&poly)[CAP], bool* dead) {
  for (int i = 0; i < CAP; ++i) poly[i] = (i < ls) ? subj[i < MA ? i : 0] : P2{0.0, 0.0};
  int len = ls;
  *dead = false;
  P2 tmp[CAP];
  int src[CAP], aprev[CAP];
  for (int ci = 0; ci < upto; ++ci) {
    const P2 r = clip[(ci == 0) ? lc - 1 : ci - 1];
    const P2 s = clip[ci];
    if (s.x - r.x == 0 && s.y - r.y == 0) continue;
    len = sh_stage<CAP>(poly, len, r, s, tmp, src, aprev);
    for (int i = 0; i < CAP; ++i) poly[i] = tmp[i];
    if (len < 3) {
      *dead = true;
      return len;
    }
  }
  return len;
}

// geometry_utils.py:82-95 polygon_area (fan from vertex 0, abs per triangle)
template <int CAP>
__device__ __forceinline__ double fan_area(const P2 (&p)[CAP], int len) {
  double area = 0.0;
  double ux = p[1].x - p[0].x, uy = p[1].y - p[0].y;
  for (int i = 2; i < len; ++i) {
    const double vx = p[0].x - p[i].x, vy = p[0].y - p[i].y;
    area += fabs(ux * vy - uy * vx);
    ux = vx;
    uy = vy;
  }
  return 0.5 * area;
}

// Adjoint of fan_area: g[i] += d(area)/d(p[i]) * garea.
template <int CAP>
__device__ __forceinline__ void fan_area_adjoint(const P2 (&p)[CAP], int len, double garea, P2 (&g)[CAP]) {
  for (int i = 2; i < len; ++i) {
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

// Adjoint of one SH stage: maps the output-vertex adjoints gout to input-vertex adjoints
// gin and the clip-edge endpoint adjoints gr, gs.
template <int CAP>
__device__ __forceinline__ void sh_stage_adjoint(const P2 (&in)[CAP], int len, P2 r, P2 s, int n_out,
                                                 const int (&src)[CAP], const int (&aprev)[CAP],
                                                 const P2 (&gout)[CAP], P2 (&gin)[CAP], P2* gr, P2* gs) {
  for (int i = 0; i < CAP; ++i) gin[i] = {0.0, 0.0};
  const double ux = s.x - r.x, uy = s.y - r.y;
  const double nx = -uy, ny = ux;
  double gnx = 0.0, gny = 0.0;
  for (int o = 0; o < n_out; ++o) {
    const P2 gp = gout[o];
    if (src[o] >= 0) {
      gin[src[o]].x += gp.x;
      gin[src[o]].y += gp.y;
      continue;
    }
    const int j = -src[o] - 1;
    const int ia = aprev[o];
    const P2 a = in[ia], b = in[j];
    const double vx = b.x - a.x, vy = b.y - a.y;
    const double wx = r.x - a.x, wy = r.y - a.y;
    const double nw = nx * wx + ny * wy;
    const double nv = nx * vx + ny * vy;
    const double t = nw / nv;
    // p = a + t V
    const double gt = gp.x * vx + gp.y * vy;
    double gax = gp.x, gay = gp.y;
    double gvx = gp.x * t, gvy = gp.y * t;
    const double gnw = gt / nv;
    const double gnv = -gt * t / nv;
    // nw = N . (r - a)
    gnx += gnw * wx;
    gny += gnw * wy;
    gr->x += gnw * nx;
    gr->y += gnw * ny;
    gax -= gnw * nx;
    gay -= gnw * ny;
    // nv = N . V
    gnx += gnv * vx;
    gny += gnv * vy;
    gvx += gnv * nx;
    gvy += gnv * ny;
    // V = b - a
    gin[j].x += gvx;
    gin[j].y += gvy;
    gax -= gvx;
    gay -= gvy;
    gin[ia].x += gax;
    gin[ia].y += gay;
  }
  // N = (r.y - s.y, s.x - r.x)
  gr->y += gnx;
  gs->y -= gnx;
  gs->x += gny;
  gr->x -= gny;
}

template <int MAXV>
__device__ __forceinline__ int load_face_idx(const int32_t* __restrict__ f, int nmax, int64_t k,
                                             const double* __restrict__ v, P2 (&poly)[MAXV], int (&vid)[MAXV]) {
  const int32_t* row = f + k * nmax;
  int len = nmax < MAXV ? nmax : MAXV;
  for (int i = 3; i < nmax && i < MAXV; ++i) {
    if (row[i] == kFill) {
      len = i;
      break;
    }
  }
  for (int i = 0; i < MAXV; ++i) {
    vid[i] = (i < len) ? row[i] : 0;
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
  constexpr int CAP = MA + MB;
  P2 a[MA], b[MB];
  int ai[MA], bi[MB];
  const int la = load_face_idx<MA>(tf, ntm, tt[k], tv, a, ai);
  const int lb = load_face_idx<MB>(sf, nsm, ss[k], sv, b, bi);
  P2 poly[CAP];
  bool dead;
  const int len = sh_run<MA, MB, CAP>(a, la, b, lb, lb, poly, &dead);
  area[k] = dead ? 0.0 : fan_area<CAP>(poly, len);
}

// Per pair: gA = sum_f gy[t,f] F[s,f] / area_t, then the reverse sweep through the clip.
template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) k_pair_backward(
    const double* __restrict__ tv, const int32_t* __restrict__ tf, int ntm, const double* __restrict__ sv,
    const int32_t* __restrict__ sf, int nsm, const int64_t* __restrict__ tt, const int64_t* __restrict__ ss,
    int64_t n, const double* __restrict__ F, const double* __restrict__ gy, const double* __restrict__ tarea,
    int nf, double* __restrict__ g_tv, double* __restrict__ g_sv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t t = tt[k], s = ss[k];
  double gA = 0.0;
  for (int f = 0; f < nf; ++f) gA += gy[t * nf + f] * F[s * nf + f];
  gA /= tarea[t];
  if (gA == 0.0) return;
  constexpr int CAP = MA + MB;
  P2 a[MA], b[MB];
  int ai[MA], bi[MB];
  const int la = load_face_idx<MA>(tf, ntm, t, tv, a, ai);
  const int lb = load_face_idx<MB>(sf, nsm, s, sv, b, bi);
  // forward to the end for the final polygon and the area adjoint
  P2 poly[CAP];
  bool dead;
  const int len = sh_run<MA, MB, CAP>(a, la, b, lb, lb, poly, &dead);
  if (dead) return;
  P2 g[CAP];
  for (int i = 0; i < CAP; ++i) g[i] = {0.0, 0.0};
  fan_area_adjoint<CAP>(poly, len, gA, g);
  P2 gclip[MB];
  for (int i = 0; i < MB; ++i) gclip[i] = {0.0, 0.0};
  // reverse sweep over the clip stages (recomputing each stage's input)
  for (int ci = lb - 1; ci >= 0; --ci) {
    const int ri = (ci == 0) ? lb - 1 : ci - 1;
    const P2 r = b[ri], sp = b[ci];
    if (sp.x - r.x == 0 && sp.y - r.y == 0) continue;
    P2 in[CAP];
    bool d2;
    const int lin = sh_run<MA, MB, CAP>(a, la, b, lb, ci, in, &d2);
    P2 out[CAP];
    int src[CAP], aprev[CAP];
    const int nout = sh_stage<CAP>(in, lin, r, sp, out, src, aprev);
    (void)nout;
    P2 gin[CAP], gr = {0.0, 0.0}, gs = {0.0, 0.0};
    sh_stage_adjoint<CAP>(in, lin, r, sp, nout, src, aprev, g, gin, &gr, &gs);
    for (int i = 0; i < CAP; ++i) g[i] = gin[i];
    gclip[ri].x += gr.x;
    gclip[ri].y += gr.y;
    gclip[ci].x += gs.x;
    gclip[ci].y += gs.y;
  }
  for (int i = 0; i < la; ++i) {
    atomicAdd(g_tv + 2 * ai[i], g[i].x);
    atomicAdd(g_tv + 2 * ai[i] + 1, g[i].y);
  }
  for (int i = 0; i < lb; ++i) {
    atomicAdd(g_sv + 2 * bi[i], gclip[i].x);
    atomicAdd(g_sv + 2 * bi[i] + 1, gclip[i].y);
  }
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

// d(area_t): g_area_t = -sum_f gy[t,f] y[t,f] / area_t, through the fan area.
template <int MAXV>
__global__ void __launch_bounds__(kBlock) k_polygon_area_backward(const double* __restrict__ v,
                                                                  const int32_t* __restrict__ f, int nmax,
                                                                  int64_t n, const double* __restrict__ y,
                                                                  const double* __restrict__ gy,
                                                                  const double* __restrict__ tarea, int nf,
                                                                  double* __restrict__ g_v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s = 0.0;
  for (int k = 0; k < nf; ++k) s += gy[t * nf + k] * y[t * nf + k];
  const double garea = -s / tarea[t];
  if (garea == 0.0) return;
  P2 p[MAXV];
  int id[MAXV];
  const int len = load_face_idx<MAXV>(f, nmax, t, v, p, id);
  P2 g[MAXV];
  for (int i = 0; i < MAXV; ++i) g[i] = {0.0, 0.0};
  fan_area_adjoint<MAXV>(p, len, garea, g);
  for (int i = 0; i < len; ++i) {
    atomicAdd(g_v + 2 * id[i], g[i].x);
    atomicAdd(g_v + 2 * id[i] + 1, g[i].y);
  }
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
  for (int64_t k = toff[t]; k < toff[t + 1]; ++k) acc += area[k] * F[ss[k] * nf + f];
  y[i] = acc / tarea[t];
}

// gF[s, f] = sum over the source's pairs (perm order) of area_k / area_t gy[t_k, f].
__global__ void __launch_bounds__(kBlock) k_apply_t(const int64_t* __restrict__ soff, const int64_t* __restrict__ perm,
                                                    const int64_t* __restrict__ tt, const double* __restrict__ area,
                                                    const double* __restrict__ tarea, const double* __restrict__ gy,
                                                    int nf, int64_t n_s, double* __restrict__ gF) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_s * nf) return;
  const int64_t s = i / nf;
  const int f = (int)(i - s * nf);
  double acc = 0.0;
  for (int64_t m = soff[s]; m < soff[s + 1]; ++m) {
    const int64_t k = perm[m];
    const int64_t t = tt[k];
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

int ct_regrid_apply_transpose(const int64_t* src_offsets, const int64_t* perm, const int64_t* tt,
                              const double* area, const double* tgt_area, const double* gy, int32_t n_fields,
                              int64_t n_src, double* gF, void* stream) {
  CT_REQUIRE(n_src >= 0 && n_fields >= 1, "bad sizes");
  if (n_src == 0) return CT_OK;
  k_apply_t<<<grid_for(n_src * n_fields), kBlock, 0, as_stream(stream)>>>(src_offsets, perm, tt, area, tgt_area,
                                                                          gy, n_fields, n_src, gF);
  CT_LAUNCH_CHECK();
  return CT_OK;
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

int ct_regrid_vertex_backward(const double* tgt_vertices, const int32_t* tgt_faces, int32_t n_max_t,
                              const double* src_vertices, const int32_t* src_faces, int32_t n_max_s,
                              const int64_t* tt, const int64_t* ss, int64_t n_pairs, int64_t n_tgt,
                              const double* F, const double* y, const double* gy, const double* tgt_area,
                              int32_t n_fields, double* g_tgt_vertices, double* g_src_vertices, void* stream) {
  CT_REQUIRE(n_pairs >= 0 && n_tgt >= 0 && n_fields >= 1 && n_max_t >= 3 && n_max_s >= 3, "bad sizes");
  hipStream_t st = as_stream(stream);
  if (n_pairs > 0) {
    const unsigned g = grid_for(n_pairs);
    CT_PAIR_DISPATCH(k_pair_backward, n_max_t, n_max_s, tgt_vertices, tgt_faces, n_max_t, src_vertices,
                     src_faces, n_max_s, tt, ss, n_pairs, F, gy, tgt_area, n_fields, g_tgt_vertices,
                     g_src_vertices);
    CT_LAUNCH_CHECK();
  }
  if (n_tgt > 0) {
    if (n_max_t <= 4)
      k_polygon_area_backward<4><<<grid_for(n_tgt), kBlock, 0, st>>>(tgt_vertices, tgt_faces, n_max_t, n_tgt, y, gy,
                                                                     tgt_area, n_fields, g_tgt_vertices);
    else
      k_polygon_area_backward<8><<<grid_for(n_tgt), kBlock, 0, st>>>(tgt_vertices, tgt_faces, n_max_t, n_tgt, y, gy,
                                                                     tgt_area, n_fields, g_tgt_vertices);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

}  // extern "C"
