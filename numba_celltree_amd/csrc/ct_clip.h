// ct_clip.h -- Sutherland-Hodgman polygon clipping held entirely in registers, shared by the
// face-pair kernels (ct_faces.hip) and the regrid forward / adjoint (ct_regrid.hip).
//
// A per-lane array indexed by a runtime value is placed in scratch memory by the compiler;
// for clipping (an output vertex is appended at a runtime position, an input vertex read at
// a runtime position) that made the kernels bound by scratch traffic.  Here every loop over
// a polygon is unrolled to its compile-time capacity with a `j < len` guard, and the two
// runtime-indexed accesses go through put() / get(): one compare + select per slot, which
// keeps the arrays in VGPRs.  The arithmetic and the branch structure are the reference's
// (sutherland_hodgman.py:56-148), operation for operation.
#pragma once

#include "ct_internal.h"

namespace ct {
namespace {

// arr[n] = v for a runtime n, as one select per slot (written as a conditional expression,
// which compiles to v_cndmask rather than a branch per slot).
__device__ __forceinline__ P2 sel(bool c, P2 a, P2 b) { return P2{c ? a.x : b.x, c ? a.y : b.y}; }
__device__ __forceinline__ int sel(bool c, int a, int b) { return c ? a : b; }

template <typename T, int CAP>
__device__ __forceinline__ void put(T (&arr)[CAP], int n, T v) {
#pragma unroll
  for (int k = 0; k < CAP; ++k) arr[k] = sel(k == n, v, arr[k]);
}

// arr[n] for a runtime n in [0, CAP).
template <typename T, int CAP>
__device__ __forceinline__ T get(const T (&arr)[CAP], int n) {
  T v = arr[0];
#pragma unroll
  for (int k = 1; k < CAP; ++k) v = sel(k == n, arr[k], v);
  return v;
}

// sutherland_hodgman.py:56-60 (strict)
__device__ __forceinline__ bool sh_in(P2 p, P2 r, double ux, double uy) {
  return ux * (p.y - r.y) > uy * (p.x - r.x);
}

// One clip stage (sutherland_hodgman.py:96-145): polygon `in` (len vertices) against the clip
// edge r -> s, into `out`; returns the output length.  With PROV, src[o] records where output
// vertex o came from: j for a copy of in[j], -(1 + j) for the intersection of input edge
// (aprev[o] -> j) with the clip line.  Writes past CAP are dropped (a convex n-gon clipped by
// a convex m-gon has at most n + m vertices).
template <int CAP, bool PROV>
__device__ __forceinline__ int sh_stage(const P2 (&in)[CAP], int len, P2 r, P2 s, P2 (&out)[CAP],
                                        int (&src)[CAP], int (&aprev)[CAP]) {
  const double ux = s.x - r.x, uy = s.y - r.y;
  const double nx = -uy, ny = ux;
  int n_out = 0;
  P2 a = get(in, len - 1);
  int ai = len - 1;
  bool a_in = sh_in(a, r, ux, uy);
#pragma unroll
  for (int j = 0; j < CAP; ++j) {
    if (j < len) {
      const P2 b = in[j];
      const double vx = b.x - a.x, vy = b.y - a.y;
      if (!(vx == 0 && vy == 0)) {  // zero-length edge: skipped, `a` kept
        bool b_in = sh_in(b, r, ux, uy);
        const double nv = nx * vx + ny * vy;
        bool cross_pt = false, copy_b = false;
        if (b_in) {
          cross_pt = !a_in && nv != 0;
          copy_b = true;
        } else if (a_in) {
          if (nv != 0) {
            cross_pt = true;
          } else {  // parallel: b is taken as inside (sutherland_hodgman.py:133-137)
            b_in = true;
            copy_b = true;
          }
        }
        if (cross_pt && n_out < CAP) {
          const double nw = nx * (r.x - a.x) + ny * (r.y - a.y);
          const double t = nw / nv;
          put(out, n_out, P2{a.x + t * vx, a.y + t * vy});
          if (PROV) {
            put(src, n_out, -(1 + j));
            put(aprev, n_out, ai);
          }
          ++n_out;
        }
        if (copy_b && n_out < CAP) {
          put(out, n_out, b);
          if (PROV) {
            put(src, n_out, j);
            put(aprev, n_out, ai);
          }
          ++n_out;
        }
        a = b;
        ai = j;
        a_in = b_in;
      }
    }
  }
  return n_out;
}

// Stages 0 .. upto-1 of clipping `subj` (ls vertices) by `clip` (lc vertices), skipping
// zero-length clip edges (sutherland_hodgman.py:105-107); returns the polygon entering stage
// `upto`, or a length < 3 as soon as a stage leaves fewer than 3 vertices (area 0).
template <int MA, int MB, int CAP>
__device__ __forceinline__ int sh_run(const P2 (&subj)[MA], int ls, const P2 (&clip)[MB], int lc, int upto,
                                      P2 (&poly)[CAP]) {
#pragma unroll
  for (int i = 0; i < CAP; ++i) poly[i] = (i < MA) ? subj[i < MA ? i : 0] : P2{0.0, 0.0};
  int len = ls;
  P2 r = get(clip, lc - 1);
  for (int ci = 0; ci < upto; ++ci) {
    const P2 s = get(clip, ci);
    if (!(s.x - r.x == 0 && s.y - r.y == 0)) {
      P2 tmp[CAP];
      int src[CAP], aprev[CAP];
      len = sh_stage<CAP, false>(poly, len, r, s, tmp, src, aprev);
#pragma unroll
      for (int i = 0; i < CAP; ++i) poly[i] = tmp[i];
      if (len < 3) return len;
    }
    r = s;
  }
  return len;
}

// The same clip with plain runtime indexing, for the large vertex classes (CAP > 16): the
// select chains grow as CAP^2 per stage, and so do code size and compile time, while faces
// of more than 8 vertices are rare.  Same operations in the same order as sh_stage.
template <int MA, int MB>
__device__ __forceinline__ double clip_area_indexed(const P2 (&subj)[MA], int ls, const P2 (&clip)[MB], int lc) {
  constexpr int CAP = MA + MB;
  P2 outp[CAP], inp[CAP];
  for (int i = 0; i < CAP; ++i) outp[i] = (i < MA) ? subj[i < MA ? i : 0] : P2{0.0, 0.0};
  int n_out = ls;
  P2 r = clip[lc - 1];
  for (int ci = 0; ci < lc; ++ci) {
    const P2 s = clip[ci];
    const double ux = s.x - r.x, uy = s.y - r.y;
    if (!(ux == 0 && uy == 0)) {
      const double nx = -uy, ny = ux;
      const int length = n_out;
      for (int j = 0; j < CAP; ++j) inp[j] = outp[j];
      n_out = 0;
      P2 a = inp[length - 1];
      bool a_in = sh_in(a, r, ux, uy);
      for (int j = 0; j < length; ++j) {
        const P2 b = inp[j];
        const double vx = b.x - a.x, vy = b.y - a.y;
        if (vx == 0 && vy == 0) continue;
        bool b_in = sh_in(b, r, ux, uy);
        const double nv = nx * vx + ny * vy;
        bool cross_pt = false, copy_b = false;
        if (b_in) {
          cross_pt = !a_in && nv != 0;
          copy_b = true;
        } else if (a_in) {
          if (nv != 0) {
            cross_pt = true;
          } else {
            b_in = true;
            copy_b = true;
          }
        }
        if (cross_pt && n_out < CAP) {
          const double nw = nx * (r.x - a.x) + ny * (r.y - a.y);
          const double t = nw / nv;
          outp[n_out++] = P2{a.x + t * vx, a.y + t * vy};
        }
        if (copy_b && n_out < CAP) outp[n_out++] = b;
        a = b;
        a_in = b_in;
      }
      if (n_out < 3) return 0.0;
    }
    r = s;
  }
  return polygon_area<CAP>(outp, n_out);
}

// sutherland_hodgman.py:84-148 polygon_polygon_clip_area.
template <int MA, int MB>
__device__ __forceinline__ double clip_area(const P2 (&subj)[MA], int ls, const P2 (&clip)[MB], int lc) {
  constexpr int CAP = MA + MB;
  if constexpr (CAP > 16) {
    return clip_area_indexed<MA, MB>(subj, ls, clip, lc);
  } else {
    P2 poly[CAP];
    const int len = sh_run<MA, MB, CAP>(subj, ls, clip, lc, lc, poly);
    return (len < 3) ? 0.0 : polygon_area<CAP>(poly, len);
  }
}

}  // namespace
}  // namespace ct
