// ct_clip.h -- Sutherland-Hodgman polygon clipping for the face-pair kernels (ct_faces.hip)
// and the regrid forward (ct_regrid.hip).
//
// A per-lane array indexed by a runtime value is placed in scratch memory by the compiler;
// for clipping (an output vertex is appended at a runtime position, an input vertex read at
// a runtime position) that made the kernels bound by scratch traffic.  Here every loop over
// a polygon is unrolled to its compile-time capacity with a `j < len` guard, so a stage reads
// its input polygon from registers with static indices, and the runtime-indexed accesses (the
// appended output vertices, the input's last vertex) go to the lane's row in LDS (its slots
// past the tenth, used only by polygons of more than 10 vertices, in private memory).  After
// each stage the row is read back into registers.  The arithmetic and the branch structure are the
// reference's (sutherland_hodgman.py:56-148), operation for operation.
#pragma once

#include "ct_internal.h"

namespace ct {
namespace {

// a ? x : y for points and ints, as selects (v_cndmask) rather than branches.
__device__ __forceinline__ P2 sel(bool c, P2 a, P2 b) { return P2{c ? a.x : b.x, c ? a.y : b.y}; }
__device__ __forceinline__ int sel(bool c, int a, int b) { return c ? a : b; }

// arr[n] for a runtime n in [0, CAP): one select per slot.
template <typename T, int CAP>
__device__ __forceinline__ T get(const T (&arr)[CAP], int n) {
  T v = arr[0];
#pragma unroll
  for (int k = 1; k < CAP; ++k) v = sel(k == n, arr[k], v);
  return v;
}

// Slots of a clip row kept in LDS.  A CU has 160 KiB of LDS: 10 slots (40 KiB per workgroup
// of kBlock lanes) let 4 workgroups share a CU, as many as the clip kernels' registers allow,
// where the 12 slots of the 8 + 4 vertex classes would let only 3.  Slots past these are
// used only by clipped polygons of more than 10 vertices (a convex n-gon clipped by a convex
// m-gon has at most n + m) and live in a small per-lane private array.
constexpr int kClipLdsSlots = 10;

// A lane's clip row of CAP slots: slot s < kLds in LDS at base[s * kBlock], lane-interleaved,
// so the lanes of a wavefront touch 64 consecutive 16-byte words whatever slot each of them
// uses (no bank conflicts); slot s >= kLds at extra[s - kLds].  A stage appends its output
// vertices here at runtime positions (one 16-byte store each); the selects that would keep
// a runtime-indexed output in registers cost CAP compares and 4 * CAP v_cndmask per appended
// vertex, and were most of the clip's instructions.
template <int CAP>
struct ClipRow {
  static constexpr int kLds = CAP < kClipLdsSlots ? CAP : kClipLdsSlots;
  static constexpr int kExtra = CAP - kLds;
  double2* base;
  double2* extra;
  __device__ __forceinline__ void put(int s, P2 v) const {
    const double2 d{v.x, v.y};
    if (kExtra == 0 || s < kLds)
      base[s * kBlock] = d;
    else
      extra[s - kLds] = d;
  }
  __device__ __forceinline__ P2 get(int s) const {
    double2 d;
    if (kExtra == 0 || s < kLds)
      d = base[s * kBlock];
    else
      d = extra[s - kLds];
    return P2{d.x, d.y};
  }
};

// Clip rows of capacity CAP for the kBlock lanes of a workgroup: min(CAP, 10) * 4 KiB of
// LDS.  Every lane touches only its own slots, so no barrier is needed.
#define CT_CLIP_ROWS(CAP, name)                                                          \
  __shared__ double2 name##_lds[ClipRow<(CAP)>::kLds * kBlock];                          \
  double2 name##_extra[ClipRow<(CAP)>::kExtra > 0 ? ClipRow<(CAP)>::kExtra : 1];         \
  const ClipRow<(CAP)> name { name##_lds + threadIdx.x, name##_extra }

// sutherland_hodgman.py:56-60 (strict)
__device__ __forceinline__ bool sh_in(P2 p, P2 r, double ux, double uy) {
  return ux * (p.y - r.y) > uy * (p.x - r.x);
}

// One clip stage (sutherland_hodgman.py:96-145): polygon `in` (len vertices, also held in
// `row` slots [0, len)) against the clip edge r -> s; the output polygon is written to `row`
// (over the input, which is read from registers) and its length returned.  Writes past CAP
// are dropped (a convex n-gon clipped by a convex m-gon has at most n + m vertices).
template <int CAP, class Row>
__device__ __forceinline__ int sh_stage(const P2 (&in)[CAP], int len, P2 r, P2 s, const Row& row) {
  const double ux = s.x - r.x, uy = s.y - r.y;
  const double nx = -uy, ny = ux;
  int n_out = 0;
  P2 a = row.get(len - 1);
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
          row.put(n_out, P2{a.x + t * vx, a.y + t * vy});
          ++n_out;
        }
        if (copy_b && n_out < CAP) {
          row.put(n_out, b);
          ++n_out;
        }
        a = b;
        a_in = b_in;
      }
    }
  }
  return n_out;
}

// Stages 0 .. upto-1 of clipping `subj` (ls vertices) by `clip` (lc vertices), skipping
// zero-length clip edges (sutherland_hodgman.py:105-107); returns the polygon entering stage
// `upto` in `poly`, or a length < 3 as soon as a stage leaves fewer than 3 vertices (area 0).
template <int MA, int MB, int CAP, class Row>
__device__ __forceinline__ int sh_run(const P2 (&subj)[MA], int ls, const P2 (&clip)[MB], int lc, int upto,
                                      const Row& row, P2 (&poly)[CAP]) {
#pragma unroll
  for (int i = 0; i < CAP; ++i) poly[i] = (i < MA) ? subj[i < MA ? i : 0] : P2{0.0, 0.0};
#pragma unroll
  for (int i = 0; i < MA; ++i) row.put(i, subj[i]);
  int len = ls;
  P2 r = get(clip, lc - 1);
  for (int ci = 0; ci < upto; ++ci) {
    const P2 s = get(clip, ci);
    if (!(s.x - r.x == 0 && s.y - r.y == 0)) {
      len = sh_stage<CAP>(poly, len, r, s, row);
      if (len < 3) return len;
#pragma unroll
      for (int i = 0; i < CAP; ++i)
        if (i < len) poly[i] = row.get(i);
    }
    r = s;
  }
  return len;
}

// The same clip with plain runtime indexing (scratch memory), for the large vertex classes
// (CAP > 16), where most of a clip row would sit past its LDS slots anyway and the unrolled
// loops would run to 32+ slots; faces of more than 8 vertices are rare.  Same operations in
// the same order as sh_stage.
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

// sutherland_hodgman.py:84-148 polygon_polygon_clip_area.  `row` is the lane's clip row of
// clip_row_cap<MA, MB>() slots (CT_CLIP_ROWS); the indexed path for the large classes does
// not use it.
template <int MA, int MB>
constexpr int clip_row_cap() {
  return MA + MB > 16 ? 1 : MA + MB;
}

template <int MA, int MB, class Row>
__device__ __forceinline__ double clip_area(const P2 (&subj)[MA], int ls, const P2 (&clip)[MB], int lc,
                                            const Row& row) {
  constexpr int CAP = MA + MB;
  if constexpr (CAP > 16) {
    return clip_area_indexed<MA, MB>(subj, ls, clip, lc);
  } else {
    static_assert(Row::kLds + Row::kExtra == CAP, "clip row of the wrong capacity");
    P2 poly[CAP];
    const int len = sh_run<MA, MB, CAP>(subj, ls, clip, lc, lc, row, poly);
    return (len < 3) ? 0.0 : polygon_area<CAP>(poly, len);
  }
}

}  // namespace
}  // namespace ct
