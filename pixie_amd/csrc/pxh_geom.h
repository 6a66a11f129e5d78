// pxh_geom.h -- per-element geometry of pixie's fill pipeline, shared by the
// host path and the HIP kernels so both produce identical float32 results.
//
//   flatten_cubic / flatten_quad   paths.nim:677-762 (adaptive subdivision)
//   quantize_y / make_segment      paths.nim:1059-1090 (shapesToSegments)
//   init_entry / solve_x / solve_y paths.nim:1127-1144
//   seg_requires_aa                paths.nim:1149-1160
//   clip_to_line                   bumpy intersects(Segment, Line) as used by
//                                  partitionSegments, paths.nim:1236-1248
//   intersects_inside              internal.nim:36-49
//   partition_range                paths.nim:1201-1213
#pragma once

#include "pxh_common.h"

#if !(defined(__HIPCC__) || defined(__HIP__))
#include <float.h>
#include <math.h>
#endif

namespace pxh {

struct V2 {
  float x, y;
};
PXH_HD V2 v2(float x, float y) {
  V2 r;
  r.x = x;
  r.y = y;
  return r;
}
PXH_HD V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
PXH_HD V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
PXH_HD V2 operator*(V2 a, float s) { return v2(a.x * s, a.y * s); }
PXH_HD V2 operator/(V2 a, float s) { return v2(a.x / s, a.y / s); }
PXH_HD bool veq(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }
PXH_HD float lensq(V2 a) { return a.x * a.x + a.y * a.y; }
PXH_HD float len(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
PXH_HD V2 normalize(V2 a) { return a / len(a); }

// Mat3 * Vec2 in vmath's column-major layout (m[3*col + row]).
PXH_HD V2 mat_apply(const float* m, V2 v) {
  return v2(m[0] * v.x + m[3] * v.y + m[6], m[1] * v.x + m[4] * v.y + m[7]);
}

constexpr float kFltEpsilon = 1.1920928955078125e-07f;  // epsilon(float32)
// strokeShapes lays its dash pattern along each segment until the segment's
// length is used up (paths.nim:2046-2059); a length that never runs out (NaN)
// or a pattern far finer than the segment would loop without bound, so both
// forms stop after this many passes over the pattern and report a capacity
// error for the call.
constexpr int kMaxDashPasses = 1 << 16;

PXH_HD V2 cubic_at(V2 at, V2 c1, V2 c2, V2 to, float t) {
  float t2 = t * t, t3 = t2 * t;
  V2 r = at * (-t3 + 3.0f * t2 - 3.0f * t + 1.0f);
  r = r + c1 * (3.0f * t3 - 6.0f * t2 + 3.0f * t);
  r = r + c2 * (-3.0f * t3 + 3.0f * t2);
  r = r + to * t3;
  return r;
}
PXH_HD V2 cubic_deriv(V2 at, V2 c1, V2 c2, V2 to, float t) {
  float t2 = t * t;
  V2 r = at * (-3.0f * t2 + 6.0f * t - 3.0f);
  r = r + c1 * (9.0f * t2 - 12.0f * t + 3.0f);
  r = r + c2 * (-9.0f * t2 + 6.0f * t);
  r = r + to * (3.0f * t2);
  return r;
}
PXH_HD V2 quad_at(V2 at, V2 c, V2 to, float t) {
  float t2 = t * t;
  V2 r = at * (t2 - 2.0f * t + 1.0f);
  r = r + c * (-2.0f * t2 + 2.0f * t);
  r = r + to * t2;
  return r;
}

// Adaptive cubic flattening.  Calls emit(prev, next) for every accepted
// piece in order; returns kErrDiscretize when the step underflows.
template <class Emit>
PXH_HD int32_t flatten_cubic(V2 at, V2 c1, V2 c2, V2 to, float errSq, Emit& emit) {
  float t = 0.0f, step = 1.0f;
  V2 prev = at;
  V2 next = cubic_at(at, c1, c2, to, t + step);
  V2 halfway = cubic_at(at, c1, c2, to, t + step / 2.0f);
  for (;;) {
    if (step <= kFltEpsilon) return kErrDiscretize;
    V2 midpoint = (prev + next) / 2.0f;
    V2 lineTangent = midpoint - prev;
    V2 curveTangent = cubic_deriv(at, c1, c2, to, t + step / 2.0f);
    V2 curveTangentScaled = normalize(curveTangent) * len(lineTangent);
    float error = lensq(midpoint - halfway);
    float errorTangent = lensq(lineTangent - curveTangentScaled);
    if (error + errorTangent > errSq) {
      next = halfway;
      halfway = cubic_at(at, c1, c2, to, t + step / 4.0f);
      step /= 2.0f;
    } else {
      emit(prev, next);
      t += step;
      if (t == 1.0f) break;
      prev = next;
      step = nim_minf(step * 2.0f, 1.0f - t);
      next = cubic_at(at, c1, c2, to, t + step);
      halfway = cubic_at(at, c1, c2, to, t + step / 2.0f);
    }
  }
  return kOk;
}

template <class Emit>
PXH_HD int32_t flatten_quad(V2 at, V2 c, V2 to, float errSq, Emit& emit) {
  float t = 0.0f, step = 1.0f;
  V2 prev = at;
  V2 next = quad_at(at, c, to, t + step);
  V2 halfway = quad_at(at, c, to, t + step / 2.0f);
  bool halfStepping = false;
  for (;;) {
    if (step <= kFltEpsilon) return kErrDiscretize;
    V2 midpoint = (prev + next) / 2.0f;
    float error = lensq(midpoint - halfway);
    if (error > errSq) {
      next = halfway;
      halfway = quad_at(at, c, to, t + step / 4.0f);
      halfStepping = true;
      step /= 2.0f;
    } else {
      emit(prev, next);
      t += step;
      if (t == 1.0f) break;
      prev = next;
      step = halfStepping ? nim_minf(step, 1.0f - t) : nim_minf(step * 2.0f, 1.0f - t);
      next = quad_at(at, c, to, t + step);
      halfway = quad_at(at, c, to, t + step / 2.0f);
    }
  }
  return kOk;
}

// vmath quantize(v, 1/256) = sign(v) * floor(abs(v) / n) * n.
PXH_HD float quantize_y(float v) {
  const float n = 1.0f / 256.0f;
  float sg = v >= 0.0f ? 1.0f : -1.0f;
  return sg * floorf(fabsf(v) / n) * n;
}

struct Seg {
  V2 at, to;
  int32_t w;
};

// One step of shapesToSegments: the segment vec1 -> vec2 (both already
// y-quantized) oriented downward with its winding; false when horizontal.
PXH_HD bool make_segment(V2 vec1, V2 vec2, Seg* out) {
  if (vec1.y == vec2.y) return false;
  if (vec1.y > vec2.y) {
    out->at = vec2;
    out->to = vec1;
    out->w = -1;
  } else {
    out->at = vec1;
    out->to = vec2;
    out->w = 1;
  }
  return true;
}

PXH_HD Entry init_entry(V2 at, V2 to) {
  Entry e;
  e.aty = at.y;
  e.toy = to.y;
  e.m = 0.0f;
  e.b = 0.0f;
  float d = at.x - to.x;
  if (d == 0.0f) {
    e.b = at.x;
  } else {
    e.m = (at.y - to.y) / d;
    e.b = at.y - e.m * at.x;
  }
  return e;
}
PXH_HD float solve_x(float m, float b, float y) { return m == 0.0f ? b : (y - b) / m; }
PXH_HD float solve_y(float m, float b, float x) { return m * x + b; }
PXH_HD float midpoint_x(V2 at, V2 to) { return (at.x + to.x) * 0.5f; }

PXH_HD bool seg_requires_aa(V2 at, V2 to) {
  return at.x != to.x || (at.x - truncf(at.x)) != 0.0f || (at.y - truncf(at.y)) != 0.0f ||
         (to.y - truncf(to.y)) != 0.0f;
}

// Intersection of the segment with the horizontal line through (0, ly) and
// (1000, ly), the point taken on the segment (u = segment parameter).
PXH_HD bool clip_to_line(V2 sat, V2 sto, float ly, V2* at) {
  V2 s1 = v2(1000.0f, 0.0f);
  V2 s2 = sto - sat;
  float denominator = (-s2.x * s1.y + s1.x * s2.y);
  float numerator = s1.x * (ly - sat.y) - s1.y * (0.0f - sat.x);
  float u = numerator / denominator;
  if (u >= 0.0f && u <= 1.0f) {
    *at = sat + s2 * u;
    return true;
  }
  return false;
}

PXH_HD bool intersects_inside(V2 aat, V2 ato, V2 bat, V2 bto) {
  V2 s1 = ato - aat, s2 = bto - bat;
  float den = (-s2.x * s1.y + s1.x * s2.y);
  float s = (-s1.y * (aat.x - bat.x) + s1.x * (aat.y - bat.y)) / den;
  float t = (s2.x * (aat.y - bat.y) - s2.y * (aat.x - bat.x)) / den;
  return s > 0.0f && s < 1.0f && t > 0.0f && t < 1.0f;
}

// Partition range of one segment; startY/bandH/nParts as in partitionSegments.
PXH_HD void partition_range(float aty, float toy, uint32_t startY, uint32_t bandH,
                            uint32_t nParts, uint32_t* lo, uint32_t* hi) {
  uint32_t a = nim_f2u32(nim_maxf(0.0f, aty - (float)startY)) / bandH;
  uint32_t b = nim_f2u32(nim_maxf(0.0f, toy - (float)startY)) / bandH;
  *lo = a < nParts - 1 ? a : nParts - 1;
  *hi = b < nParts - 1 ? b : nParts - 1;
}

// Per-fill locals of fillShapes (paths.nim:1602-1619) from the segment
// bounds.  Returns kErrIntOverflow when pathWidth < 0 (as pixie raises for
// finite bounds beyond the int range).
struct Bounds {
  float xMin, xMax, yMin, yMax;
};
PXH_HD Bounds bounds_init() {
  Bounds b;
  b.xMin = INFINITY;
  b.xMax = -INFINITY;
  b.yMin = INFINITY;
  b.yMax = -INFINITY;
  return b;
}
PXH_HD void bounds_add(Bounds* b, const Seg& s) {
  b->xMin = nim_minf(b->xMin, nim_minf(s.at.x, s.to.x));
  b->xMax = nim_maxf(b->xMax, nim_maxf(s.at.x, s.to.x));
  b->yMin = nim_minf(b->yMin, s.at.y);
  b->yMax = nim_maxf(b->yMax, s.to.y);
}
// Merge two partial bounds; nim_minf/maxf are order-sensitive only for NaN,
// and NaN bounds are discarded as a whole below.
PXH_HD Bounds bounds_merge(Bounds a, const Bounds& b) {
  a.xMin = nim_minf(a.xMin, b.xMin);
  a.xMax = nim_maxf(a.xMax, b.xMax);
  a.yMin = nim_minf(a.yMin, b.yMin);
  a.yMax = nim_maxf(a.yMax, b.yMax);
  return a;
}

PXH_HD int32_t fill_locals(Bounds bd, bool anyNaN, int32_t W, int32_t H, FillInfo* f) {
  f->canvasW = W;
  f->canvasH = H;
  if (bd.xMin == INFINITY && bd.xMax == -INFINITY) {
    // No segment (an empty or all-horizontal path): computeBounds keeps its
    // infinite initial extremes, whose int conversions in fillShapes are
    // undefined.  The reference's own tests fill such paths and expect
    // nothing drawn and no error (a horizontal <line> with the default
    // fill in test_svg's dashes.svg, a lone moveto in its tiger, a span of
    // one space in test_fonts' spans6): nothing is drawn here either.
    f->startX = f->startY = f->pathWidth = f->pathHeight = 0;
    f->xlo = f->xhi = 0;
    return kOk;
  }
  // computeBounds 1098-1117 then snapToPixels common.nim:92-101
  float rx = 0, ry = 0, rw = 0, rh = 0;
  if (!(anyNaN || bd.xMin != bd.xMin || bd.xMax != bd.xMax || bd.yMin != bd.yMin ||
        bd.yMax != bd.yMax)) {
    rx = bd.xMin;
    ry = bd.yMin;
    rw = bd.xMax - bd.xMin;
    rh = bd.yMax - bd.yMin;
  }
  float sx = floorf(rx), sw = ceilf(rx + rw) - sx;
  float sy = floorf(ry), sh = ceilf(ry + rh) - sy;
  int64_t startX = nim_f2i(sx);
  if (startX < 0) startX = 0;
  int64_t startY = nim_f2i(sy);
  if (startY < 0) startY = 0;
  int64_t pathWidth = 0;
  if (startX < W) {
    int64_t bw = nim_f2i(sw);
    pathWidth = bw < (int64_t)W - startX ? bw : (int64_t)W - startX;
  }
  int64_t ph = nim_f2i(sy + sh);
  int64_t pathHeight = (int64_t)H < ph ? (int64_t)H : ph;
  f->startX = (int32_t)startX;
  f->startY = (int32_t)(startY > (int64_t)H ? (int64_t)H : startY);
  f->pathWidth = pathWidth < 0 ? -1 : (int32_t)pathWidth;
  f->pathHeight = (int32_t)(pathHeight < -1 ? -1 : pathHeight);
  int64_t xlo = startX - kWriteMargin, xhi = startX + (pathWidth > 0 ? pathWidth : 0) + kWriteMargin;
  f->xlo = (int32_t)(xlo < 0 ? 0 : xlo);
  f->xhi = (int32_t)(xhi > W ? W : xhi);
  if (pathWidth < 0) return kErrIntOverflow;
  return kOk;
}

// Band layout of partitionSegments for a fill with nSeg segments.
PXH_HD void band_layout(FillInfo* f, int64_t nSeg) {
  int64_t height = (int64_t)f->pathHeight - f->startY;
  int64_t hd4 = height / 4;
  uint32_t maxP = (uint32_t)(hd4 > 1 ? hd4 : 1);
  int64_t s2 = nSeg / 2;
  uint32_t sP = (uint32_t)(s2 > 1 ? s2 : 1);
  uint32_t nP = maxP < sP ? maxP : sP;
  f->nParts = (int32_t)nP;
  f->bandH = (int32_t)((uint32_t)height / nP);
}

}  // namespace pxh
