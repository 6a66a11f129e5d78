// pxh_scene.hip -- the batched form: device geometry, band building and the
// two raster phases of a recorded scene (see pxh_scene.h).
//
// Every stage is exact with respect to the per-call form: flattening,
// stroking, segments, bounds and bands use the same pxh_geom.h functions in
// the same float32 order, and the row program is the same device code.
#include <hip/hip_runtime.h>

#include "pxh_raster.h"
#include "pxh_scene.h"

namespace pxh {

constexpr int kBlock = 256;
constexpr int kBigGeomBlock = 1024;  // geometry workgroup of a call of more than kBlock commands
constexpr int kStagePts = 4096;  // LDS points of the geometry stage (32 KB)
constexpr int kBigSegGrid = 1024;  // workgroups of the big-fill segment kernels (grid-stride)

// A value that is the same in every lane, moved to a scalar register.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
constexpr int kWinA = 512;    // phase-A column window
constexpr int kWinB = 512;    // phase-B column window (2 chunks of 256 px)
constexpr int kChunksB = kWinB / 256;

// ---------------------------------------------------------------------------
// block helpers (256 threads = 4 waves)
// ---------------------------------------------------------------------------
template <int NB>
__device__ __forceinline__ int block_excl_scan_n(int v, int* total, int* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int wt;
  int x = wave_excl_scan(v, &wt);
  __syncthreads();
  if (lane == 0) sh[w] = wt;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NB / 64; i++) {
    int t = sh[i];
    if (i < w) off += t;
    tot += t;
  }
  *total = tot;
  return off + x;
}

__device__ __forceinline__ int block_excl_scan(int v, int* total, int* sh) {
  return block_excl_scan_n<kBlock>(v, total, sh);
}

// Single-workgroup scans over every band / call / bin use the largest
// workgroup so that they take few serial steps.
constexpr int kScanBlock = 1024;

__device__ int block_sum(int v, int* sh) {
  int t;
  block_excl_scan(v, &t, sh);
  return t;
}

// ---------------------------------------------------------------------------
// commands (SceneCmd) -> points, commandsToShapes (paths.nim:654-1057)
// ---------------------------------------------------------------------------
__device__ __forceinline__ V2 cmd_end(const SceneCmd& c) {
  switch ((int)c.kind) {
    case 1: case 2: return v2(c.p[0], c.p[1]);  // Move, Line
    case 5: return v2(c.p[4], c.p[5]);          // Cubic
    case 7: return v2(c.p[2], c.p[3]);          // Quad
    default: return v2(c.sx, c.sy);             // Close
  }
}

struct CountEmit {
  int n = 0;
  __device__ void operator()(V2 a, V2 b) {
    V2 d = a - b;
    if (d.x != 0.0f || d.y != 0.0f) n++;  // addSegment skips zero-length pieces
  }
};

// Counts like CountEmit and keeps the first `cap` points in an LDS slice,
// transformed like StoreEmit, so that the emit stage can copy them instead of
// flattening the command again.
struct StageEmit {
  float2* buf;
  int cap;
  const float* mat;
  bool xf;
  bool writer = true;  // lanes of a flattening group count alike; one writes
  int n = 0;
  __device__ void operator()(V2 a, V2 b) {
    V2 d = a - b;
    if (d.x == 0.0f && d.y == 0.0f) return;  // addSegment skips zero-length pieces
    if (writer && n < cap) {
      V2 p = xf ? mat_apply(mat, b) : b;
      buf[n] = make_float2(p.x, p.y);
    }
    n++;
  }
};

// flatten_cubic (paths.nim:677-721) run by a group of kFlatGroup lanes of
// one wave (lanes gbase .. gbase + kFlatGroup - 1, u = lane - gbase).  Each
// round lane u tries the step halved u times; the first candidate in
// halving order that is accepted, or that reaches epsilon, decides the round,
// which is the decision the sequential loop reaches after u halvings (each
// halving and its reuse of the halfway point are exact, so the candidates
// evaluate the same points).  Every lane of the group calls emit alike.
constexpr int kFlatGroup = 8;
template <class Emit>
__device__ __forceinline__ int32_t flatten_cubic_group(V2 at, V2 c1, V2 c2, V2 to, float errSq, Emit& emit, int u,
                                       int gbase) {
  float t = 0.0f, step = 1.0f;
  V2 prev = at;
  for (;;) {
    float s = step;
    for (int q = 0; q < u; q++) s /= 2.0f;
    const bool eps = s <= kFltEpsilon;
    bool ok = false;
    V2 next = v2(0.0f, 0.0f);
    if (!eps) {
      next = cubic_at(at, c1, c2, to, t + s);
      const V2 halfway = cubic_at(at, c1, c2, to, t + s / 2.0f);
      const V2 midpoint = (prev + next) / 2.0f;
      const V2 lineTangent = midpoint - prev;
      const V2 curveTangent = cubic_deriv(at, c1, c2, to, t + s / 2.0f);
      const V2 curveTangentScaled = normalize(curveTangent) * len(lineTangent);
      const float error = lensq(midpoint - halfway);
      const float errorTangent = lensq(lineTangent - curveTangentScaled);
      ok = !(error + errorTangent > errSq);
    }
    const uint64_t bm = (__ballot(eps || ok) >> gbase) & ((1ull << kFlatGroup) - 1ull);
    if (bm == 0) {  // every candidate was refined: go on from step / 2^kFlatGroup
      for (int q = 0; q < kFlatGroup; q++) step /= 2.0f;
      continue;
    }
    const int src = gbase + __builtin_ctzll(bm);
    if (__shfl(eps ? 1 : 0, src, 64)) return kErrDiscretize;
    const float sw = __shfl(s, src, 64);
    const V2 nw = v2(__shfl(next.x, src, 64), __shfl(next.y, src, 64));
    emit(prev, nw);
    t += sw;
    if (t == 1.0f) break;
    prev = nw;
    step = nim_minf(sw * 2.0f, 1.0f - t);
  }
  return kOk;
}

struct StoreEmit {
  float2* out;
  int32_t* outPoly;
  int32_t poly;
  const float* mat;
  bool xf;
  int n = 0;
  __device__ void put(V2 p) {
    if (xf) p = mat_apply(mat, p);
    out[n] = make_float2(p.x, p.y);
    outPoly[n] = poly;
    n++;
  }
  __device__ void operator()(V2 a, V2 b) {
    V2 d = a - b;
    if (d.x != 0.0f || d.y != 0.0f) put(b);
  }
};

// Points command i emits (its non-degenerate pieces' end points).
template <class Emit>
__device__ __forceinline__ int32_t cmd_emit(const SceneCmd* cmds, int i, float errSq, Emit& e) {
  const SceneCmd& c = cmds[i];
  int k = (int)c.kind;
  if (k == 1) return kOk;  // Move
  V2 at = cmd_end(cmds[i - 1]);  // a subpath always starts with a Move
  switch (k) {
    case 2: e(at, v2(c.p[0], c.p[1])); return kOk;
    case 5: return flatten_cubic(at, v2(c.p[0], c.p[1]), v2(c.p[2], c.p[3]), v2(c.p[4], c.p[5]), errSq, e);
    case 7: return flatten_quad(at, v2(c.p[0], c.p[1]), v2(c.p[2], c.p[3]), errSq, e);
    case 0: {  // Close: addSegment(at, start) when at != start
      V2 st = v2(c.sx, c.sy);
      if (!veq(at, st)) e(at, st);
      return kOk;
    }
    default: return kErrArg;
  }
}

// ---------------------------------------------------------------------------
// stroke polygons (strokeShapes, paths.nim:1922-2082) -> segments
// ---------------------------------------------------------------------------
struct SegSink {
  // counting when out == nullptr
  float4* out = nullptr;
  int8_t* outW = nullptr;
  int n = 0;
  int err = 0;  // a dash pattern ran past kMaxDashPasses
  Bounds bd;
  __device__ void seg(V2 v1, V2 v2_) {
    Seg s;
    if (!make_segment(v1, v2_, &s)) return;
    if (out) {
      out[n] = make_float4(s.at.x, s.at.y, s.to.x, s.to.y);
      outW[n] = (int8_t)s.w;
      bounds_add(&bd, s);
    }
    n++;
  }
  // shapesToSegments (paths.nim:1069-1090) of one polygon q[0..m)
  __device__ void poly(const V2* q, int m, const SceneItem& I) {
    if (m <= 0) return;
    V2 last = I.identity ? q[m - 1] : mat_apply(I.mat, q[m - 1]);
    V2 vec1 = v2(last.x, quantize_y(last.y));
    for (int i = 0; i < m; i++) {
      V2 p = I.identity ? q[i] : mat_apply(I.mat, q[i]);
      V2 vec2 = v2(p.x, quantize_y(p.y));
      if (i == 0 && veq(vec1, vec2)) continue;
      seg(vec1, vec2);
      vec1 = vec2;
    }
  }
};

// Streams the polygon of makeCircle (paths.nim:1938-1941): the ellipse
// path's first shape.  Its last point equals its first bit for bit (the last
// cubic ends on the Move point), so the closing candidate is always skipped.
struct CircleEmit {
  SegSink* sink;
  const SceneItem* I;
  V2 prevq;
  __device__ void add(V2 p) {
    V2 t = I->identity ? p : mat_apply(I->mat, p);
    V2 q = v2(t.x, quantize_y(t.y));
    sink->seg(prevq, q);
    prevq = q;
  }
  __device__ void operator()(V2 a, V2 b) {
    V2 d = a - b;
    if (d.x != 0.0f || d.y != 0.0f) add(b);
  }
};

__device__ __forceinline__ void circle_segments(V2 at, const SceneItem& I, SegSink& sink) {
  const double K = 4.0 * (-1.0 + 1.4142135623730951) / 3.0;  // splineCircleK (float64)
  float cx = at.x, cy = at.y, rx = I.halfStroke, ry = I.halfStroke;
  double mx = K * (double)rx, my = K * (double)ry;
  V2 s = v2(cx + rx, cy);
  V2 st = I.identity ? s : mat_apply(I.mat, s);
  CircleEmit ce{&sink, &I, v2(st.x, quantize_y(st.y))};
  V2 a = s;
  V2 c1 = v2(cx + rx, (float)((double)cy + my)), c2 = v2((float)((double)cx + mx), cy + ry), to = v2(cx, cy + ry);
  flatten_cubic(a, c1, c2, to, I.errSq, ce);
  a = to;
  c1 = v2((float)((double)cx - mx), cy + ry); c2 = v2(cx - rx, (float)((double)cy + my)); to = v2(cx - rx, cy);
  flatten_cubic(a, c1, c2, to, I.errSq, ce);
  a = to;
  c1 = v2(cx - rx, (float)((double)cy - my)); c2 = v2((float)((double)cx - mx), cy - ry); to = v2(cx, cy - ry);
  flatten_cubic(a, c1, c2, to, I.errSq, ce);
  a = to;
  c1 = v2((float)((double)cx + mx), cy - ry); c2 = v2(cx + rx, (float)((double)cy - my)); to = v2(cx + rx, cy);
  flatten_cubic(a, c1, c2, to, I.errSq, ce);
}

__device__ __forceinline__ void rect_segments(V2 at, V2 to, const SceneItem& I, SegSink& sink) {  // makeRect
  float hs = I.halfStroke;
  V2 tangent = normalize(to - at);
  V2 normal = v2(tangent.y, tangent.x);
  V2 q[5];
  q[0] = v2(at.x + normal.x * hs, at.y - normal.y * hs);
  q[1] = v2(to.x + normal.x * hs, to.y - normal.y * hs);
  q[2] = v2(to.x - normal.x * hs, to.y + normal.y * hs);
  q[3] = v2(at.x - normal.x * hs, at.y + normal.y * hs);
  q[4] = q[0];
  sink.poly(q, 5, I);
}

__device__ __forceinline__ float fix_angle_d(float a) {  // vmath fixAngle
  const float PI = 3.14159265358979323846f;
  while (a > PI) a -= PI * 2.0f;
  while (a < -PI) a += PI * 2.0f;
  return a;
}

// A round join's circle is left to the caller (*circle = true): it is always
// the last polygon a stroke point emits, so stroke_point draws every circle
// from one place.
__device__ __forceinline__ void join_segments(V2 prevPos, V2 pos, V2 nextPos, const SceneItem& I,
                                              SegSink& sink, bool* circle) {
  const float PI = 3.14159265358979323846f;
  float minArea = I.minArea, hs = I.halfStroke;
  if (I.join == 1) {  // RoundJoin
    float area = PI * hs * hs;
    if (area > minArea) *circle = true;
    return;
  }
  V2 dn = nextPos - pos, dp = prevPos - pos;
  float angle = fix_angle_d(atan2f(dn.y, dn.x) - atan2f(dp.y, dp.x));
  if (fabsf(fabsf(angle) - PI) > 0.00031415926535897936f) {
    V2 a = normalize(pos - prevPos) * hs;
    V2 b = normalize(pos - nextPos) * hs;
    if (angle >= 0.0f) {
      a = v2(-a.y, a.x);
      b = v2(b.y, -b.x);
    } else {
      a = v2(a.y, -a.x);
      b = v2(-b.y, b.x);
    }
    int j = I.join;
    if (j == 0 && fabsf(angle) < I.miterAngleLimit) j = 2;
    if (j == 0) {
      V2 la0 = prevPos + a, la1 = pos + a, lb0 = nextPos + b, lb1 = pos + b;
      V2 s1 = la1 - la0, s2 = lb1 - lb0;
      float den = (-s2.x * s1.y + s1.x * s2.y);
      if (den != 0.0f) {
        float t = (s2.x * (la0.y - lb0.y) - s2.y * (la0.x - lb0.x)) / den;
        V2 at = la0 + s1 * t;
        float bl = lensq(at - pos);
        float areaSq = 0.25f * (lensq(a) * bl + lensq(b) * bl);
        if (areaSq > minArea * minArea) {
          V2 q[5] = {pos + a, at, pos + b, pos, pos + a};
          sink.poly(q, 5, I);
        }
      }
    } else if (j == 2) {
      float areaSq = 0.25f * lensq(a) * lensq(b);
      if (areaSq > minArea * minArea) {
        V2 q[4] = {a + pos, b + pos, pos, a + pos};
        sink.poly(q, 4, I);
      }
    }
  }
}

// Stroke polygons attributed to point k of a shape with n points, in
// strokeShapes order.
// Segments of point k of a stroked polyline P[0..n): its cap or its
// segment rectangle and join.  A round cap or round join comes last, so the
// one circle a point may need is drawn at the end, from a single call site
// (which keeps I and sink in registers once inlined).
__device__ __forceinline__ void stroke_point(const float2* P, int k, int n, const SceneItem& I,
                                             const float* dash, SegSink& sink) {
  auto pt = [&](int i) { return v2(P[i].x, P[i].y); };
  bool closed = veq(pt(0), pt(n - 1));
  bool circle = false;
  V2 center = v2(0.0f, 0.0f);
  if (k == 0) {
    if (!closed) {
      if (I.cap == 1) {
        circle = true;
        center = pt(0);
      } else if (I.cap == 2) {
        V2 tangent = normalize(pt(1) - pt(0));
        rect_segments(pt(0) - tangent * I.halfStroke, pt(0), I, sink);
      }
    }
  } else {
    if (I.dashCount > 0) {
      // the dash pattern restarts on every segment (paths.nim:2046-2059)
      const V2 prevPos = pt(k - 1), pos = pt(k);
      float distance = len(prevPos - pos);
      const V2 dir = normalize(pos - prevPos);
      V2 currPos = prevPos;
      bool done = false;
      for (int pass = 0; !done; pass++) {
        if (pass == kMaxDashPasses) {
          sink.err = 1;
          break;
        }
        for (int j = 0; j < I.dashCount; j++) {
          const float d = dash[j];
          if ((j & 1) == 0) rect_segments(currPos, currPos + dir * nim_minf(distance, d), I, sink);
          currPos = currPos + dir * d;
          distance -= d;
          if (distance <= 0.0f) {
            done = true;
            break;
          }
        }
      }
    } else {
      rect_segments(pt(k - 1), pt(k), I, sink);
    }
    if (k < n - 1) {
      join_segments(pt(k - 1), pt(k), pt(k + 1), I, sink, &circle);
      center = pt(k);
    }
    if (k == n - 1) {
      if (closed) {
        if (n >= 2) {
          join_segments(pt(n - 2), pt(n - 1), pt(1), I, sink, &circle);
          center = pt(n - 1);
        }
      } else {
        if (I.cap == 1) {
          circle = true;
          center = pt(n - 1);
        } else if (I.cap == 2) {
          V2 tangent = normalize(pt(n - 1) - pt(n - 2));
          rect_segments(pt(n - 1) + tangent * I.halfStroke, pt(n - 1), I, sink);
        }
      }
    }
  }
  if (circle) circle_segments(center, I, sink);
}

// ---------------------------------------------------------------------------
// geometry: one workgroup per call
// ---------------------------------------------------------------------------
// Stages 6-7 of the geometry kernel: the call's bounds (merged from every
// thread's `bd`) give its fill locals and band layout (fill_locals,
// band_layout), then every segment gets its band range and AA bit, and every
// 64-segment chunk the band range of its segments.
template <int NB>
__device__ __forceinline__ void geom_finish(const ScenePools& P, const SceneItem& I, int it, Bounds bd, int nseg, int sBase,
                            Bounds* sBd, FillInfo& sF) {
  const int tid = threadIdx.x;
  // 6. bounds -> fill locals, band layout
  for (int o = 32; o > 0; o >>= 1) {
    Bounds ob;
    ob.xMin = __shfl_xor(bd.xMin, o, 64);
    ob.xMax = __shfl_xor(bd.xMax, o, 64);
    ob.yMin = __shfl_xor(bd.yMin, o, 64);
    ob.yMax = __shfl_xor(bd.yMax, o, 64);
    bd = bounds_merge(bd, ob);
  }
  if ((tid & 63) == 0) sBd[tid >> 6] = bd;
  __syncthreads();
  if (tid == 0) {
    Bounds b = sBd[0];
    for (int w = 1; w < NB / 64; w++) b = bounds_merge(b, sBd[w]);
    FillInfo f = {};
    f.rgbx = I.rgbx;
    f.mode = I.mode;
    f.rule = I.stroke ? kNonZero : I.rule;
    int32_t r = fill_locals(b, false, P.W, P.H, &f);
    f.status = r;
    if (r == kOk && f.pathWidth > 0) band_layout(&f, nseg);
    P.fills[it] = f;
    sF = f;
  }
  __syncthreads();
  // 7. band range of every segment (partitionSegments 1201-1213) and its
  // requiresAntiAliasing bit, and the band range of each 64-segment chunk
  const FillInfo fl = sF;
  if (fl.status == kOk && fl.pathWidth > 0) {
    // four rounds at a time, their loads first
    for (int j0 = 0; j0 < nseg; j0 += 4 * NB) {
      float4 sg[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = j0 + u * NB + tid;
        sg[u] = j < nseg ? P.segs[sBase + j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = j0 + u * NB + tid;
        int lo = 0x7fffffff, hi = -1;
        if (j < nseg) {
          uint32_t l = 0, h = 0;
          if (fl.nParts > 1)
            partition_range(sg[u].y, sg[u].w, (uint32_t)fl.startY, (uint32_t)fl.bandH, (uint32_t)fl.nParts,
                            &l, &h);
          const bool aa = seg_requires_aa(v2(sg[u].x, sg[u].y), v2(sg[u].z, sg[u].w));
          P.segBand[sBase + j] = make_int2((int)l, (int)h | (aa ? kSegAA : 0));
          lo = (int)l;
          hi = (int)h;
        }
        for (int o = 32; o > 0; o >>= 1) {
          lo = min(lo, __shfl_xor(lo, o, 64));
          hi = max(hi, __shfl_xor(hi, o, 64));
        }
        if ((tid & 63) == 0 && j < nseg) P.segChunk[(sBase + j) >> 6] = make_int2(lo, hi);
      }
    }
  }
}

// Thread per command of the calls of more than kBlock commands (the first
// nBig calls of itemOrder; bigCmdStart holds their command prefix): counts
// its points (emit = 0; a negative count is an error code), or writes them
// at the offsets the front half computed (emit = 1).
__global__ void __launch_bounds__(256) scene_big_cmd_kernel(ScenePools P, int emit) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  if (gid >= P.bigCmdStart[P.nBig]) return;
  int lo = 0, hi = P.nBig;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (P.bigCmdStart[mid] <= gid) lo = mid;
    else hi = mid;
  }
  const int it = P.itemOrder[lo];
  const SceneItem& I = P.items[it];
  const int i = I.cmdBase + (gid - P.bigCmdStart[lo]);
  const bool xf = !I.stroke && !I.identity;
  if (!emit) {
    CountEmit ce;
    const int32_t r = cmd_emit(P.cmds, i, I.errSq, ce);
    P.cmdCount[i] = r != kOk ? -r : ce.n;
    return;
  }
  if (P.geo[it].status != kOk || P.cmdCount[i] <= 0) return;
  const int o = P.geo[it].ptBase + P.cmdOff[i];
  StoreEmit se{P.pts + o, P.ptPoly + o, P.cmds[i].pad[0], I.mat, xf};
  cmd_emit(P.cmds, i, I.errSq, se);
}

// ---- the segment stage of big fills, spread over the device ----
// Stage 5 of a fill makes a segment from each point and its predecessor
// (the first point of a subpath pairs with its last), skipping horizontal
// ones; segments keep point order.  For a big fill the points are cut into
// chunks of kBigChunk (256 threads x 16 consecutive points): the chunks
// count their segments and bounds, one workgroup per call scans the counts,
// merges the bounds and lays out the call's bands (stage 6), the chunks
// write their segments at the scanned offsets, and one wave per 64-segment
// group finds the band ranges (stage 7).  The result is the one-workgroup
// back half's, segment for segment.

__device__ __forceinline__ bool fill_point_seg(const ScenePools& P, const SceneItem& I, const float2* pts,
                                               const int32_t* ptPoly, int j, Seg* s) {
  const int q = I.polyBase + ptPoly[j];
  const int pb = P.polyPt[q], np = P.polyN[q];
  const float2 pa = j == pb ? pts[pb + np - 1] : pts[j - 1];
  const float2 pc = pts[j];
  const V2 v1 = v2(pa.x, quantize_y(pa.y)), v2_ = v2(pc.x, quantize_y(pc.y));
  return !(j == pb && veq(v1, v2_)) && make_segment(v1, v2_, s);
}

// The big call whose range of `start` (an exclusive prefix, nBig + 1
// entries) holds index c.
__device__ __forceinline__ int big_call_of(const int32_t* start, int nBig, int c) {
  int lo = 0, hi = nBig;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (start[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Block-wide bounds of 256 threads (valid in every thread after the call).
__device__ __forceinline__ Bounds block_bounds256(Bounds bd, Bounds* sBd) {
  for (int o = 32; o > 0; o >>= 1) {
    Bounds ob;
    ob.xMin = __shfl_xor(bd.xMin, o, 64);
    ob.xMax = __shfl_xor(bd.xMax, o, 64);
    ob.yMin = __shfl_xor(bd.yMin, o, 64);
    ob.yMax = __shfl_xor(bd.yMax, o, 64);
    bd = bounds_merge(bd, ob);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sBd[threadIdx.x >> 6] = bd;
  __syncthreads();
  Bounds b = sBd[0];
  for (int w = 1; w < 4; w++) b = bounds_merge(b, sBd[w]);
  return b;
}

// Prefixes over the big calls: subpath chunks of every call (what = 2),
// point chunks of the fills whose points were laid out (what = 0), and
// 64-segment groups of the fills that draw (what = 1).
__global__ void __launch_bounds__(kScanBlock) scene_big_plan_kernel(ScenePools P, int what) {
  __shared__ int sh[kScanBlock / 64];
  int32_t* out = what == 2 ? P.bigPolyStart : what == 1 ? P.bigGroupStart : P.bigChunkStart;
  int carry = 0;
  for (int i0 = 0; i0 < P.nBig; i0 += kScanBlock) {
    const int i = i0 + threadIdx.x;
    int n = 0;
    if (i < P.nBig) {
      const int it = P.itemOrder[i];
      if (what == 2) {
        n = (P.items[it].polyCount + kBigPolyChunk - 1) / kBigPolyChunk;
      } else {
        const ItemGeom g = P.geo[it];
        if (!P.items[it].stroke && g.status == kOk) {
          if (what == 0) n = (g.ptCount + kBigChunk - 1) / kBigChunk;
          else if (P.fills[it].pathWidth > 0) n = (g.segCount + 63) / 64;
        }
      }
    }
    int t;
    const int off = block_excl_scan_n<kScanBlock>(n, &t, sh);
    if (i < P.nBig) out[i] = carry + off;
    carry += t;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[P.nBig] = carry;
}

// ---- the point layout of big calls (geometry stages 2-3), spread over
// the device: each chunk of kBigPolyChunk subpaths (256 threads x 4)
// counts its subpaths' points and finds count errors, one workgroup per call
// scans the chunk totals and allocates the call's points, and the chunks
// place their subpaths (offsets, per-command offsets, start and closing
// points).  The layout is the one-workgroup front half's, point for point.

// Points of subpath q (stage 2): start + pieces, plus the implicit closing
// point of a fill; *err gets a command's error code.
__device__ __forceinline__ int big_poly_points(const ScenePools& P, const SceneItem& I, int q, int* err) {
  const ScenePoly sp = P.polys[q];
  int n = 0;
  for (int c = 0; c < sp.numCmds; c++) {
    const int k = P.cmdCount[sp.firstCmd + c];
    if (k < 0) *err = max(*err, -k);
    else n += k;
  }
  if (n <= 0) return 0;
  int np = n + 1;
  const SceneCmd& last = P.cmds[sp.firstCmd + sp.numCmds - 1];
  if (!I.stroke && (int)last.kind != 0 && !sp.noClose) {
    V2 d = cmd_end(last) - v2(last.sx, last.sy);
    if (d.x != 0.0f || d.y != 0.0f) np++;
  }
  return np;
}

// emit = 0: points and errors per chunk; emit = 1: subpath placement.
__global__ void __launch_bounds__(256) scene_big_front_kernel(ScenePools P, int emit) {
  __shared__ int sh[4];
  const int nChunks = P.bigPolyStart[P.nBig];
  for (int c = blockIdx.x; c < nChunks; c += gridDim.x) {
    const int bi = big_call_of(P.bigPolyStart, P.nBig, c);
    const int it = P.itemOrder[bi];
    const SceneItem& I = P.items[it];
    const int ql0 = (c - P.bigPolyStart[bi]) * kBigPolyChunk + threadIdx.x * 4;
    if (!emit) {
      int sum = 0, err = kOk;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int ql = ql0 + u;
        if (ql < I.polyCount) {
          const int np = big_poly_points(P, I, I.polyBase + ql, &err);
          P.polyN[I.polyBase + ql] = np;
          sum += np;
        }
      }
      int t;
      block_excl_scan_n<256>(sum, &t, sh);
      for (int o = 32; o > 0; o >>= 1) err = max(err, __shfl_xor(err, o, 64));
      __syncthreads();
      if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = err;
      __syncthreads();
      if (threadIdx.x == 0) P.bigPolyChunk[c] = make_int4(t, max(max(sh[0], sh[1]), max(sh[2], sh[3])), 0, 0);
      __syncthreads();
      continue;
    }
    const ItemGeom g = P.geo[it];
    if (g.status != kOk) continue;  // the same for the whole block
    int np[4], sum = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int ql = ql0 + u;
      np[u] = ql < I.polyCount ? P.polyN[I.polyBase + ql] : 0;
      sum += np[u];
    }
    int t;
    int pb = P.bigPolyChunk[c].z + block_excl_scan_n<256>(sum, &t, sh);
    const bool xf = !I.stroke && !I.identity;
    float2* pts = P.pts + g.ptBase;
    int32_t* ptPoly = P.ptPoly + g.ptBase;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int ql = ql0 + u;
      if (ql < I.polyCount) {  // stage 3
        const int q = I.polyBase + ql;
        P.polyPt[q] = pb;
        const ScenePoly sp = P.polys[q];
        int off = pb + 1;
        for (int k = 0; k < sp.numCmds; k++) {
          P.cmdOff[sp.firstCmd + k] = off;
          off += P.cmdCount[sp.firstCmd + k];
        }
        if (np[u] > 0) {
          const SceneCmd& mv = P.cmds[sp.firstCmd];
          V2 s = v2(mv.sx, mv.sy);
          V2 w = xf ? mat_apply(I.mat, s) : s;
          pts[pb] = make_float2(w.x, w.y);
          ptPoly[pb] = ql;
          if (off < pb + np[u]) {  // implicit closing point of a fill
            pts[off] = make_float2(w.x, w.y);
            ptPoly[off] = ql;
          }
        }
        pb += np[u];
      }
    }
    __syncthreads();
  }
}

// One workgroup per big call: its point total and errors from the chunks,
// the chunks' point offsets, and the call's point allocation.
__global__ void __launch_bounds__(256) scene_big_front_alloc_kernel(ScenePools P) {
  __shared__ int sh[4];
  const int bi = blockIdx.x;
  const int it = P.itemOrder[bi];
  const SceneItem& I = P.items[it];
  const int c0 = P.bigPolyStart[bi], nc = P.bigPolyStart[bi + 1] - c0;
  int carry = 0, err = kOk;
  for (int k0 = 0; k0 < nc; k0 += 256) {
    const int k = k0 + threadIdx.x;
    int n = 0;
    if (k < nc) {
      const int4 ch = P.bigPolyChunk[c0 + k];
      n = ch.x;
      err = max(err, ch.y);
    }
    int t;
    const int off = block_excl_scan_n<256>(n, &t, sh);
    if (k < nc) P.bigPolyChunk[c0 + k].z = carry + off;
    carry += t;
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) err = max(err, __shfl_xor(err, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = err;
  __syncthreads();
  if (threadIdx.x != 0) return;
  err = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
  const int total = carry;
  int base = 0;
  if (err == kOk) {
    base = atomicAdd(&P.counters[kCntPoints], total);
    if (base + total > P.ptCap) {
      atomicOr(&P.counters[kCntError], 1);
      err = kErrCapacity;
    }
    // strokeShapes checks the dash values once it meets a shape with
    // points (paths.nim:2029-2031)
    if (I.stroke && !I.strokeNone && I.dashBad && total > 0 && err == kOk) err = kErrDash;
  }
  ItemGeom g = {};
  if (err != kOk) {
    g.status = err;
    P.fills[it].status = err;
    P.fills[it].pathWidth = 0;
  } else {
    g.ptBase = base;
    g.ptCount = total;
    g.status = kOk;
  }
  P.geo[it] = g;
}

// Segments and bounds of each chunk (emit = 0), or its segments written at
// the call's segment base plus the chunk's scanned offset (emit = 1).
__global__ void __launch_bounds__(256) scene_big_seg_kernel(ScenePools P, int emit) {
  __shared__ int sh[4];
  __shared__ Bounds sBd[4];
  const int nChunks = P.bigChunkStart[P.nBig];
  for (int c = blockIdx.x; c < nChunks; c += gridDim.x) {
    const int bi = big_call_of(P.bigChunkStart, P.nBig, c);
    const int it = P.itemOrder[bi];
    const SceneItem& I = P.items[it];
    const ItemGeom g = P.geo[it];
    if (emit && g.status != kOk) continue;  // the same for the whole block
    const float2* pts = P.pts + g.ptBase;
    const int32_t* ptPoly = P.ptPoly + g.ptBase;
    const int j0 = (c - P.bigChunkStart[bi]) * kBigChunk + threadIdx.x * 16;
    float4 sg[16];
    int8_t sw[16];
    uint32_t valid = 0;
    Bounds bd = bounds_init();
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int j = j0 + u;
      Seg s;
      if (j < g.ptCount && fill_point_seg(P, I, pts, ptPoly, j, &s)) {
        sg[u] = make_float4(s.at.x, s.at.y, s.to.x, s.to.y);
        sw[u] = (int8_t)s.w;
        bounds_add(&bd, s);
        valid |= 1u << u;
      }
    }
    const int n = __popc(valid);
    int t;
    const int off = block_excl_scan_n<256>(n, &t, sh);
    if (!emit) {
      const Bounds b = block_bounds256(bd, sBd);
      if (threadIdx.x == 0) {
        P.bigChunkCnt[c] = make_int2(t, 0);
        P.bigChunkBd[c] = make_float4(b.xMin, b.xMax, b.yMin, b.yMax);
      }
    } else {
      const int w = g.segBase + P.bigChunkCnt[c].y + off;
      int k = 0;
#pragma unroll
      for (int u = 0; u < 16; u++)
        if ((valid >> u) & 1u) {
          P.segs[w + k] = sg[u];
          P.segW[w + k] = sw[u];
          k++;
        }
    }
    __syncthreads();
  }
}

// One workgroup per big call: scans its chunk counts, merges their bounds,
// allocates the segments and lays out the fill (stage 6).
__global__ void __launch_bounds__(256) scene_big_seg_alloc_kernel(ScenePools P) {
  __shared__ int sh[4];
  __shared__ Bounds sBd[4];
  const int bi = blockIdx.x;
  const int it = P.itemOrder[bi];
  const SceneItem& I = P.items[it];
  if (I.stroke) return;
  ItemGeom g = P.geo[it];
  if (g.status != kOk) return;  // the front half wrote the error
  const int c0 = P.bigChunkStart[bi], nc = P.bigChunkStart[bi + 1] - c0;
  int carry = 0;
  Bounds bd = bounds_init();
  for (int k0 = 0; k0 < nc; k0 += 256) {
    const int k = k0 + threadIdx.x;
    int n = 0;
    if (k < nc) {
      n = P.bigChunkCnt[c0 + k].x;
      const float4 b = P.bigChunkBd[c0 + k];
      Bounds cb;
      cb.xMin = b.x;
      cb.xMax = b.y;
      cb.yMin = b.z;
      cb.yMax = b.w;
      bd = bounds_merge(bd, cb);
    }
    int t;
    const int off = block_excl_scan_n<256>(n, &t, sh);
    if (k < nc) P.bigChunkCnt[c0 + k].y = carry + off;
    carry += t;
    __syncthreads();
  }
  const Bounds b = block_bounds256(bd, sBd);
  if (threadIdx.x != 0) return;
  const int nseg = carry;
  // segments start on a 64-segment chunk of their own
  const int base = atomicAdd(&P.counters[kCntSegs], (nseg + 63) & ~63);
  FillInfo f = {};
  f.rgbx = I.rgbx;
  f.mode = I.mode;
  f.rule = I.rule;
  if (base + nseg > P.segCap) {
    atomicOr(&P.counters[kCntError], 1);
    f.status = kErrCapacity;
    f.pathWidth = 0;
    g.segBase = 0;
    g.segCount = 0;
    g.status = kErrCapacity;
  } else {
    const int32_t r = fill_locals(b, false, P.W, P.H, &f);
    f.status = r;
    if (r == kOk && f.pathWidth > 0) band_layout(&f, nseg);
    g.segBase = base;
    g.segCount = nseg;
    g.status = r;
  }
  P.fills[it] = f;
  P.geo[it] = g;
}

// Stage 7 of the big fills: one wave per 64-segment group.
__global__ void __launch_bounds__(256) scene_big_band_kernel(ScenePools P) {
  const int nGroups = P.bigGroupStart[P.nBig];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int gi = blockIdx.x * 4 + wave; gi < nGroups; gi += gridDim.x * 4) {
    const int bi = big_call_of(P.bigGroupStart, P.nBig, gi);
    const int it = P.itemOrder[bi];
    const ItemGeom g = P.geo[it];
    const FillInfo fl = P.fills[it];
    const int j = (gi - P.bigGroupStart[bi]) * 64 + lane;
    int lo = 0x7fffffff, hi = -1;
    if (j < g.segCount) {
      const float4 sg = P.segs[g.segBase + j];
      uint32_t l = 0, h = 0;
      if (fl.nParts > 1)
        partition_range(sg.y, sg.w, (uint32_t)fl.startY, (uint32_t)fl.bandH, (uint32_t)fl.nParts, &l, &h);
      const bool aa = seg_requires_aa(v2(sg.x, sg.y), v2(sg.z, sg.w));
      P.segBand[g.segBase + j] = make_int2((int)l, (int)h | (aa ? kSegAA : 0));
      lo = (int)l;
      hi = (int)h;
    }
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if (lane == 0) P.segChunk[(g.segBase + j) >> 6] = make_int2(lo, hi);
  }
}

// Geometry kernel stages by MODE: kGeomAll runs every stage for a call.
// Calls of more than kBlock commands spread their stages over the CUs:
// scene_big_cmd_kernel counts every command's points, scene_big_front_*
// lay the points out (stages 2-3), scene_big_cmd_kernel emits them, and the
// segment stages run in scene_big_seg_* / scene_big_band_kernel for fills
// and in one kGeomBack workgroup per call for strokes (stages 5-7).
constexpr int kGeomAll = 0, kGeomBack = 2;

template <int NB, int MODE>
__global__ void __launch_bounds__(NB) scene_geom_kernel(ScenePools P, int orderBase) {
  __shared__ int sh[NB / 64];
  __shared__ int sBase;
  __shared__ Bounds sBd[NB / 64];
  __shared__ FillInfo sF;
  __shared__ int sErr;
  __shared__ float2 sPts[kStagePts];
  __shared__ int sPre[NB + 1];
  const int it = P.itemOrder[orderBase + blockIdx.x];
  const int tid = threadIdx.x;
  const SceneItem& I = P.items[it];
  const bool xf = !I.stroke && !I.identity;  // fills transform points here
  if (tid == 0) sErr = kOk;
  __syncthreads();

  int total = 0, ptBase = 0;
  if (MODE == kGeomBack) {
    const ItemGeom gf = P.geo[it];
    if (gf.status != kOk) return;  // the front wrote the error
    if (!I.stroke) return;         // big fills: scene_big_seg_* kernels
    total = gf.ptCount;
    ptBase = gf.ptBase;
    goto segments;
  }
  {
  // 1. points per command.  A call of at most NB commands keeps each
  // command's points in an LDS slice of its own (thread tid <-> command tid
  // in stages 1 and 4), so stage 4 copies them unless a slice overflowed.
  const bool stage = I.cmdCount <= NB;
  const int slice = stage && I.cmdCount > 0 ? kStagePts / I.cmdCount : 0;
  bool staged = false;
  int myN = 0;
  if (stage && I.cmdCount <= NB / kFlatGroup) {
    // few commands: kFlatGroup lanes per command, so a cubic's adaptive
    // steps are tried kFlatGroup halvings at a time
    const int c = tid / kFlatGroup, u = tid % kFlatGroup;
    if (c < I.cmdCount) {
      const int i = I.cmdBase + c;
      StageEmit se{sPts + c * slice, slice, I.mat, xf};
      se.writer = u == 0;
      int32_t r;
      const SceneCmd& cm = P.cmds[i];
      if ((int)cm.kind == 5)
        r = flatten_cubic_group(cmd_end(P.cmds[i - 1]), v2(cm.p[0], cm.p[1]), v2(cm.p[2], cm.p[3]),
                                v2(cm.p[4], cm.p[5]), I.errSq, se, u, tid & 63 & ~(kFlatGroup - 1));
      else
        r = cmd_emit(P.cmds, i, I.errSq, se);
      if (u == 0) {
        if (r != kOk) atomicMax(&sErr, r);
        P.cmdCount[i] = se.n;
        sPre[c] = se.n;
      }
    }
    __syncthreads();
    if (tid < I.cmdCount) {
      myN = sPre[tid];
      staged = myN <= slice;
    }
  } else if (MODE == kGeomAll) {
    for (int c0 = 0; c0 < I.cmdCount; c0 += NB) {
      int i = I.cmdBase + c0 + tid;
      if (c0 + tid < I.cmdCount) {
        StageEmit se{sPts + tid * slice, slice, I.mat, xf};
        int32_t r = cmd_emit(P.cmds, i, I.errSq, se);
        if (r != kOk) atomicMax(&sErr, r);
        P.cmdCount[i] = se.n;
        myN = se.n;
        staged = stage && se.n <= slice;
      }
    }
  }
  __syncthreads();
  // With one command per thread: points before each command of the call
  // (sPre[c], and the total in sPre[cmdCount]), so that subpath sums and
  // command offsets take two LDS reads instead of a loop over commands.
  if (stage) {
    int tot;
    const int pre = block_excl_scan_n<NB>(tid < I.cmdCount ? myN : 0, &tot, sh);
    if (tid < I.cmdCount) sPre[tid] = pre;
    if (tid == 0) sPre[I.cmdCount] = tot;
    __syncthreads();
  }
  // 2. points per subpath: start + pieces (+ the implicit closing point of a fill)
  for (int q0 = 0; q0 < I.polyCount; q0 += NB) {
    int q = I.polyBase + q0 + tid;
    int np = 0;
    if (q0 + tid < I.polyCount) {
      ScenePoly sp = P.polys[q];
      int n = 0;
      if (stage) {
        const int cl = sp.firstCmd - I.cmdBase;
        n = sPre[cl + sp.numCmds] - sPre[cl];
      } else {
        for (int c = 0; c < sp.numCmds; c++) n += P.cmdCount[sp.firstCmd + c];
      }
      if (n > 0) {
        np = n + 1;
        const SceneCmd& last = P.cmds[sp.firstCmd + sp.numCmds - 1];
        if (!I.stroke && (int)last.kind != 0 && !sp.noClose) {
          V2 d = cmd_end(last) - v2(last.sx, last.sy);
          if (d.x != 0.0f || d.y != 0.0f) np++;
        }
      }
    }
    int t;
    int off = block_excl_scan_n<NB>(np, &t, sh);
    if (q0 + tid < I.polyCount) {
      P.polyPt[q] = total + off;
      P.polyN[q] = np;
    }
    total += t;
  }
  if (tid == 0) {
    int base = atomicAdd(&P.counters[kCntPoints], total);
    if (base + total > P.ptCap) {
      atomicOr(&P.counters[kCntError], 1);
      sErr = kErrCapacity;
    }
    // strokeShapes checks the dash values once it meets a shape with points
    // (paths.nim:2029-2031)
    if (I.stroke && !I.strokeNone && I.dashBad && total > 0 && sErr == kOk) sErr = kErrDash;
    sBase = base;
  }
  __syncthreads();
  ptBase = sBase;
  if (sErr != kOk) {
    if (tid == 0) {
      P.geo[it].status = sErr;
      P.geo[it].segCount = 0;
      P.fills[it].status = sErr;
      P.fills[it].pathWidth = 0;
    }
    return;
  }
  float2* pts = P.pts + ptBase;
  int32_t* ptPoly = P.ptPoly + ptBase;
  int32_t* ptSeg = P.ptSeg + ptBase;  // segments of each point (stage 5)

  // 3. per-command point offsets, start and closing points (thread per subpath)
  for (int q0 = 0; q0 < I.polyCount; q0 += NB) {
    int ql = q0 + tid;
    if (ql < I.polyCount) {
      int q = I.polyBase + ql;
      ScenePoly sp = P.polys[q];
      int np = P.polyN[q], pb = P.polyPt[q];
      int off = pb + 1;
      if (stage) {  // stage 4 derives each command's offset from sPre
        const int cl = sp.firstCmd - I.cmdBase;
        off += sPre[cl + sp.numCmds] - sPre[cl];
      } else {
        for (int c = 0; c < sp.numCmds; c++) {
          P.cmdOff[sp.firstCmd + c] = off;
          off += P.cmdCount[sp.firstCmd + c];
        }
      }
      if (np > 0) {
        const SceneCmd& mv = P.cmds[sp.firstCmd];
        V2 s = v2(mv.sx, mv.sy);
        V2 w = xf ? mat_apply(I.mat, s) : s;
        pts[pb] = make_float2(w.x, w.y);
        ptPoly[pb] = ql;
        if (off < pb + np) {  // implicit closing point of a fill
          pts[off] = make_float2(w.x, w.y);
          ptPoly[off] = ql;
        }
      }
    }
  }
  __syncthreads();
  // 4. emit points (thread per command)
  for (int c0 = 0; c0 < I.cmdCount; c0 += NB) {
    int i = I.cmdBase + c0 + tid;
    if (c0 + tid < I.cmdCount && P.cmdCount[i] > 0) {
      // subpath of command i (the host stores it in pad[0] of every command)
      int ql = P.cmds[i].pad[0];
      int o;
      if (stage) {
        const int q = I.polyBase + ql;
        o = P.polyPt[q] + 1 + sPre[c0 + tid] - sPre[P.polys[q].firstCmd - I.cmdBase];
      } else {
        o = P.cmdOff[i];
      }
      if (staged) {
        const int n = P.cmdCount[i];
        const float2* src = sPts + tid * slice;
        for (int k = 0; k < n; k++) {
          pts[o + k] = src[k];
          ptPoly[o + k] = ql;
        }
      } else {
        StoreEmit se{pts + o, ptPoly + o, ql, I.mat, xf};
        cmd_emit(P.cmds, i, I.errSq, se);
      }
    }
  }
  __syncthreads();
  }
segments:
  float2* pts = P.pts + ptBase;
  int32_t* ptPoly = P.ptPoly + ptBase;
  int32_t* ptSeg = P.ptSeg + ptBase;  // segments of each point (stage 5)

  // 5. segments: count, allocate, emit (thread per point).  A fill makes at
  // most one segment per point, so it allocates that bound and emits in one
  // pass; a stroke counts first.
  int nseg = 0;
  const bool onePass = !I.stroke;
  if (onePass) {
    if (tid == 0) {
      // segments start on a 64-segment chunk of their own
      int base = atomicAdd(&P.counters[kCntSegs], (total + 63) & ~63);
      if (base + total > P.segCap) {
        atomicOr(&P.counters[kCntError], 1);
        sErr = kErrCapacity;
      }
      sBase = base;
    }
    __syncthreads();
    if (sErr == kOk) {
      // four consecutive points per thread per round, their loads first
      int segOff = 0;
      Bounds bd = bounds_init();
      for (int j0 = 0; j0 < total; j0 += 4 * NB) {
        const int jb = j0 + 4 * tid;
        Seg fs[4];
        int c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int j = jb + u;
          c[u] = 0;
          if (j < total) {
            const int q = I.polyBase + ptPoly[j];
            const int pb = P.polyPt[q], np = P.polyN[q];
            // shapesToSegments: candidate (prev point -> this point); the
            // first point pairs with the polygon's last point
            const float2 pa = j == pb ? pts[pb + np - 1] : pts[j - 1];
            const float2 pc = pts[j];
            const V2 v1 = v2(pa.x, quantize_y(pa.y)), v2_ = v2(pc.x, quantize_y(pc.y));
            if (!(j == pb && veq(v1, v2_)) && make_segment(v1, v2_, &fs[u])) c[u] = 1;
          }
        }
        int t;
        int w = segOff + block_excl_scan_n<NB>(c[0] + c[1] + c[2] + c[3], &t, sh);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (c[u]) {
            P.segs[sBase + w] = make_float4(fs[u].at.x, fs[u].at.y, fs[u].to.x, fs[u].to.y);
            P.segW[sBase + w] = (int8_t)fs[u].w;
            bounds_add(&bd, fs[u]);
            w++;
          }
        }
        segOff += t;
      }
      nseg = segOff;
      geom_finish<NB>(P, I, it, bd, nseg, sBase, sBd, sF);
    }
  } else {
    // a stroke: count the segments of every point, allocate, then emit
    for (int pass = 0; pass < 2 && sErr == kOk; pass++) {
      int segOff = 0;
      Bounds bd = bounds_init();
      for (int j0 = 0; j0 < total; j0 += NB) {
        int j = j0 + tid;
        int cnt = 0;
        if (pass == 1) {
          // the counts of pass 0 (this thread wrote them)
          if (j < total) cnt = ptSeg[j];
        } else if (j < total && !I.strokeNone) {
          int q = I.polyBase + ptPoly[j];
          int pb = P.polyPt[q], np = P.polyN[q];
          SegSink sink;
          stroke_point(pts + pb, j - pb, np, I, P.dashes + I.dashBase, sink);
          cnt = sink.n;
          if (sink.err) atomicMax(&sErr, kErrCapacity);
        }
        if (pass == 0 && j < total) ptSeg[j] = cnt;
        int t;
        int off = block_excl_scan_n<NB>(cnt, &t, sh);
        if (pass == 1 && j < total && cnt > 0) {
          SegSink out;
          out.out = P.segs + sBase + segOff + off;
          out.outW = P.segW + sBase + segOff + off;
          out.bd = bd;
          int q = I.polyBase + ptPoly[j];
          int pb = P.polyPt[q], np = P.polyN[q];
          stroke_point(pts + pb, j - pb, np, I, P.dashes + I.dashBase, out);
          bd = out.bd;
        }
        segOff += t;
      }
      if (pass == 0) {
        nseg = segOff;
        __syncthreads();
        if (tid == 0) {
          // segments start on a 64-segment chunk of their own
          int base = atomicAdd(&P.counters[kCntSegs], (nseg + 63) & ~63);
          if (base + nseg > P.segCap) {
            atomicOr(&P.counters[kCntError], 1);
            sErr = kErrCapacity;
          }
          sBase = base;
        }
        __syncthreads();
      } else {
        geom_finish<NB>(P, I, it, bd, nseg, sBase, sBd, sF);
      }
    }
  }
  if (tid == 0) {
    ItemGeom g = {};
    g.ptBase = ptBase;
    g.ptCount = total;
    g.segBase = sErr == kOk ? sBase : 0;
    g.segCount = sErr == kOk ? nseg : 0;
    g.status = sErr != kOk ? sErr : P.fills[it].status;
    if (sErr != kOk) {
      P.fills[it].status = sErr;
      P.fills[it].pathWidth = 0;
    }
    P.geo[it] = g;
  }
}

// ---------------------------------------------------------------------------
// planning (one block): bands, phase-A tasks, band -> call map
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kScanBlock) scene_plan_kernel(ScenePools P) {
  __shared__ int sh[kScanBlock / 64];
  const int tid = threadIdx.x;
  int bandTot = 0, taskTot = 0;
  for (int i0 = 0; i0 < P.nItems; i0 += kScanBlock) {
    int it = i0 + tid;
    int nb = 0, nt = 0;
    ItemGeom g;
    FillInfo f;
    int4 rows = make_int4(0, 0, 0, 0);
    if (it < P.nItems) {
      g = P.geo[it];
      f = P.fills[it];
      g.rowStart = 0;
      g.rowCount = 0;
      g.nWin = 0;
      if (g.status == kOk && f.pathWidth > 0) {
        nb = f.nParts;
        int rs = f.startY, re = f.pathHeight;
        if (re > P.H) re = P.H;
        g.rowStart = rs;
        g.rowCount = re > rs ? re - rs : 0;
        int lo = f.mode == kMask ? 0 : (f.xlo & ~3), hi = f.mode == kMask ? P.W : f.xhi;
        g.winBase = lo;
        g.nWin = hi > lo ? (hi - lo + kWinA - 1) / kWinA : 0;
        nt = g.rowCount * g.nWin;
        // a mask call touches every row (it clears the rows outside it)
        rows = f.mode == kMask ? make_int4(0, P.H, 0, P.W)
                               : make_int4(g.rowStart, g.rowStart + g.rowCount, f.xlo, f.xhi);
      }
      P.itemRows[it] = rows;
    }
    int tb, tt;
    int bOff = block_excl_scan_n<kScanBlock>(nb, &tb, sh);
    int tOff = block_excl_scan_n<kScanBlock>(nt, &tt, sh);
    if (it < P.nItems) {
      g.bandBase = bandTot + bOff;
      g.taskBase = taskTot + tOff;
      P.geo[it] = g;
      // compact per-call record for phase B
      P.itemPaint[it] = make_int4((int)f.rgbx, f.mode, g.taskBase, g.nWin);
      P.itemSpan[it] = make_int4(g.rowStart, g.rowCount, f.xlo, f.xhi);
    }
    bandTot += tb;
    taskTot += tt;
  }
  if (tid == 0) {
    P.counters[kCntBands] = bandTot;
    P.counters[kCntTasks] = taskTot;
    if (bandTot > P.bandCap || taskTot > P.taskCap) atomicOr(&P.counters[kCntError], 2);
  }
}

// Band b of its call: top / bottom rows (partitionSegments 1183-1192).
__device__ __forceinline__ void band_rows(const FillInfo& f, int k, float* top, float* bot) {
  *top = (float)(f.startY + k * f.bandH);
  *bot = (k == f.nParts - 1) ? (float)f.pathHeight : (float)(f.startY + (k + 1) * f.bandH);
}


// A grid-stride loop over items [0, n) laid out by XCD: workgroups go to
// the kXcds XCDs round robin, and XCD x takes the x-th contiguous share of
// the items, so that neighbouring items (the bands of one call, which read
// the same segments) share its L2.
constexpr int kXcds = 8;
struct XcdSpan {
  int begin, end, step;
};
__device__ __forceinline__ XcdSpan xcd_span(int n) {
  const int per = (int)gridDim.x / kXcds;
  if (per == 0) return XcdSpan{(int)blockIdx.x, n, (int)gridDim.x};
  const int xcd = (int)blockIdx.x % kXcds, li = (int)blockIdx.x / kXcds;
  if (li >= per) return XcdSpan{0, 0, 1};  // beyond the last full round of XCDs
  return XcdSpan{(int)((int64_t)n * xcd / kXcds) + li, (int)((int64_t)n * (xcd + 1) / kXcds), per};
}

// One wave per band: entry count and requiresAntiAliasing.
__global__ void __launch_bounds__(64) scene_band_count_kernel(ScenePools P) {
  const int lane = threadIdx.x;
  const int nb = P.counters[kCntBands];
  if (nb > P.bandCap) return;
  const XcdSpan sp = xcd_span(nb);
  for (int b = sp.begin; b < sp.end; b += sp.step) {
    const int it = uni(P.bandItem[b]);
    ItemGeom g = P.geo[it];
    g.bandBase = uni(g.bandBase);
    g.segBase = uni(g.segBase);
    g.segCount = uni(g.segCount);
    const int k = b - g.bandBase;
    int cnt = 0;
    bool aa = false;
    const int s0 = g.segBase, s1 = g.segBase + g.segCount;  // s0 is 64-aligned
    const int cEnd = (s1 + 63) >> 6;
    for (int cb = s0 >> 6; cb < cEnd; cb += 64) {
      // chunks that may reach band k, 64 chunk ranges at a time
      bool may = false;
      if (cb + lane < cEnd) {
        const int2 cr = P.segChunk[cb + lane];
        may = k >= cr.x && k <= cr.y;
      }
      uint64_t cm = __ballot(may);
      while (cm) {  // four candidate chunks at a time: their loads overlap
        int2 sb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          sb[u] = make_int2(1, 0);  // empty band range
          if (cm) {
            const int j = ((cb + __builtin_ctzll(cm)) << 6) + lane;
            cm &= cm - 1;
            if (j < s1) sb[u] = P.segBand[j];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const bool in = k >= sb[u].x && k <= (sb[u].y & ~kSegAA);
          cnt += __popcll(__ballot(in));
          if (__ballot(in && (sb[u].y & kSegAA))) aa = true;
        }
      }
    }
    if (lane == 0) {
      PartInfo pi;
      pi.entryBase = 0;
      pi.count = cnt;
      pi.flags = aa ? 1 : 0;
      pi.pad = 0;
      P.parts[b] = pi;
    }
  }
}

// One block: entry offsets of all bands, largest band.
// Band -> call map (a workgroup per call, its bands across the threads).
__global__ void __launch_bounds__(kBlock) scene_band_map_kernel(ScenePools P) {
  const int tid = threadIdx.x;
  for (int it = blockIdx.x; it < P.nItems; it += gridDim.x) {
    const ItemGeom g = P.geo[it];
    const FillInfo f = P.fills[it];
    const int nb = (g.status == kOk && f.pathWidth > 0) ? f.nParts : 0;
    if (g.bandBase + nb > P.bandCap) continue;
    for (int b = tid; b < nb; b += kBlock) P.bandItem[g.bandBase + b] = it;
  }
}

// Band-entry offsets: a scan of the band entry counts over kEntryTiles
// contiguous tiles of bands.  1. tile sums (and the largest band).
__device__ __forceinline__ int2 entry_tile(int nb) {
  const int per = (nb + kEntryTiles - 1) / kEntryTiles;
  const int lo = (int)blockIdx.x * per;
  return make_int2(lo, lo + per < nb ? lo + per : nb);
}

__global__ void __launch_bounds__(kBlock) scene_entry_tile_kernel(ScenePools P) {
  __shared__ int sh[kBlock / 64];
  const int nb = P.counters[kCntBands];
  if (nb > P.bandCap) return;
  const int2 r = entry_tile(nb);
  int sum = 0, mx = 0;
  for (int b = r.x + (int)threadIdx.x; b < r.y; b += kBlock) {
    const int c = P.parts[b].count;
    sum += c;
    mx = c > mx ? c : mx;
  }
  sum = block_sum(sum, sh);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0 && mx > 0) atomicMax(&P.counters[kCntMaxBand], mx);
  if (threadIdx.x == 0) P.counters[kCntTileSum + blockIdx.x] = sum;
}

// 2. tile offsets (in place of the tile sums) and the entry total.
__global__ void __launch_bounds__(kEntryTiles) scene_entry_scan_kernel(ScenePools P) {
  __shared__ int sh[kEntryTiles / 64];
  const int nb = P.counters[kCntBands];
  if (nb > P.bandCap) return;
  const int t = threadIdx.x;
  int tot;
  const int off = block_excl_scan_n<kEntryTiles>(P.counters[kCntTileSum + t], &tot, sh);
  P.counters[kCntTileSum + t] = off;
  if (t == 0) {
    P.counters[kCntEntries] = tot;
    if (tot > P.entryCap) atomicOr(&P.counters[kCntError], 4);
  }
}

// 3. entry offset of every band.
__global__ void __launch_bounds__(kBlock) scene_entry_apply_kernel(ScenePools P) {
  __shared__ int sh[kBlock / 64];
  const int nb = P.counters[kCntBands];
  if (nb > P.bandCap) return;
  const int2 r = entry_tile(nb);
  int base = P.counters[kCntTileSum + blockIdx.x];
  for (int b0 = r.x; b0 < r.y; b0 += kBlock) {
    const int b = b0 + (int)threadIdx.x;
    const int c = b < r.y ? P.parts[b].count : 0;
    int t;
    const int off = block_excl_scan(c, &t, sh);
    if (b < r.y) P.parts[b].entryBase = base + off;
    base += t;
  }
}

// Entry of segment s in band [top, bot]: slope-intercept form of the whole
// segment, y range clipped to the band when the segment spans it
// (partitionSegments 1236-1248).  *clipped receives the clipped segment.
__device__ __forceinline__ Entry band_entry(float4 s, float top, float bot, float4* clipped) {
  V2 at = v2(s.x, s.y), to = v2(s.z, s.w);
  Entry e = init_entry(at, to);
  if (at.y <= top && to.y >= bot) {
    V2 c = v2(0, 0);
    clip_to_line(at, to, top, &c);
    at = c;
    clip_to_line(at, to, bot, &c);
    to = c;
    e.aty = at.y;
    e.toy = to.y;
  }
  *clipped = make_float4(at.x, at.y, to.x, to.y);
  return e;
}

// One wave per band: ordered entries, clipping, two-spanning flag
// (partitionSegments 1194-1262).  Lanes exchange data only through ballots
// and shuffles, which every lane executes.
__global__ void __launch_bounds__(64) scene_band_write_kernel(ScenePools P) {
  const int lane = threadIdx.x;
  const int nb = P.counters[kCntBands];
  if (nb > P.bandCap || P.counters[kCntEntries] > P.entryCap) return;
  const XcdSpan sp = xcd_span(nb);
  for (int b = sp.begin; b < sp.end; b += sp.step) {
    const int it = uni(P.bandItem[b]);
    ItemGeom g = P.geo[it];
    g.bandBase = uni(g.bandBase);
    g.segBase = uni(g.segBase);
    g.segCount = uni(g.segCount);
    FillInfo f = P.fills[it];
    f.startY = uni(f.startY);
    f.bandH = uni(f.bandH);
    f.nParts = uni(f.nParts);
    f.pathHeight = uni(f.pathHeight);
    const int k = b - g.bandBase;
    PartInfo pi = P.parts[b];
    pi.entryBase = uni(pi.entryBase);
    pi.count = uni(pi.count);
    pi.flags = uni(pi.flags);
    float top, bot;
    band_rows(f, k, &top, &bot);
    const int s0 = g.segBase, s1 = g.segBase + g.segCount;
    if (pi.count != 2) {
      int pos = 0;
      const int cEnd = (s1 + 63) >> 6;
      for (int cb = s0 >> 6; cb < cEnd && pos < pi.count; cb += 64) {
        bool may = false;
        if (cb + lane < cEnd) {
          const int2 cr = P.segChunk[cb + lane];
          may = k >= cr.x && k <= cr.y;
        }
        uint64_t cm = __ballot(may);
        while (cm && pos < pi.count) {  // four candidate chunks at a time
          int2 sb[4];
          int js[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            sb[u] = make_int2(1, 0);  // empty band range
            js[u] = 0;
            if (cm) {
              js[u] = ((cb + __builtin_ctzll(cm)) << 6) + lane;
              cm &= cm - 1;
              if (js[u] < s1) sb[u] = P.segBand[js[u]];
            }
          }
          bool in[4];
          float4 sg[4];
          int8_t sw[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            in[u] = k >= sb[u].x && k <= (sb[u].y & ~kSegAA);
            sg[u] = make_float4(0, 0, 0, 0);
            sw[u] = 0;
            if (in[u]) {
              sg[u] = P.segs[js[u]];
              sw[u] = P.segW[js[u]];
            }
          }
          const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const uint64_t m = __ballot(in[u]);
            if (in[u]) {
              const int p = pos + __popcll(m & below);
              float4 cl;
              P.entries[pi.entryBase + p] = band_entry(sg[u], top, bot, &cl);
              P.entW[pi.entryBase + p] = sw[u];
            }
            pos += __popcll(m);
          }
        }
      }
      continue;
    }
    // two entries: gather both segments into every lane, decide the
    // two-spanning flag and order (1250-1262), then write
    float4 g0 = make_float4(0, 0, 0, 0), g1 = g0;
    int w0 = 0, w1 = 0, found = 0;
    const int cEnd2 = (s1 + 63) >> 6;
    for (int cb = s0 >> 6; cb < cEnd2 && found < 2; cb += 64) {
      bool may = false;
      if (cb + lane < cEnd2) {
        const int2 cr = P.segChunk[cb + lane];
        may = k >= cr.x && k <= cr.y;
      }
      uint64_t cm = __ballot(may);
      while (cm && found < 2) {
      const int c = cb + __builtin_ctzll(cm);
      cm &= cm - 1;
      const int j = (c << 6) + lane;
      bool in = false;
      float4 sg = make_float4(0, 0, 0, 0);
      int wv = 0;
      if (j < s1) {
        const int2 sb = P.segBand[j];
        in = k >= sb.x && k <= (sb.y & ~kSegAA);
        if (in) {
          sg = P.segs[j];
          wv = P.segW[j];
        }
      }
      uint64_t m = __ballot(in);
      while (m != 0 && found < 2) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        float4 t = make_float4(__shfl(sg.x, src, 64), __shfl(sg.y, src, 64), __shfl(sg.z, src, 64),
                               __shfl(sg.w, src, 64));
        int tw = __shfl(wv, src, 64);
        if (found == 0) {
          g0 = t;
          w0 = tw;
        } else {
          g1 = t;
          w1 = tw;
        }
        found++;
      }
      }
    }
    float4 a, c;
    Entry e0 = band_entry(g0, top, bot, &a);
    Entry e1 = band_entry(g1, top, bot, &c);
    int32_t flags = pi.flags;
    bool swap = false;
    V2 a0 = v2(a.x, a.y), a1 = v2(a.z, a.w), c0 = v2(c.x, c.y), c1 = v2(c.z, c.w);
    if (!intersects_inside(a0, a1, c0, c1)) {
      if (a0.y <= top && a1.y >= bot && c0.y <= top && c1.y >= bot) {
        flags |= 2;
        swap = midpoint_x(a0, a1) > midpoint_x(c0, c1);
      }
    }
    // every lane stores the same values (conditional copies, not ?: on the
    // two lvalues, which would keep both entries in memory)
    Entry first = e0, second = e1;
    if (swap) {
      first = e1;
      second = e0;
    }
    P.entries[pi.entryBase] = first;
    P.entries[pi.entryBase + 1] = second;
    P.entW[pi.entryBase] = (int8_t)(swap ? w1 : w0);
    P.entW[pi.entryBase + 1] = (int8_t)(swap ? w0 : w1);
    P.parts[b].flags = flags;
  }
}

// ---------------------------------------------------------------------------
// row bins for phase B: one wave per bin, calls compacted in call order
// ---------------------------------------------------------------------------
template <bool kWrite>
__global__ void __launch_bounds__(64) scene_bin_kernel(ScenePools P) {
  const int lane = threadIdx.x;
  const int nBins = (P.H + kBinRows - 1) / kBinRows;
  if (kWrite && P.counters[kCntBinEntries] > P.binCap) return;
  for (int bin = blockIdx.x; bin < nBins; bin += gridDim.x) {
    const int r0 = bin * kBinRows, r1 = r0 + kBinRows;
    int pos = kWrite ? P.binOff[bin] : 0;
    for (int i0 = 0; i0 < P.nItems; i0 += 64) {
      const int it = i0 + lane;
      bool in = false;
      if (it < P.nItems) {
        int4 r = P.itemRows[it];
        in = r.x < r1 && r.y > r0;
      }
      uint64_t m = __ballot(in);
      if (kWrite && in) P.binList[pos + __popcll(m & ((1ull << lane) - 1ull))] = it;
      pos += __popcll(m);
    }
    if (!kWrite) P.binCount[bin] = pos;  // every lane stores the same value
  }
}

__global__ void __launch_bounds__(kScanBlock) scene_bin_scan_kernel(ScenePools P) {
  __shared__ int sh[kScanBlock / 64];
  const int nBins = (P.H + kBinRows - 1) / kBinRows;
  int tot = 0;
  for (int b0 = 0; b0 < nBins; b0 += kScanBlock) {
    int b = b0 + threadIdx.x;
    int c = b < nBins ? P.binCount[b] : 0;
    int t;
    int off = block_excl_scan_n<kScanBlock>(c, &t, sh);
    if (b < nBins) P.binOff[b] = tot + off;
    tot += t;
  }
  if (threadIdx.x == 0) {
    P.binOff[nBins] = tot;
    P.counters[kCntBinEntries] = tot;
    if (tot > P.binCap) atomicOr(&P.counters[kCntError], 32);
  }
}

// ---------------------------------------------------------------------------
// phase A: row programs of every (call, row, window) -> sparse runs
// ---------------------------------------------------------------------------

// Next work item of this wave's queue, or a value >= n when it is empty.
// The items are dealt to the kQueues queues in granules of G consecutive
// items, round robin; a queue hands its items out in order.  Workgroups go
// to the XCDs round robin, so queue q is served by XCD q % 8 alone: a
// granule's items (neighbouring rows of one call, in phase A) share that
// XCD's L2.
// Every lane takes part in the atomic (lane 0 adds 1, the others 0) and lane
// 0's old value is broadcast with a shuffle: broadcasting it through LDS
// from a lane-0-only branch is not safe, because __syncthreads() compiles to
// nothing in a one-wave workgroup and the compiler may then let lanes 1..63
// loop back to the LDS read without lane 0, spinning on a stale value.
template <int G = 1>
__device__ __forceinline__ int queue_next(int32_t* heads, int n) {
  const int lane = threadIdx.x & 63;
  const int q = (int)(blockIdx.x % kQueues);
  int k = atomicAdd(heads + q * kQueueStride, lane == 0 ? 1 : 0);
  k = __shfl(k, 0, 64);
  // the same in every lane: keep it (and what derives from it) scalar
  k = __builtin_amdgcn_readfirstlane(k);
  // the item is increasing in k, so a queue is empty once one is >= n
  const int64_t item = ((int64_t)(k / G) * kQueues + q) * G + k % G;
  return item < n ? (int)item : n;
}

// Takes n entries of a pool counter for the whole wave (same convergent
// form as queue_next).
__device__ __forceinline__ int queue_take(int32_t* counter, int n) {
  const int lane = threadIdx.x & 63;
  int base = atomicAdd(counter, lane == 0 ? n : 0);
  return __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
}

// Band of row y of a banded call (partitionSegments' layout).
__device__ __forceinline__ int band_of_row(const FillInfo& F, int y) {
  const int pi = (y - F.startY) / F.bandH;
  return pi > F.nParts - 1 ? F.nParts - 1 : pi;
}
// Cost class of band pi of a call: the bit length of its entry count.
__device__ __forceinline__ int band_class(const ScenePools& P, const ItemGeom& g, int pi) {
  const int cnt = P.parts[g.bandBase + pi].count;
  return cnt <= 0 ? 0 : 32 - __clz(cnt);
}

// Maps work items to calls and records their cost classes, counting them
// per class (per block in LDS, then one global atomic per class and block).
// The classes of a call's bands are staged in LDS first.
constexpr int kLdsBandClasses = 4096;
constexpr int kTaskMapBlock = 1024;
__global__ void __launch_bounds__(kTaskMapBlock) scene_task_map_kernel(ScenePools P) {
  __shared__ int32_t hist[kCostClasses];
  __shared__ int8_t cls[kLdsBandClasses];
  const int tid = threadIdx.x, lane = tid & 63;
  if (P.counters[kCntError]) return;
  if (tid < kCostClasses) hist[tid] = 0;
  __syncthreads();
  for (int it = blockIdx.x; it < P.nItems; it += gridDim.x) {
    const ItemGeom g = P.geo[it];
    const FillInfo F = P.fills[it];
    const int nt = g.rowCount * g.nWin;
    if (nt <= 0) continue;
    const bool banded = F.nParts > 0 && F.bandH > 0;
    const int nb = banded ? F.nParts : 0;
    for (int b = tid; b < nb && b < kLdsBandClasses; b += kTaskMapBlock) {
      const int cnt = P.parts[g.bandBase + b].count;
      cls[b] = (int8_t)(cnt <= 0 ? 0 : 32 - __clz(cnt));
    }
    __syncthreads();
    for (int t0 = 0; t0 < nt; t0 += kTaskMapBlock) {
      const int t = t0 + tid;
      const bool on = t < nt;
      int c = -1;
      bool grouped = false;
      if (on) {
        // classes of this task's row and of its group's rows (the bands
        // from the group's first row to its last; every row of a call lies
        // in [startY, pathHeight))
        const int r = t / g.nWin, rg = r - r % kRowGroup;
        const int re = (rg + kRowGroup < g.rowCount ? rg + kRowGroup : g.rowCount) - 1;
        int own = 0, chunk = 0;
        if (banded) {
          const int pr = band_of_row(F, g.rowStart + r), p0 = band_of_row(F, g.rowStart + rg),
                    p1 = band_of_row(F, g.rowStart + re);
          for (int pi = p0; pi <= p1; pi++) {
            const int rc = pi < kLdsBandClasses ? cls[pi] : band_class(P, g, pi);
            if (pi == pr) own = rc;
            chunk = rc > chunk ? rc : chunk;
          }
        }
        grouped = chunk <= kGroupMaxClass;
        c = !grouped ? own : (r == rg ? chunk : -1);
        P.taskItem[g.taskBase + t] = grouped ? (int32_t)((uint32_t)it | 0x80000000u) : it;
        P.taskClass[g.taskBase + t] = (int8_t)c;
      }
      // one LDS atomic per class present in the wave
      uint64_t pending = __ballot(c >= 0);
      while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int cl = __builtin_amdgcn_readlane(c, leader);
        const uint64_t m = __ballot(c == cl);
        if (lane == leader) atomicAdd(&hist[cl], (int)__popcll(m));
        pending &= ~m;
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (tid < kCostClasses && hist[tid]) atomicAdd(&P.counters[kCntClassHist + tid], hist[tid]);
}

// Block b's contiguous share of the work items.
__device__ __forceinline__ int2 task_share(int ntasks) {
  const int per = ((ntasks + (int)gridDim.x - 1) / (int)gridDim.x + 63) & ~63;
  const int lo = (int)blockIdx.x * per;
  return make_int2(lo, lo + per < ntasks ? lo + per : ntasks);
}

// Start of each cost class in the work-item order, largest class first, and
// the number of items in classes >= bigClass (the large-band launch's share).
__global__ void scene_task_scan_kernel(int32_t* counters, int bigClass) {
  if (threadIdx.x != 0) return;
  int off = 0;
  for (int c = kCostClasses - 1; c >= 0; c--) {
    if (c == bigClass - 1) counters[kCntBigTasks] = off;
    counters[kCntClassCursor + c] = off;
    off += counters[kCntClassHist + c];
  }
  if (bigClass <= 0) counters[kCntBigTasks] = off;
  counters[kCntGroups] = off;
}

// Places every work item in the order by its cost class: each block
// reserves a range of every class for its share of the items, then deals it
// out with LDS atomics.
__global__ void __launch_bounds__(64) scene_task_order_kernel(ScenePools P) {
  __shared__ int32_t cnt[kCostClasses], cur[kCostClasses];
  const int lane = threadIdx.x;
  if (P.counters[kCntError]) return;
  if (lane < kCostClasses) cnt[lane] = 0;
  __syncthreads();
  const int2 r = task_share(P.counters[kCntTasks]);
  for (int t = r.x + lane; t < r.y; t += 64)
    if (P.taskClass[t] >= 0) atomicAdd(&cnt[P.taskClass[t]], 1);
  __syncthreads();
  if (lane < kCostClasses) cur[lane] = cnt[lane] ? atomicAdd(&P.counters[kCntClassCursor + lane], cnt[lane]) : 0;
  __syncthreads();
  for (int t = r.x + lane; t < r.y; t += 64)
    if (P.taskClass[t] >= 0) P.taskOrder[atomicAdd(&cur[P.taskClass[t]], 1)] = t;
}

// kBig: the large-band launch (tasks [0, nBig) of the task order), else the
// small-band one; two instantiations so that profiles tell them apart.
template <bool kBig>
__global__ void __launch_bounds__(64, 5) scene_rows_kernel(ScenePools P) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x;
  const size_t fixedBytes = row_lds_fixed_bytes(kWinA);
  RowLds Ll = row_lds_carve(smem, smem + fixedBytes, P.ldsCap, kWinA);
  RowLds Lg = row_lds_carve(smem, P.gscratch + (size_t)blockIdx.x * row_cap_bytes(P.gcap), P.gcap, kWinA);
  // the wide path stages its keys through AX and the LDS arrays, which follow it
  Ll.wide = Lg.wide = (uint64_t*)Ll.AX;
  Ll.wideCap = Lg.wideCap = (int)((kWinA * 4 + row_cap_bytes(P.ldsCap)) / 8);
  for (int i = lane; i < kWinA + 4; i += 64) Ll.D[i] = 0;
  __syncthreads();
  if (P.counters[kCntError]) return;
  // this launch's share of the work-item order: the large-band items come first
  const int nBig = P.counters[kCntBigTasks];
  const int tBegin = kBig ? 0 : nBig;
  const int tEnd = kBig ? nBig : P.counters[kCntGroups];
  int32_t* heads = P.queues + (kBig ? 0 : kQueues * kQueueStride);
  // Each wave appends its runs to chunks of the run pool it takes one at a
  // time (one atomic per chunk, not per task).
  int cursor = 0, chunkEnd = 0, used = 0;
  for (;;) {
    const int q = tBegin + queue_next<kRowGranule>(heads, tEnd - tBegin);
    if (q >= tEnd) break;
    // Work item, call and the call's records are the same in every lane:
    // read them into scalar registers.
    const int first = uni(P.taskOrder[q]);
    const int itw = uni(P.taskItem[first]);
    const int it = itw & 0x7fffffff;
    ItemGeom g = P.geo[it];
    g.rowStart = uni(g.rowStart);
    g.nWin = uni(g.nWin);
    g.winBase = uni(g.winBase);
    g.taskBase = uni(g.taskBase);
    g.bandBase = uni(g.bandBase);
    g.rowCount = uni(g.rowCount);
    FillInfo Fu = P.fills[it];
    Fu.startX = uni(Fu.startX);
    Fu.startY = uni(Fu.startY);
    Fu.pathWidth = uni(Fu.pathWidth);
    Fu.pathHeight = uni(Fu.pathHeight);
    Fu.nParts = uni(Fu.nParts);
    Fu.bandH = uni(Fu.bandH);
    Fu.mode = uni(Fu.mode);
    Fu.rule = uni(Fu.rule);
    Fu.xlo = uni(Fu.xlo);
    Fu.xhi = uni(Fu.xhi);
    Fu.canvasW = uni(Fu.canvasW);
    const FillInfo& F = Fu;
    const int local = first - g.taskBase;
    const int win = local % g.nWin;
    const int r0 = local / g.nWin;
    const int r1 = itw >= 0 ? r0 + 1 : (r0 + kRowGroup < g.rowCount ? r0 + kRowGroup : g.rowCount);
    const bool mask = F.mode == kMask;
    for (int r = r0; r < r1; r++) {
      const int y = g.rowStart + r;
      const int task = g.taskBase + r * g.nWin + win;
      RowCtx C;
      C.F = &Fu;
      C.parts = P.parts + g.bandBase;
      C.E = P.entries;
      C.EW = P.entW;
      C.y = y;
      C.cx0 = mask ? 0 : F.xlo;
      C.cx1 = mask ? P.W : F.xhi;
      C.wx0 = g.winBase + win * kWinA;
      C.wx1 = C.wx0 + kWinA < C.cx1 ? C.wx0 + kWinA : C.cx1;
      // band of this row decides where the row scratch lives
      int pi = (y - F.startY) / F.bandH;
      if (pi > F.nParts - 1) pi = F.nParts - 1;
      const int cnt = uni(C.parts[pi].count);
      // a large band's row whose selection fits LDS still runs there
      bool inLds = cnt <= P.ldsCap;
      if (!inLds && cnt <= P.gcap) inLds = row_selection_count(C) <= P.ldsCap;
      RowLds& L = inLds ? Ll : Lg;
      int32_t kind = (inLds || cnt <= P.gcap) ? row_program(C, L) : -1;
      RowRuns rr;
      rr.off = 0;
      rr.n = 0;
      rr.kind = kind;
      rr.pad = 0;
      if (kind < 0) {
        atomicOr(&P.counters[kCntError], 8);
      } else if (kind != kRowNone) {
        // compact D[0..wlen] (nonzero deltas) and, for ops rows, the area
        // factors of the pixels whose code is an area code, in one pass.
        // Coverage rows need no codes, only the nonzero deltas.  The wave's
        // run chunk is first made to hold the most the row can write (wlen + 1
        // deltas, and wlen area factors on ops rows).  D has ww + 4 entries
        // and is zero beyond wlen, so a lane with i0 <= wlen moves its 4
        // entries as one 16-byte access.
        const int wlen = C.wx1 - C.wx0;
        const bool ops = kind == kRowOps;
        const int most = ops ? 2 * wlen + 1 : wlen + 1;
        if (cursor + most > chunkEnd) {
          cursor = queue_take(&P.counters[kCntRuns], kRunChunk);
          chunkEnd = cursor + kRunChunk;
        }
        rr.off = cursor;
        int written = 0;
        if (chunkEnd > P.runCap) {
          atomicOr(&P.counters[kCntError], 16);
          for (int i = lane; i <= wlen; i += 64) L.D[i] = 0;
        } else {
          int carry = 0;
          for (int base = 0; base <= wlen; base += 4 * kWave) {
            const int i0 = base + 4 * lane;
            int4 dv = make_int4(0, 0, 0, 0);
            if (i0 <= wlen) dv = *(const int4*)(Ll.D + i0);
            const int32_t d[4] = {dv.x, dv.y, dv.z, dv.w};
            bool area[4] = {false, false, false, false};
            if (ops) {
              const int32_t c0 = d[0], c1 = c0 + d[1], c2 = c1 + d[2], c3 = c2 + d[3];
              int32_t tot;
              const int32_t off = carry + wave_excl_scan(c3, &tot);
              carry += tot;
              const int32_t code[4] = {off + c0, off + c1, off + c2, off + c3};
#pragma unroll
              for (int q = 0; q < 4; q++) area[q] = i0 + q < wlen && code[q] >= kArea && code[q] < kClear;
            }
            int n = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) n += (d[q] != 0 ? 1 : 0) + (area[q] ? 1 : 0);
            int ntot;
            const int p = wave_excl_scan(n, &ntot);
            if (i0 <= wlen) {
              int2* out = P.runs + rr.off + written + p;
              int w = 0;
#pragma unroll
              for (int q = 0; q < 4; q++) {
                const int x = C.wx0 + i0 + q;
                if (d[q] != 0) out[w++] = make_int2(x, d[q]);
                if (area[q]) out[w++] = make_int2(x | kRunArea, (int32_t)L.AX[i0 + q]);
              }
              *(int4*)(Ll.D + i0) = make_int4(0, 0, 0, 0);
            }
            written += ntot;
          }
        }
        rr.n = written;
        cursor += written;
        used += written;
        __syncthreads();
      }
      P.tasks[task] = rr;  // every lane stores the same value
    }
  }
  if (threadIdx.x == 0 && used) atomicAdd(&P.counters[kCntRunsUsed], used);
}

// ---------------------------------------------------------------------------
// phase B: canvas row windows in registers, calls applied in order
// ---------------------------------------------------------------------------
// The canvas window lives in LDS (1024 px) while the calls are applied, so
// that one copy of the per-pixel blend code serves every pixel: the kernel
// stays small enough for the instruction cache.
template <bool kA8>
__global__ void __launch_bounds__(64) scene_compose_kernel(ScenePools P, void* canvas,
                                                          unsigned long long* touchedOut,
                                                          int32_t* errOut) {
  __shared__ __align__(16) int32_t D[kWinB + 4];
  __shared__ uint32_t AX[kWinB];
  __shared__ uint32_t WIN[kWinB];
  __shared__ int32_t Q[2 * kWave];  // pixels queued for the blend
  const int lane = threadIdx.x;
  for (int i = lane; i < kWinB + 4; i += 64) D[i] = 0;
  const int W = P.W, H = P.H;
  const int winsPerRow = (W + kWinB - 1) / kWinB;
  const int nUnits = H * winsPerRow;
  const bool vec = ((W & 3) == 0);
  uint32_t ntouch = 0;
  if (const int32_t err = P.counters[kCntError]) {
    // an earlier stage ran out of a pool: the canvas stays untouched and the
    // error reaches the host through the lane's error word
    if (blockIdx.x == 0 && lane == 0) atomicOr(errOut, err);
    return;
  }
  for (;;) {
    const int unit = queue_next<kComposeGranule>(P.queues + 2 * kQueues * kQueueStride, nUnits);
    if (unit >= nUnits) break;
    const int y = unit / winsPerRow;
    const int b0 = (unit % winsPerRow) * kWinB;
    const int b1 = b0 + kWinB < W ? b0 + kWinB : W;
    bool loaded = false, dirty = false;
    const int bin = y / kBinRows;
    const int l1 = P.binOff[bin + 1];
    for (int i0 = P.binOff[bin]; i0 < l1; i0 += 64) {
      const int li = i0 + lane;
      bool cov = false;
      int4 paint = make_int4(0, 0, 0, 0), span = paint;
      if (li < l1) {
        const int it = P.binList[li];
        const int4 r = P.itemRows[it];
        cov = y >= r.x && y < r.y && r.w > b0 && r.z < b1;
        if (cov) {  // each lane fetches its call's record: the loads overlap
          paint = P.itemPaint[it];
          span = P.itemSpan[it];
        }
      }
      // each covering lane also fetches its call's first phase-A window
      // record for this row and window (pre.w = that window, -1 if none)
      int4 pre = make_int4(0, 0, 0, -1);
      if (cov && (paint.y != kMask || (y >= span.x && y < span.x + span.y))) {
        const int winBase = paint.y == kMask ? 0 : (span.z & ~3);
        const int d = b0 - kWinA - winBase;  // windows ending before b0 are skipped
        const int w = d > 0 ? (d + kWinA - 1) / kWinA : 0;
        if (w < paint.w && winBase + w * kWinA < b1) {
          const RowRuns rr = P.tasks[paint.z + (y - span.x) * paint.w + w];
          pre = make_int4(rr.off, rr.n, rr.kind, w);
        }
      }
      uint64_t m = __ballot(cov);
      while (m) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        const uint32_t rgbx = (uint32_t)__builtin_amdgcn_readlane(paint.x, src);
        const int mode = __builtin_amdgcn_readlane(paint.y, src);
        const int taskBase = __builtin_amdgcn_readlane(paint.z, src);
        const int nWin = __builtin_amdgcn_readlane(paint.w, src);
        const int rowStart = __builtin_amdgcn_readlane(span.x, src);
        const int rowCount = __builtin_amdgcn_readlane(span.y, src);
        const int xlo = __builtin_amdgcn_readlane(span.z, src);
        const int xhi = __builtin_amdgcn_readlane(span.w, src);
        const int preW = __builtin_amdgcn_readlane(pre.w, src);
        if (!loaded) {  // canvas window -> LDS
#pragma unroll
          for (int c = 0; c < kChunksB; c++) {
            const int x0 = b0 + c * 256 + 4 * lane;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (kA8) {
              const uint8_t* row = (const uint8_t*)canvas + (size_t)y * W;
              uint32_t u[4] = {0, 0, 0, 0};
              if (vec && x0 + 3 < b1) {
                uint32_t w4 = *(const uint32_t*)(row + x0);
#pragma unroll
                for (int q = 0; q < 4; q++) u[q] = ((w4 >> (8 * q)) & 255u) * 0x01010101u;
              } else {
#pragma unroll
                for (int q = 0; q < 4; q++) u[q] = (x0 + q < b1) ? row[x0 + q] * 0x01010101u : 0u;
              }
              v = make_uint4(u[0], u[1], u[2], u[3]);
            } else {
              const uint32_t* row = (const uint32_t*)canvas + (size_t)y * W;
              if (vec && x0 + 3 < b1) {
                v = *(const uint4*)(row + x0);
              } else {
                v.x = (x0 < b1) ? row[x0] : 0u;
                v.y = (x0 + 1 < b1) ? row[x0 + 1] : 0u;
                v.z = (x0 + 2 < b1) ? row[x0 + 2] : 0u;
                v.w = (x0 + 3 < b1) ? row[x0 + 3] : 0u;
              }
            }
            const int i4 = c * 256 + 4 * lane;
            WIN[i4] = v.x;
            WIN[i4 + 1] = v.y;
            WIN[i4 + 2] = v.z;
            WIN[i4 + 3] = v.w;
          }
          loaded = true;
        }
        const bool mask = mode == kMask;
        if (mask && !(y >= rowStart && y < rowStart + rowCount)) {
          // rows outside a mask fill's path are cleared (paths.nim:1910-1912)
          for (int c = 0; c < kChunksB; c++) {
            const int i4 = c * 256 + 4 * lane;
            WIN[i4] = WIN[i4 + 1] = WIN[i4 + 2] = WIN[i4 + 3] = 0u;
            const int x0 = b0 + c * 256 + 4 * lane;
            const int n = b1 - x0;
            ntouch += n <= 0 ? 0 : (n >= 4 ? 4 : n);
          }
          dirty = true;
          continue;
        }
        // gather this call's runs for row y into D / AX.  Every phase-A
        // window's deltas sum to zero, so windows that end before b0 or
        // start at b1 or later change nothing here.
        const int tb = taskBase + (y - rowStart) * nWin;
        const int winBase = mask ? 0 : (xlo & ~3);
        int carry = 0;
        int kind = kRowNone;
        for (int w = 0; w < nWin; w++) {
          const int wx0 = winBase + w * kWinA;
          if (wx0 >= b1) break;
          if (wx0 + kWinA < b0) continue;
          RowRuns rr;
          if (w == preW) {
            rr.off = __builtin_amdgcn_readlane(pre.x, src);
            rr.n = __builtin_amdgcn_readlane(pre.y, src);
            rr.kind = __builtin_amdgcn_readlane(pre.z, src);
          } else {
            rr = P.tasks[tb + w];
            rr.off = uni(rr.off);
            rr.n = uni(rr.n);
            rr.kind = uni(rr.kind);
          }
          if (rr.kind > kind) kind = rr.kind;
          for (int e0 = 0; e0 < rr.n; e0 += 64) {
            int e = e0 + lane;
            if (e < rr.n) {
              int2 r = P.runs[rr.off + e];
              int x = r.x & kRunXMask;
              if (r.x & kRunArea) {
                if (x >= b0 && x < b1) AX[x - b0] = (uint32_t)r.y;
              } else if (x < b0) {
                carry += r.y;
              } else if (x < b1) {
                atomicAdd(&D[x - b0], r.y);
              }
            }
          }
        }
        carry = wave_incl_sum(carry);
        carry = __builtin_amdgcn_readlane(carry, kWave - 1);
        __syncthreads();
        if (kind == kRowNone) continue;
        const int lo = mask ? b0 : (xlo > b0 ? xlo : b0);
        const int hi = mask ? b1 : (xhi < b1 ? xhi : b1);
        // The call's deltas lie in [xlo & ~3, xhi] (phase-A windows start on
        // a multiple of 4, and a run may end at xhi): the 64-px slices from
        // lo's to hi's hold them all, since xlo & ~3 shares xlo's slice.
        // Pixel codes in 64-px slices (lane l takes pixel l of the slice).
        // The pixels this call changes are queued, packed as code << 11 |
        // index, and the blend runs on full batches of 64 queued pixels, so
        // partly covered slices do not leave lanes idle in the blend.  A
        // call changes each pixel at most once, so a batch never holds the
        // same pixel twice.
        const bool maskCov = kind == kRowCoverage && mask;  // every pixel is written
        int s = (lo - b0) / 64;
        const int sEnd = (hi - b0) / 64 + 1 < kWinB / 64 ? (hi - b0) / 64 + 1 : kWinB / 64;
        int nq = 0;
        for (;;) {
          while (s < sEnd && nq < kWave) {
            const int idx = s * 64 + lane;
            const int x = b0 + idx;
            const int32_t d = D[idx];
            D[idx] = 0;
            const int32_t incl = wave_incl_sum(d);
            const int32_t code = carry + incl;
            carry += __builtin_amdgcn_readlane(incl, kWave - 1);
            const bool work = x >= lo && x < hi && (code != 0 || maskCov);
            const uint64_t wm = __ballot(work);
            const int pos = nq + __builtin_amdgcn_mbcnt_hi((uint32_t)(wm >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)wm, 0u));
            if (work) Q[pos] = code * 2048 + idx;
            nq += __popcll(wm);
            s++;
          }
          if (nq == 0) break;
          if (lane < nq) {
            const int32_t v = Q[lane];
            const int idx = v & 2047;
            const int32_t code = v >> 11;
            uint32_t ax = (kind == kRowOps && code >= kArea && code < kClear) ? AX[idx] : 0u;
            bool t = false;
            WIN[idx] = apply_code(kind, code, ax, WIN[idx], rgbx, mode, &t);
            if (t) {
              ntouch++;
              dirty = true;
            }
          }
          // keep the overflow (queue positions 64 .. nq-1) for the next batch
          nq = nq > kWave ? nq - kWave : 0;
          if (lane < nq) Q[lane] = Q[kWave + lane];
        }
        __syncthreads();
      }
    }
    if (__ballot(dirty) != 0) {  // LDS window -> canvas
#pragma unroll
      for (int c = 0; c < kChunksB; c++) {
        const int x0 = b0 + c * 256 + 4 * lane;
        const int i4 = c * 256 + 4 * lane;
        const uint4 v = make_uint4(WIN[i4], WIN[i4 + 1], WIN[i4 + 2], WIN[i4 + 3]);
        if (kA8) {
          uint8_t* row = (uint8_t*)canvas + (size_t)y * W;
          if (vec && x0 + 3 < b1) {
            *(uint32_t*)(row + x0) = ch_a(v.x) | (ch_a(v.y) << 8) | (ch_a(v.z) << 16) | (ch_a(v.w) << 24);
          } else {
            if (x0 < b1) row[x0] = (uint8_t)ch_a(v.x);
            if (x0 + 1 < b1) row[x0 + 1] = (uint8_t)ch_a(v.y);
            if (x0 + 2 < b1) row[x0 + 2] = (uint8_t)ch_a(v.z);
            if (x0 + 3 < b1) row[x0 + 3] = (uint8_t)ch_a(v.w);
          }
        } else {
          uint32_t* row = (uint32_t*)canvas + (size_t)y * W;
          if (vec && x0 + 3 < b1) {
            *(uint4*)(row + x0) = v;
          } else {
            if (x0 < b1) row[x0] = v.x;
            if (x0 + 1 < b1) row[x0 + 1] = v.y;
            if (x0 + 2 < b1) row[x0 + 2] = v.z;
            if (x0 + 3 < b1) row[x0 + 3] = v.w;
          }
        }
      }
    }
  }
  if (touchedOut) {
    ntouch = (uint32_t)wave_incl_sum((int32_t)ntouch);
    ntouch = __builtin_amdgcn_readlane(ntouch, kWave - 1);
    if (lane == 0 && ntouch) atomicAdd(touchedOut, (unsigned long long)ntouch);
  }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
int scene_set_lds_limits(size_t bytes) {
  // the kernel's static LDS (its task slot) counts against the same limit
  if (hipFuncSetAttribute((const void*)scene_rows_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)(bytes - 1024)) != hipSuccess ||
      hipFuncSetAttribute((const void*)scene_rows_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)(bytes - 1024)) != hipSuccess)
    return kErrDevice;
  return 0;
}

size_t scene_gscratch_bytes(int rowGrid, int gcap) { return (size_t)rowGrid * row_cap_bytes(gcap); }

// Clears the counters and queue heads of a render.  A kernel rather than a
// memset, so that a render captured in a graph holds kernel nodes only.
__global__ void __launch_bounds__(256) scene_clear_kernel(int32_t* counters, int nc, int32_t* queues,
                                                          int nq) {
  for (int i = threadIdx.x; i < nc; i += blockDim.x) counters[i] = 0;
  for (int i = threadIdx.x; i < nq; i += blockDim.x) queues[i] = 0;
}

// Text calls: one wave per glyph instance writes its commands and subpaths
// into the scene's recording (GlyphInst).  Points are transformed as
// Path.transform transforms them (paths.nim:264-333): m * (x, y) in float32.
__global__ void __launch_bounds__(256) scene_expand_kernel(const GlyphInst* inst, int n,
                                                           const SceneCmd* aCmds, const ScenePoly* aPolys,
                                                           SceneCmd* cmds, ScenePoly* polys) {
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < n; w += gridDim.x * 4) {
    const GlyphInst g = inst[w];
    const float* m = g.m;
    if (g.srcCmd < 0) {  // path.rect(x, y, w, h), clockwise unless m[4] (paths.nim:506-521)
      const float x = m[0], y = m[1], rw = m[2], rh = m[3];
      const bool ccw = m[4] != 0.0f;
      if (lane < 5) {
        SceneCmd c = {};
        c.kind = lane == 0 ? 1.0f : lane == 4 ? 0.0f : 2.0f;
        if (lane == 0) { c.p[0] = x; c.p[1] = y; }
        if (lane == 1) { c.p[0] = ccw ? x : x + rw; c.p[1] = ccw ? y + rh : y; }
        if (lane == 2) { c.p[0] = x + rw; c.p[1] = y + rh; }
        if (lane == 3) { c.p[0] = ccw ? x + rw : x; c.p[1] = ccw ? y : y + rh; }
        c.sx = x;
        c.sy = y;
        c.pad[0] = g.polyIdx;
        cmds[g.dstCmd + lane] = c;
      }
      if (lane == 0) {
        ScenePoly sp = {};
        sp.firstCmd = g.dstCmd;
        sp.numCmds = 5;
        polys[g.dstPoly] = sp;
      }
      continue;
    }
    for (int k = lane; k < g.nCmds; k += 64) {
      SceneCmd c = aCmds[g.srcCmd + k];
      if (!g.identity) {
        const int kind = (int)c.kind;
        const int pts = kind == 5 ? 3 : kind == 7 ? 2 : (kind == 1 || kind == 2) ? 1 : 0;
        for (int q = 0; q < pts; q++) {
          const float x = c.p[2 * q], y = c.p[2 * q + 1];
          c.p[2 * q] = m[0] * x + m[2] * y + m[4];
          c.p[2 * q + 1] = m[1] * x + m[3] * y + m[5];
        }
        const float sx = c.sx, sy = c.sy;
        c.sx = m[0] * sx + m[2] * sy + m[4];
        c.sy = m[1] * sx + m[3] * sy + m[5];
      }
      c.pad[0] += g.polyIdx;
      cmds[g.dstCmd + k] = c;
    }
    for (int k = lane; k < g.nPolys; k += 64) {
      ScenePoly sp = aPolys[g.srcPoly + k];
      sp.firstCmd = sp.firstCmd - g.srcCmd + g.dstCmd;
      polys[g.dstPoly + k] = sp;
    }
  }
}

int scene_launch_expand(const GlyphInst* inst, int32_t n, const SceneCmd* arenaCmds, const ScenePoly* arenaPolys,
                        SceneCmd* cmds, ScenePoly* polys, hipStream_t s) {
  if (n <= 0) return 0;
  int grid = (n + 3) / 4;
  if (grid > 2048) grid = 2048;
  hipLaunchKernelGGL(scene_expand_kernel, dim3(grid), dim3(256), 0, s, inst, n, arenaCmds, arenaPolys, cmds,
                     polys);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int scene_launch_clear(const ScenePools& p, hipStream_t s) {
  hipLaunchKernelGGL(scene_clear_kernel, dim3(1), dim3(256), 0, s, p.counters, (int)kCounterInts,
                     p.queues, kQueueSets * kQueues * kQueueStride);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int scene_launch_geometry(const ScenePools& p, hipStream_t s) {
  if (p.nItems <= 0) return 0;
  // calls of more than kBlock commands (they come first in itemOrder):
  // commands flattened by a thread each over the whole device, the rest in
  // 1024-thread workgroups
  if (p.nBig > 0) {
    const int grid = (p.bigCmds + 255) / 256;
    hipLaunchKernelGGL(scene_big_cmd_kernel, dim3(grid), dim3(256), 0, s, p, 0);
    // point layout: subpath chunks over the device
    hipLaunchKernelGGL(scene_big_plan_kernel, dim3(1), dim3(kScanBlock), 0, s, p, 2);
    hipLaunchKernelGGL(scene_big_front_kernel, dim3(kBigSegGrid), dim3(256), 0, s, p, 0);
    hipLaunchKernelGGL(scene_big_front_alloc_kernel, dim3(p.nBig), dim3(256), 0, s, p);
    hipLaunchKernelGGL(scene_big_front_kernel, dim3(kBigSegGrid), dim3(256), 0, s, p, 1);
    hipLaunchKernelGGL(scene_big_cmd_kernel, dim3(grid), dim3(256), 0, s, p, 1);
    // fills: segments in chunks over the device; strokes: one workgroup
    hipLaunchKernelGGL(scene_big_plan_kernel, dim3(1), dim3(kScanBlock), 0, s, p, 0);
    hipLaunchKernelGGL(scene_big_seg_kernel, dim3(kBigSegGrid), dim3(256), 0, s, p, 0);
    hipLaunchKernelGGL(scene_big_seg_alloc_kernel, dim3(p.nBig), dim3(256), 0, s, p);
    hipLaunchKernelGGL(scene_big_plan_kernel, dim3(1), dim3(kScanBlock), 0, s, p, 1);
    hipLaunchKernelGGL(scene_big_seg_kernel, dim3(kBigSegGrid), dim3(256), 0, s, p, 1);
    hipLaunchKernelGGL(scene_big_band_kernel, dim3(kBigSegGrid), dim3(256), 0, s, p);
    hipLaunchKernelGGL((scene_geom_kernel<kBigGeomBlock, kGeomBack>), dim3(p.nBig), dim3(kBigGeomBlock), 0, s,
                       p, 0);
  }
  if (p.nItems > p.nBig)
    hipLaunchKernelGGL((scene_geom_kernel<kBlock, kGeomAll>), dim3(p.nItems - p.nBig), dim3(kBlock), 0, s, p,
                       p.nBig);
  hipLaunchKernelGGL(scene_plan_kernel, dim3(1), dim3(kScanBlock), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int scene_launch_bands(const ScenePools& p, int grid, hipStream_t s) {
  hipLaunchKernelGGL(scene_band_map_kernel, dim3(grid / 4 > 0 ? grid / 4 : 1), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(scene_band_count_kernel, dim3(grid), dim3(64), 0, s, p);
  hipLaunchKernelGGL(scene_entry_tile_kernel, dim3(kEntryTiles), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(scene_entry_scan_kernel, dim3(1), dim3(kEntryTiles), 0, s, p);
  hipLaunchKernelGGL(scene_entry_apply_kernel, dim3(kEntryTiles), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(scene_band_write_kernel, dim3(grid), dim3(64), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int scene_launch_bins(const ScenePools& p, int grid, hipStream_t s) {
  hipLaunchKernelGGL(scene_task_map_kernel, dim3(grid / 4 > 0 ? grid / 4 : 1), dim3(kTaskMapBlock), 0, s, p);
  hipLaunchKernelGGL(scene_task_scan_kernel, dim3(1), dim3(64), 0, s, p.counters, p.bigClass);
  hipLaunchKernelGGL(scene_task_order_kernel, dim3(grid >= 8 ? grid / 4 : grid), dim3(64), 0, s, p);
  hipLaunchKernelGGL(scene_bin_kernel<false>, dim3(grid), dim3(64), 0, s, p);
  hipLaunchKernelGGL(scene_bin_scan_kernel, dim3(1), dim3(kScanBlock), 0, s, p);
  hipLaunchKernelGGL(scene_bin_kernel<true>, dim3(grid), dim3(64), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int scene_launch_rows(const ScenePools& p, const RowShape& shape, int numCUs, hipStream_t s,
                      hipStream_t aux, hipEvent_t fork, hipEvent_t join) {
  const int gridSmall = scene_queue_grid(numCUs * shape.wavesPerCU);
  ScenePools small = p;
  small.ldsCap = shape.ldsCap;
  const size_t smallLds = row_lds_fixed_bytes(kWinA) + row_cap_bytes(shape.ldsCap);
  if (shape.mode == kRowsSingle) {  // every task; big ones use the row scratch
    hipLaunchKernelGGL(scene_rows_kernel<false>, dim3(gridSmall), dim3(64), smallLds, s, small);
    return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
  }
  const int gridBig = scene_queue_grid(numCUs * shape.bigWaves);
  ScenePools big = p;
  big.ldsCap = shape.bigCap;
  const size_t bigLds = row_lds_fixed_bytes(kWinA) + row_cap_bytes(shape.bigCap);
  // the small-band launch never needs global row scratch (its bands hold
  // fewer entries than its LDS capacity); with no slots, a row that did would
  // report a capacity error instead of reaching another wave's slot
  small.gscratch = nullptr;
  small.gcap = 0;
  if (shape.mode == kRowsSequential) {
    hipLaunchKernelGGL(scene_rows_kernel<true>, dim3(gridBig), dim3(64), bigLds, s, big);
    hipLaunchKernelGGL(scene_rows_kernel<false>, dim3(gridSmall), dim3(64), smallLds, s, small);
    return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
  }
  if (hipEventRecord(fork, s) != hipSuccess || hipStreamWaitEvent(aux, fork, 0) != hipSuccess)
    return kErrDevice;
  hipLaunchKernelGGL(scene_rows_kernel<false>, dim3(gridSmall), dim3(64), smallLds, aux, small);
  hipLaunchKernelGGL(scene_rows_kernel<true>, dim3(gridBig), dim3(64), bigLds, s, big);
  if (hipEventRecord(join, aux) != hipSuccess || hipStreamWaitEvent(s, join, 0) != hipSuccess)
    return kErrDevice;
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int scene_launch_compose(const ScenePools& p, void* canvas, bool a8, unsigned long long* touched,
                         int32_t* errOut, int grid, hipStream_t s) {
  if (a8)
    hipLaunchKernelGGL(scene_compose_kernel<true>, dim3(grid), dim3(64), 0, s, p, canvas, touched,
                       errOut);
  else
    hipLaunchKernelGGL(scene_compose_kernel<false>, dim3(grid), dim3(64), 0, s, p, canvas, touched,
                       errOut);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

}  // namespace pxh
