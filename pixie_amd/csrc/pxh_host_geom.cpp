// pxh_host_geom.cpp -- host geometry of the fill/stroke path (see header).
//
// Float math is float32 in pixie's order (built with -ffp-contract=off);
// where pixie mixes a float32 with the float64 PI the arithmetic is float64
// and narrowed, as Nim's implicit conversions do.
#include "pxh_host_geom.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

namespace pxh {

namespace {

enum Cmd : int32_t {
  kClose = 0, kMove, kLine, kHLine, kVLine, kCubic, kSCubic, kQuad, kTQuad, kArc,
  kRMove, kRLine, kRHLine, kRVLine, kRCubic, kRSCubic, kRQuad, kRTQuad, kRArc
};

int32_t param_count(int32_t k) {  // paths.nim:73-81
  switch (k) {
    case kClose: return 0;
    case kMove: case kLine: case kRMove: case kRLine: case kTQuad: case kRTQuad: return 2;
    case kHLine: case kVLine: case kRHLine: case kRVLine: return 1;
    case kCubic: case kRCubic: return 6;
    case kSCubic: case kRSCubic: case kQuad: case kRQuad: return 4;
    case kArc: case kRArc: return 7;
    default: return -1;
  }
}

const float kEpsilon = 0.00031415926535897936f;  // 0.0001 * PI, paths.nim:44
const float kPixelErrorMargin = 0.2f;            // paths.nim:45
const double kPi = 3.141592653589793;

// ---- parsePath state machine, paths.nim:119-262 ----
struct Parser {
  const char* s;
  int32_t p = 0, numberStart = 0, kind = kClose;
  bool armed = false, hitDecimal = false, bad = false;
  std::vector<float> nums;
  std::vector<float>* out;

  void finish_number() {
    if (numberStart > 0) {
      std::string t(s + numberStart, s + p);
      char* end = nullptr;
      double d = strtod(t.c_str(), &end);
      if (t.empty() || end != t.c_str() + t.size()) bad = true;
      nums.push_back((float)d);
    }
    numberStart = 0;
    hitDecimal = false;
  }
  void finish_command() {
    finish_number();
    if (armed) {
      int32_t pc = param_count(kind);
      if (pc == 0) {
        if (!nums.empty()) bad = true;
        out->push_back((float)kind);
      } else {
        if (nums.size() % pc != 0) bad = true;
        for (size_t batch = 0; batch < nums.size() / pc; batch++) {
          if (batch > 0) {
            if (kind == kMove) kind = kLine;
            else if (kind == kRMove) kind = kRLine;
          }
          out->push_back((float)kind);
          for (int32_t i = 0; i < pc; i++) out->push_back(nums[batch * pc + i]);
        }
        nums.clear();
      }
    }
    armed = true;
  }
  bool expects_arc_flag() const {
    if (kind != kArc && kind != kRArc) return false;
    size_t m = nums.size() % 7;
    return m == 3 || m == 4;
  }
};

int32_t letter_kind(char c) {
  switch (c) {
    case 'm': return kRMove; case 'l': return kRLine; case 'h': return kRHLine;
    case 'v': return kRVLine; case 'c': return kRCubic; case 's': return kRSCubic;
    case 'q': return kRQuad; case 't': return kRTQuad; case 'a': return kRArc;
    case 'z': case 'Z': return kClose;
    case 'M': return kMove; case 'L': return kLine; case 'H': return kHLine;
    case 'V': return kVLine; case 'C': return kCubic; case 'S': return kSCubic;
    case 'Q': return kQuad; case 'T': return kTQuad; case 'A': return kArc;
    default: return -1;
  }
}

// ---- commandsToShapes helpers ----
struct Poly {
  std::vector<V2> p;
  void add_segment(V2 at, V2 to) {  // paths.nim:669-675
    V2 d = at - to;
    if (d.x != 0.0f || d.y != 0.0f) {
      if (p.empty() || !veq(p.back(), at)) p.push_back(at);
      p.push_back(to);
    }
  }
};

struct PolyEmit {
  Poly* poly;
  void operator()(V2 a, V2 b) { poly->add_segment(a, b); }
};

struct ArcParams {
  V2 radii;
  float rot[9];
  V2 center;
  float theta, delta;
};

float svg_angle(V2 u, V2 v) {
  float dot = u.x * v.x + u.y * v.y;
  float l = len(u) * len(v);
  float q = dot / l;
  if (q < -1.0f) q = -1.0f;
  if (q > 1.0f) q = 1.0f;
  float r = acosf(q);
  if ((u.x * v.y - u.y * v.x) < 0.0f) r = -r;
  return r;
}

ArcParams arc_params(V2 at, V2 radii, float rotation, bool large, bool sweep, V2 to) {  // 778-846
  radii = v2(fabsf(radii.x), fabsf(radii.y));
  V2 radiiSq = v2(radii.x * radii.x, radii.y * radii.y);
  float radians = (float)((double)(rotation / 180.0f) * kPi);
  V2 d = v2((at.x - to.x) / 2.0f, (at.y - to.y) / 2.0f);
  float cr_ = cosf(radians), sr_ = sinf(radians);
  V2 p = v2(cr_ * d.x + sr_ * d.y, -sr_ * d.x + cr_ * d.y);
  V2 pSq = v2(p.x * p.x, p.y * p.y);
  float cr = pSq.x / radiiSq.x + pSq.y / radiiSq.y;
  if (cr > 1.0f) {
    radii = radii * sqrtf(cr);
    radiiSq = v2(radii.x * radii.x, radii.y * radii.y);
  }
  float dq = radiiSq.x * pSq.y + radiiSq.y * pSq.x;
  float pq = (radiiSq.x * radiiSq.y - dq) / dq;
  float q = sqrtf(nim_maxf(0.0f, pq));
  if (large == sweep) q = -q;
  V2 cp = v2(q * radii.x * p.y / radii.y, -q * radii.y * p.x / radii.x);
  V2 center = v2(cr_ * cp.x - sr_ * cp.y + (at.x + to.x) / 2.0f,
                 sr_ * cp.x + cr_ * cp.y + (at.y + to.y) / 2.0f);
  float theta = svg_angle(v2(1, 0), v2((p.x - cp.x) / radii.x, (p.y - cp.y) / radii.y));
  float delta = svg_angle(v2((p.x - cp.x) / radii.x, (p.y - cp.y) / radii.y),
                          v2((-p.x - cp.x) / radii.x, (-p.y - cp.y) / radii.y));
  const double kTwoPi = 6.283185307179586;
  delta = (float)fmod((double)delta, kTwoPi);
  if (sweep && delta < 0.0f) delta += (float)kTwoPi;
  else if (!sweep && delta > 0.0f) delta -= (float)kTwoPi;
  while ((double)delta > kTwoPi) delta -= (float)kTwoPi;
  while ((double)delta < -kTwoPi) delta += (float)kTwoPi;
  ArcParams a;
  a.radii = radii;
  float c = cosf(radians), s = sinf(radians);
  float rot[9] = {c, s, 0, -s, c, 0, 0, 0, 1};  // vmath rotate(angle)
  memcpy(a.rot, rot, sizeof(rot));
  a.center = center;
  a.theta = theta;
  a.delta = delta;
  return a;
}

V2 arc_at(const ArcParams& a, float ang) {
  V2 r = v2(cosf(ang) * a.radii.x, sinf(ang) * a.radii.y);
  return mat_apply(a.rot, r) + a.center;
}

int32_t add_arc(Poly* sh, V2 at, V2 radii, float rotation, bool large, bool sweep, V2 to,
                float errSq) {  // 764-884
  ArcParams arc = arc_params(at, radii, rotation, large, sweep, to);
  float t = 0.0f, step = 1.0f;
  V2 prev = at;
  while (t != 1.0f) {
    if (step <= kFltEpsilon) return kErrDiscretize;
    float aPrev = arc.theta + arc.delta * t;
    float a = arc.theta + arc.delta * (t + step);
    V2 next = arc_at(arc, a);
    V2 halfway = arc_at(arc, aPrev + (a - aPrev) / 2.0f);
    V2 midpoint = (prev + next) / 2.0f;
    float error = lensq(midpoint - halfway);
    if (error > errSq) {
      V2 quarterway = arc_at(arc, aPrev + (a - aPrev) / 4.0f);
      V2 mid2 = (prev + halfway) / 2.0f;
      float halfwayError = lensq(mid2 - quarterway);
      if (halfwayError < errSq) {
        sh->add_segment(prev, halfway);
        prev = halfway;
        t += step / 2.0f;
        step = nim_minf(step / 2.0f, 1.0f - t);
      } else {
        step = step / 4.0f;
      }
    } else {
      sh->add_segment(prev, next);
      prev = next;
      t += step;
      step = nim_minf(step * 2.0f, 1.0f - t);
    }
  }
  return kOk;
}

const char* discretize_msg(int32_t kind) {
  switch (kind) {
    case kCubic: case kSCubic: case kRCubic: case kRSCubic: return "Unable to discretize cubic";
    case kQuad: case kTQuad: case kRQuad: case kRTQuad: return "Unable to discretize quadratic";
    default: return "Unable to discretize arc";
  }
}

bool is_cubic(int32_t k) { return k == kCubic || k == kSCubic || k == kRCubic || k == kRSCubic; }
bool is_quad(int32_t k) { return k == kQuad || k == kTQuad || k == kRQuad || k == kRTQuad; }

}  // namespace

int32_t parse_path(const char* s, std::vector<float>* out, std::string* err) {
  out->clear();
  size_t n = strlen(s);
  if (n == 0) return kOk;
  Parser P;
  P.s = s;
  P.out = out;
  while ((size_t)P.p < n) {
    char c = s[P.p];
    int32_t k = letter_kind(c);
    if (k >= 0) {
      P.finish_command();
      P.kind = k;
    } else if (c == '-' || c == '+') {
      if (!(P.numberStart > 0 && (s[P.p - 1] == 'e' || s[P.p - 1] == 'E'))) {
        P.finish_number();
        P.numberStart = P.p;
      }
    } else if (c == '.') {
      if (P.hitDecimal || P.expects_arc_flag()) P.finish_number();
      P.hitDecimal = true;
      if (P.numberStart == 0) P.numberStart = P.p;
    } else if (c == ' ' || c == ',' || c == '\r' || c == '\n' || c == '\t') {
      P.finish_number();
    } else {
      if (P.numberStart > 0 && P.expects_arc_flag()) P.finish_number();
      if (P.p - 1 == P.numberStart && s[P.p - 1] == '0') P.finish_number();
      if (P.numberStart == 0) P.numberStart = P.p;
    }
    P.p++;
  }
  P.finish_command();
  if (P.bad) {
    out->clear();
    *err = "Invalid path";
    return kErrParse;
  }
  return kOk;
}

float pixel_scale(const float* m) {  // paths.nim:61-66
  return nim_maxf(len(v2(m[0], m[1])), len(v2(m[3], m[4])));
}

bool mat_is_identity(const float* m) {
  return m[0] == 1 && m[1] == 0 && m[2] == 0 && m[3] == 0 && m[4] == 1 && m[5] == 0 && m[6] == 0 &&
         m[7] == 0 && m[8] == 1;
}

int32_t commands_to_shapes(const float* c, int32_t n, bool closeSubpaths, float pixelScale,
                           Shapes* out, std::string* err) {  // paths.nim:654-1057
  V2 start = v2(0, 0), at = v2(0, 0), prevCtrl = v2(0, 0), prevCtrl2 = v2(0, 0);
  int32_t prevKind = kMove;
  Poly shape;
  float errSq = powf(kPixelErrorMargin / pixelScale, 2.0f);
  int32_t i = 0;
  while (i < n) {
    int32_t kind = (int32_t)c[i];
    i++;
    int32_t pc = param_count(kind);
    if (pc < 0 || i + pc > n) {
      *err = "Invalid path command";
      return kErrArg;
    }
    int32_t rc = kOk;
    PolyEmit emit{&shape};
    switch (kind) {
      case kMove:
        if (!shape.p.empty()) {
          if (closeSubpaths) shape.add_segment(at, start);
          out->add(shape.p.data(), (int32_t)shape.p.size());
          shape.p.clear();
        }
        at = v2(c[i], c[i + 1]);
        start = at;
        break;
      case kLine: {
        V2 to = v2(c[i], c[i + 1]);
        shape.add_segment(at, to);
        at = to;
      } break;
      case kHLine: {
        V2 to = v2(c[i], at.y);
        shape.add_segment(at, to);
        at = to;
      } break;
      case kVLine: {
        V2 to = v2(at.x, c[i]);
        shape.add_segment(at, to);
        at = to;
      } break;
      case kCubic: {
        V2 c1 = v2(c[i], c[i + 1]), c2 = v2(c[i + 2], c[i + 3]), to = v2(c[i + 4], c[i + 5]);
        rc = flatten_cubic(at, c1, c2, to, errSq, emit);
        at = to;
        prevCtrl2 = c2;
      } break;
      case kSCubic: {
        V2 c2 = v2(c[i], c[i + 1]), to = v2(c[i + 2], c[i + 3]);
        V2 c1 = is_cubic(prevKind) ? at * 2.0f - prevCtrl2 : at;
        rc = flatten_cubic(at, c1, c2, to, errSq, emit);
        at = to;
        prevCtrl2 = c2;
      } break;
      case kQuad: {
        V2 cc = v2(c[i], c[i + 1]), to = v2(c[i + 2], c[i + 3]);
        rc = flatten_quad(at, cc, to, errSq, emit);
        at = to;
        prevCtrl = cc;
      } break;
      case kTQuad: {
        V2 to = v2(c[i], c[i + 1]);
        V2 cc = is_quad(prevKind) ? at * 2.0f - prevCtrl : at;
        rc = flatten_quad(at, cc, to, errSq, emit);
        at = to;
        prevCtrl = cc;
      } break;
      case kArc: {
        V2 radii = v2(c[i], c[i + 1]), to = v2(c[i + 5], c[i + 6]);
        rc = add_arc(&shape, at, radii, c[i + 2], c[i + 3] == 1.0f, c[i + 4] == 1.0f, to, errSq);
        at = to;
      } break;
      case kRMove:
        if (!shape.p.empty()) {
          out->add(shape.p.data(), (int32_t)shape.p.size());
          shape.p.clear();
        }
        at.x += c[i];
        at.y += c[i + 1];
        start = at;
        break;
      case kRLine: {
        V2 to = v2(at.x + c[i], at.y + c[i + 1]);
        shape.add_segment(at, to);
        at = to;
      } break;
      case kRHLine: {
        V2 to = v2(at.x + c[i], at.y);
        shape.add_segment(at, to);
        at = to;
      } break;
      case kRVLine: {
        V2 to = v2(at.x, at.y + c[i]);
        shape.add_segment(at, to);
        at = to;
      } break;
      case kRCubic: {
        V2 c1 = v2(at.x + c[i], at.y + c[i + 1]);
        V2 c2 = v2(at.x + c[i + 2], at.y + c[i + 3]);
        V2 to = v2(at.x + c[i + 4], at.y + c[i + 5]);
        rc = flatten_cubic(at, c1, c2, to, errSq, emit);
        at = to;
        prevCtrl2 = c2;
      } break;
      case kRSCubic: {
        V2 c2 = v2(at.x + c[i], at.y + c[i + 1]);
        V2 to = v2(at.x + c[i + 2], at.y + c[i + 3]);
        V2 c1 = is_cubic(prevKind) ? at * 2.0f - prevCtrl2 : at;
        rc = flatten_cubic(at, c1, c2, to, errSq, emit);
        at = to;
        prevCtrl2 = c2;
      } break;
      case kRQuad: {
        V2 cc = v2(at.x + c[i], at.y + c[i + 1]);
        V2 to = v2(at.x + c[i + 2], at.y + c[i + 3]);
        rc = flatten_quad(at, cc, to, errSq, emit);
        at = to;
        prevCtrl = cc;
      } break;
      case kRTQuad: {
        V2 to = v2(at.x + c[i], at.y + c[i + 1]);
        V2 cc = is_quad(prevKind) ? at * 2.0f - prevCtrl : at;
        rc = flatten_quad(at, cc, to, errSq, emit);
        at = to;
        prevCtrl = cc;
      } break;
      case kRArc: {
        V2 radii = v2(c[i], c[i + 1]);
        V2 to = v2(at.x + c[i + 5], at.y + c[i + 6]);
        rc = add_arc(&shape, at, radii, c[i + 2], c[i + 3] == 1.0f, c[i + 4] == 1.0f, to, errSq);
        at = to;
      } break;
      case kClose:
        if (!veq(at, start)) {
          shape.add_segment(at, start);
          at = start;
        }
        if (!shape.p.empty()) {
          out->add(shape.p.data(), (int32_t)shape.p.size());
          shape.p.clear();
        }
        break;
    }
    if (rc != kOk) {
      *err = discretize_msg(kind);
      return rc;
    }
    i += pc;
    prevKind = kind;
  }
  if (!shape.p.empty()) {
    if (closeSubpaths) shape.add_segment(at, start);
    out->add(shape.p.data(), (int32_t)shape.p.size());
  }
  return kOk;
}

void transform_shapes(Shapes* s, const float* m) {  // paths.nim:1092-1096
  if (!m || mat_is_identity(m)) return;
  for (auto& p : s->pts) p = mat_apply(m, p);
}

// ---- strokeShapes, paths.nim:1914-2082 ----
namespace {

const double kSplineCircleK = 4.0 * (-1.0 + 1.4142135623730951) / 3.0;  // paths.nim:529
const float kPiF = 3.14159265358979323846f;

float fix_angle(float a) {  // vmath fixAngle
  while (a > kPiF) a -= kPiF * 2.0f;
  while (a < -kPiF) a += kPiF * 2.0f;
  return a;
}
float vangle(V2 v) { return atan2f(v.y, v.x); }  // vmath angle

bool line_intersects(V2 aa, V2 ab, V2 ba, V2 bb, V2* at) {  // bumpy Line-Line
  V2 s1 = ab - aa, s2 = bb - ba;
  float den = (-s2.x * s1.y + s1.x * s2.y);
  if (den == 0.0f) return false;
  float t = (s2.x * (aa.y - ba.y) - s2.y * (aa.x - ba.x)) / den;
  *at = aa + s1 * t;
  return true;
}

struct Stroker {
  float halfStroke, miterAngleLimit, pixelScale;
  int32_t cap, join;
  Shapes* out;
  int32_t rc = kOk;
  std::string err;

  void add_poly(const V2* p, int32_t n) { out->add(p, n); }

  void add_circle(V2 at) {  // makeCircle 1938-1941: ellipse path, first shape
    float cx = at.x, cy = at.y, rx = halfStroke, ry = halfStroke;
    double mx = kSplineCircleK * (double)rx, my = kSplineCircleK * (double)ry;
    float c[32];
    int n = 0;
    c[n++] = kMove; c[n++] = cx + rx; c[n++] = cy;
    c[n++] = kCubic;
    c[n++] = cx + rx; c[n++] = (float)((double)cy + my);
    c[n++] = (float)((double)cx + mx); c[n++] = cy + ry;
    c[n++] = cx; c[n++] = cy + ry;
    c[n++] = kCubic;
    c[n++] = (float)((double)cx - mx); c[n++] = cy + ry;
    c[n++] = cx - rx; c[n++] = (float)((double)cy + my);
    c[n++] = cx - rx; c[n++] = cy;
    c[n++] = kCubic;
    c[n++] = cx - rx; c[n++] = (float)((double)cy - my);
    c[n++] = (float)((double)cx - mx); c[n++] = cy - ry;
    c[n++] = cx; c[n++] = cy - ry;
    c[n++] = kCubic;
    c[n++] = (float)((double)cx + mx); c[n++] = cy - ry;
    c[n++] = cx + rx; c[n++] = (float)((double)cy - my);
    c[n++] = cx + rx; c[n++] = cy;
    c[n++] = kClose;
    Shapes tmp;
    int32_t r = commands_to_shapes(c, n, true, pixelScale, &tmp, &err);
    if (r != kOk) {
      rc = r;
      return;
    }
    if (tmp.count() > 0) add_poly(tmp.pts.data(), tmp.ofs[1] - tmp.ofs[0]);
  }

  void add_rect(V2 at, V2 to) {  // makeRect 1943-1965
    float hs = halfStroke;
    V2 tangent = normalize(to - at);
    V2 normal = v2(tangent.y, tangent.x);
    V2 a = v2(at.x + normal.x * hs, at.y - normal.y * hs);
    V2 b = v2(to.x + normal.x * hs, to.y - normal.y * hs);
    V2 c = v2(to.x - normal.x * hs, to.y + normal.y * hs);
    V2 d = v2(at.x - normal.x * hs, at.y + normal.y * hs);
    V2 pts[5] = {a, b, c, d, a};
    add_poly(pts, 5);
  }

  void add_join(V2 prevPos, V2 pos, V2 nextPos) {  // 1967-2013
    float minArea = kPixelErrorMargin / pixelScale;
    float hs = halfStroke;
    if (join == 1) {  // RoundJoin
      float area = kPiF * hs * hs;
      if (area > minArea) add_circle(pos);
      return;
    }
    float angle = fix_angle(vangle(nextPos - pos) - vangle(prevPos - pos));
    if (fabsf(fabsf(angle) - kPiF) > kEpsilon) {
      V2 a = normalize(pos - prevPos) * hs;
      V2 b = normalize(pos - nextPos) * hs;
      if (angle >= 0.0f) {
        a = v2(-a.y, a.x);
        b = v2(b.y, -b.x);
      } else {
        a = v2(a.y, -a.x);
        b = v2(-b.y, b.x);
      }
      int32_t j = join;
      if (j == 0 && fabsf(angle) < miterAngleLimit) j = 2;
      if (j == 0) {
        V2 at;
        if (line_intersects(prevPos + a, pos + a, nextPos + b, pos + b, &at)) {
          float bl = lensq(at - pos);
          float areaSq = 0.25f * (lensq(a) * bl + lensq(b) * bl);
          if (areaSq > minArea * minArea) {
            V2 pts[5] = {pos + a, at, pos + b, pos, pos + a};
            add_poly(pts, 5);
          }
        }
      } else if (j == 2) {
        float areaSq = 0.25f * lensq(a) * lensq(b);
        if (areaSq > minArea * minArea) {
          V2 pts[4] = {a + pos, b + pos, pos, a + pos};
          add_poly(pts, 4);
        }
      }
    }
  }
};

}  // namespace

int32_t stroke_shapes(const Shapes& in, const StrokeParams& sp, float pixelScale, Shapes* out,
                      std::string* err) {
  if (sp.width <= 0.0f) return kOk;
  Stroker S;
  S.halfStroke = sp.width / 2.0f;
  S.miterAngleLimit = asinf(1.0f / sp.miterLimit) * 2.0f;
  S.pixelScale = pixelScale;
  S.cap = sp.cap;
  S.join = sp.join;
  S.out = out;
  std::vector<float> dashes(sp.dashes, sp.dashes + sp.ndashes);
  if (dashes.size() % 2 != 0) {
    std::vector<float> d2 = dashes;
    dashes.insert(dashes.end(), d2.begin(), d2.end());
  }
  for (int32_t k = 0; k < in.count() && S.rc == kOk; k++) {
    const V2* sh = in.pts.data() + in.ofs[k];
    int32_t n = in.ofs[k + 1] - in.ofs[k];
    if (n == 0) continue;
    if (!veq(sh[0], sh[n - 1])) {
      if (sp.cap == 1) S.add_circle(sh[0]);
      else if (sp.cap == 2) {
        V2 tangent = normalize(sh[1] - sh[0]);
        S.add_rect(sh[0] - tangent * S.halfStroke, sh[0]);
      }
    }
    for (float d : dashes)
      if (d <= 0.0f) {
        *err = "Invalid line dash value";
        return kErrDash;
      }
    for (int32_t i = 1; i < n; i++) {
      V2 pos = sh[i], prevPos = sh[i - 1];
      if (!dashes.empty()) {
        float distance = len(prevPos - pos);
        V2 dir = normalize(pos - prevPos);
        V2 currPos = prevPos;
        bool done = false;
        for (int pass = 0; !done; pass++) {
          if (pass == kMaxDashPasses) {
            *err = "dash pattern too fine for a segment";
            return kErrCapacity;
          }
          for (size_t j = 0; j < dashes.size(); j++) {
            float d = dashes[j];
            if (j % 2 == 0) {
              float dd = nim_minf(distance, d);
              S.add_rect(currPos, currPos + dir * dd);
            }
            currPos = currPos + dir * d;
            distance -= d;
            if (distance <= 0.0f) {
              done = true;
              break;
            }
          }
        }
      } else {
        S.add_rect(prevPos, pos);
      }
      if (i < n - 1) S.add_join(prevPos, pos, sh[i + 1]);
    }
    if (veq(sh[0], sh[n - 1])) {
      if (n >= 2) S.add_join(sh[n - 2], sh[n - 1], sh[1]);
    } else {
      if (sp.cap == 1) S.add_circle(sh[n - 1]);
      else if (sp.cap == 2) {
        V2 tangent = normalize(sh[n - 1] - sh[n - 2]);
        S.add_rect(sh[n - 1] + tangent * S.halfStroke, sh[n - 1]);
      }
    }
  }
  if (S.rc != kOk) *err = S.err;
  return S.rc;
}

// ---- shapesToSegments + bounds + partitionSegments ----
int32_t build_fill(const Shapes& s, int32_t W, int32_t H, uint32_t rgbx, int32_t mode,
                   int32_t rule, FillGeom* g, std::string* err) {
  std::vector<Seg> segs;
  segs.reserve(s.pts.size());
  Bounds bd = bounds_init();
  for (int32_t k = 0; k < s.count(); k++) {  // paths.nim:1059-1090
    int32_t a = s.ofs[k], b = s.ofs[k + 1];
    if (b <= a) continue;
    V2 vec1 = s.pts[b - 1];
    vec1.y = quantize_y(vec1.y);
    for (int32_t i = a; i < b; i++) {
      V2 vec2 = s.pts[i];
      vec2.y = quantize_y(vec2.y);
      if (i == a && veq(vec1, vec2)) continue;
      Seg sg;
      bool ok = make_segment(vec1, vec2, &sg);
      vec1 = vec2;
      if (!ok) continue;
      segs.push_back(sg);
      bounds_add(&bd, sg);
    }
  }
  FillInfo& f = g->info;
  memset(&f, 0, sizeof(f));
  f.rgbx = rgbx;
  f.mode = mode;
  f.rule = rule;
  int32_t rc = fill_locals(bd, false, W, H, &f);
  g->parts.clear();
  g->entries.clear();
  g->windings.clear();
  g->maxCount = 0;
  if (rc != kOk) {
    *err = "Path int overflow detected";
    f.status = rc;
    return rc;
  }
  if (f.pathWidth == 0) return kOk;  // nothing to draw (paths.nim:1615-1616)

  band_layout(&f, (int64_t)segs.size());
  const uint32_t nP = (uint32_t)f.nParts;
  const int64_t top = f.startY;
  std::vector<uint32_t> lo(segs.size()), hi(segs.size());
  std::vector<int32_t> cnt(nP, 0);
  for (size_t i = 0; i < segs.size(); i++) {
    if (nP == 1) {
      lo[i] = hi[i] = 0;
    } else {
      partition_range(segs[i].at.y, segs[i].to.y, (uint32_t)top, (uint32_t)f.bandH, nP, &lo[i],
                      &hi[i]);
    }
    for (uint32_t p = lo[i]; p <= hi[i]; p++) cnt[p]++;
  }
  g->parts.resize(nP);
  int32_t base = 0;
  for (uint32_t p = 0; p < nP; p++) {
    g->parts[p].entryBase = base;
    g->parts[p].count = cnt[p];
    g->parts[p].flags = 0;
    g->parts[p].pad = 0;
    base += cnt[p];
    g->maxCount = std::max(g->maxCount, cnt[p]);
  }
  g->entries.resize(base);
  g->windings.resize(base);
  // unclipped segments per entry, for the AA test and the two-spanning test
  std::vector<Seg> eseg(base);
  std::vector<int32_t> fillPos(nP, 0);
  for (size_t i = 0; i < segs.size(); i++) {
    Entry e = init_entry(segs[i].at, segs[i].to);
    for (uint32_t p = lo[i]; p <= hi[i]; p++) {
      int32_t k = g->parts[p].entryBase + fillPos[p]++;
      g->entries[k] = e;
      g->windings[k] = (int8_t)segs[i].w;
      eseg[k] = segs[i];
    }
  }
  for (uint32_t p = 0; p < nP; p++) {
    PartInfo& P = g->parts[p];
    float ftop = (float)(top + (int64_t)p * f.bandH);
    float fbot = (p == nP - 1) ? (float)f.pathHeight : (float)(top + (int64_t)(p + 1) * f.bandH);
    bool aa = false;
    for (int32_t i = 0; i < P.count; i++) {
      const Seg& sg = eseg[P.entryBase + i];
      if (seg_requires_aa(sg.at, sg.to)) {
        aa = true;
        break;
      }
    }
    // clip entries spanning the band to its top / bottom lines
    for (int32_t i = 0; i < P.count; i++) {
      Seg& sg = eseg[P.entryBase + i];
      if (sg.at.y <= ftop && sg.to.y >= fbot) {
        V2 at = v2(0, 0);
        clip_to_line(sg.at, sg.to, ftop, &at);
        sg.at = at;
        clip_to_line(sg.at, sg.to, fbot, &at);
        sg.to = at;
        g->entries[P.entryBase + i].aty = sg.at.y;
        g->entries[P.entryBase + i].toy = sg.to.y;
      }
    }
    bool two = false;
    if (P.count == 2) {
      Seg& s0 = eseg[P.entryBase];
      Seg& s1 = eseg[P.entryBase + 1];
      if (!intersects_inside(s0.at, s0.to, s1.at, s1.to)) {
        if (s0.at.y <= ftop && s0.to.y >= fbot && s1.at.y <= ftop && s1.to.y >= fbot) {
          two = true;
          if (midpoint_x(s0.at, s0.to) > midpoint_x(s1.at, s1.to)) {
            std::swap(g->entries[P.entryBase], g->entries[P.entryBase + 1]);
            std::swap(g->windings[P.entryBase], g->windings[P.entryBase + 1]);
            std::swap(s0, s1);
          }
        }
      }
    }
    P.flags = (aa ? 1 : 0) | (two ? 2 : 0);
  }
  return kOk;
}

}  // namespace pxh
