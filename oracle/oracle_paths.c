/*
 * oracle_paths.c -- pixie v6.0.0 paths.nim restated in C (scalar semantics).
 * TEST INFRASTRUCTURE ONLY (see pixie_oracle.h).
 *
 * All float arithmetic is float32 in pixie's order, built with
 * -ffp-contract=off.  Assumptions about un-vendored dependencies (vmath,
 * bumpy) are marked "parity unpinned".
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"
#include "pixie_oracle.h"

/* ---------------------------------------------------------------------- */
/* errors                                                                 */
/* ---------------------------------------------------------------------- */
static char g_err[256];
void orc_set_error(const char* msg) {
  strncpy(g_err, msg, sizeof(g_err) - 1);
  g_err[sizeof(g_err) - 1] = 0;
}
const char* orc_last_error(void) { return g_err; }
void orc_free(void* p) { free(p); }

/* ---------------------------------------------------------------------- */
/* vmath Vec2 (parity of length/normalize unpinned: vmath >= 3.0.0)        */
/* ---------------------------------------------------------------------- */
typedef struct { float x, y; } v2;
static inline v2 V(float x, float y) { v2 r = {x, y}; return r; }
static inline v2 vadd(v2 a, v2 b) { return V(a.x + b.x, a.y + b.y); }
static inline v2 vsub(v2 a, v2 b) { return V(a.x - b.x, a.y - b.y); }
static inline v2 vmul(v2 a, float s) { return V(a.x * s, a.y * s); }
static inline v2 vdiv(v2 a, float s) { return V(a.x / s, a.y / s); }
static inline int veq(v2 a, v2 b) { return a.x == b.x && a.y == b.y; }
static inline float vlensq(v2 a) { return a.x * a.x + a.y * a.y; }
static inline float vlen(v2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
static inline v2 vnorm(v2 a) { return vdiv(a, vlen(a)); }

static const float EPSILON = 0.00031415926535897936f; /* 0.0001 * PI, paths.nim:44 */
static const float PIXEL_ERROR_MARGIN = 0.2f;         /* paths.nim:45 */
static const float PI_F = 3.14159265358979323846f;

/* Mat3 (vmath): m[i][j] with i the column; stored m9[3*i + j].  Mat3 * Vec2
 * = (m00*x + m10*y + m20, m01*x + m11*y + m21).  Parity unpinned (vmath). */
static inline v2 mat_apply(const float* m, v2 v) {
  return V(m[0] * v.x + m[3] * v.y + m[6], m[1] * v.x + m[4] * v.y + m[7]);
}
static int mat_is_identity(const float* m) {
  return m[0] == 1 && m[1] == 0 && m[2] == 0 && m[3] == 0 && m[4] == 1 &&
         m[5] == 0 && m[6] == 0 && m[7] == 0 && m[8] == 1;
}
/* paths.nim:61-66 */
static float pixel_scale(const float* m) {
  return nim_maxf(vlen(V(m[0], m[1])), vlen(V(m[3], m[4])));
}

/* ---------------------------------------------------------------------- */
/* growable arrays                                                        */
/* ---------------------------------------------------------------------- */
typedef struct { v2* p; int n, cap; } vbuf;
static void vb_push(vbuf* b, v2 v) {
  if (b->n == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 16;
    b->p = (v2*)realloc(b->p, sizeof(v2) * (size_t)b->cap);
  }
  b->p[b->n++] = v;
}

typedef struct { vbuf pts; int* ofs; int npoly, capoly; } shapes_t;
static void sh_init(shapes_t* s) {
  memset(s, 0, sizeof(*s));
  s->capoly = 8;
  s->ofs = (int*)malloc(sizeof(int) * 9);
  s->ofs[0] = 0;
}
static void sh_free(shapes_t* s) {
  free(s->pts.p);
  free(s->ofs);
}
static void sh_add(shapes_t* s, const vbuf* poly) {
  for (int i = 0; i < poly->n; i++) vb_push(&s->pts, poly->p[i]);
  if (s->npoly + 1 >= s->capoly) {
    s->capoly *= 2;
    s->ofs = (int*)realloc(s->ofs, sizeof(int) * (size_t)(s->capoly + 1));
  }
  s->npoly++;
  s->ofs[s->npoly] = s->pts.n;
}

/* ---------------------------------------------------------------------- */
/* parsePath  paths.nim:119-262                                           */
/* ---------------------------------------------------------------------- */
static int param_count(int kind) { /* paths.nim:73-81 */
  switch (kind) {
    case ORC_CMD_CLOSE: return 0;
    case ORC_CMD_MOVE: case ORC_CMD_LINE: case ORC_CMD_RMOVE: case ORC_CMD_RLINE:
    case ORC_CMD_TQUAD: case ORC_CMD_RTQUAD: return 2;
    case ORC_CMD_HLINE: case ORC_CMD_VLINE: case ORC_CMD_RHLINE: case ORC_CMD_RVLINE: return 1;
    case ORC_CMD_CUBIC: case ORC_CMD_RCUBIC: return 6;
    case ORC_CMD_SCUBIC: case ORC_CMD_RSCUBIC: case ORC_CMD_QUAD: case ORC_CMD_RQUAD: return 4;
    case ORC_CMD_ARC: case ORC_CMD_RARC: return 7;
  }
  return -1;
}

typedef struct {
  const char* s;
  int p, numberStart, armed, hitDecimal, kind, err;
  float* nums; int nn, capn;
  float* out; int no, capo;
} parser_t;

static void po_push(parser_t* P, float v) {
  if (P->no == P->capo) {
    P->capo = P->capo ? P->capo * 2 : 64;
    P->out = (float*)realloc(P->out, sizeof(float) * (size_t)P->capo);
  }
  P->out[P->no++] = v;
}
static void finish_number(parser_t* P) {
  if (P->numberStart > 0) {
    int len = P->p - P->numberStart;
    char buf[128];
    if (len <= 0 || len >= (int)sizeof(buf)) {
      P->err = 1;
    } else {
      memcpy(buf, P->s + P->numberStart, (size_t)len);
      buf[len] = 0;
      char* end = NULL;
      double d = strtod(buf, &end);
      if (end != buf + len) P->err = 1;
      if (P->nn == P->capn) {
        P->capn = P->capn ? P->capn * 2 : 16;
        P->nums = (float*)realloc(P->nums, sizeof(float) * (size_t)P->capn);
      }
      P->nums[P->nn++] = (float)d;
    }
  }
  P->numberStart = 0;
  P->hitDecimal = 0;
}
static void finish_command(parser_t* P) {
  finish_number(P);
  if (P->armed) {
    int pc = param_count(P->kind);
    if (pc == 0) {
      if (P->nn != 0) P->err = 1;
      po_push(P, (float)P->kind);
    } else {
      if (P->nn % pc != 0) P->err = 1;
      for (int batch = 0; batch < P->nn / pc; batch++) {
        if (batch > 0) {
          if (P->kind == ORC_CMD_MOVE) P->kind = ORC_CMD_LINE;
          else if (P->kind == ORC_CMD_RMOVE) P->kind = ORC_CMD_RLINE;
        }
        po_push(P, (float)P->kind);
        for (int i = 0; i < pc; i++) po_push(P, P->nums[batch * pc + i]);
      }
      P->nn = 0;
    }
  }
  P->armed = 1;
}
static int expects_arc_flag(parser_t* P) {
  if (P->kind != ORC_CMD_ARC && P->kind != ORC_CMD_RARC) return 0;
  int m = P->nn % 7;
  return m == 3 || m == 4;
}

int orc_parse_path(const char* s, float** out, int* n) {
  parser_t P;
  memset(&P, 0, sizeof(P));
  P.s = s;
  int len = (int)strlen(s);
  if (len == 0) { *out = NULL; *n = 0; return ORC_OK; }
  static const char* letters = "mlhvcsqtazMLHVCSQTAZ";
  static const int kinds[] = {
      ORC_CMD_RMOVE, ORC_CMD_RLINE, ORC_CMD_RHLINE, ORC_CMD_RVLINE, ORC_CMD_RCUBIC,
      ORC_CMD_RSCUBIC, ORC_CMD_RQUAD, ORC_CMD_RTQUAD, ORC_CMD_RARC, ORC_CMD_CLOSE,
      ORC_CMD_MOVE, ORC_CMD_LINE, ORC_CMD_HLINE, ORC_CMD_VLINE, ORC_CMD_CUBIC,
      ORC_CMD_SCUBIC, ORC_CMD_QUAD, ORC_CMD_TQUAD, ORC_CMD_ARC, ORC_CMD_CLOSE};
  while (P.p < len) {
    char c = s[P.p];
    const char* lp = c ? strchr(letters, c) : NULL;
    if (lp) {
      finish_command(&P);
      P.kind = kinds[lp - letters];
    } else if (c == '-' || c == '+') {
      if (P.numberStart > 0 && (s[P.p - 1] == 'e' || s[P.p - 1] == 'E')) {
      } else {
        finish_number(&P);
        P.numberStart = P.p;
      }
    } else if (c == '.') {
      if (P.hitDecimal || expects_arc_flag(&P)) finish_number(&P);
      P.hitDecimal = 1;
      if (P.numberStart == 0) P.numberStart = P.p;
    } else if (c == ' ' || c == ',' || c == '\r' || c == '\n' || c == '\t') {
      finish_number(&P);
    } else {
      if (P.numberStart > 0 && expects_arc_flag(&P)) finish_number(&P);
      if (P.p - 1 == P.numberStart && s[P.p - 1] == '0') finish_number(&P);
      if (P.numberStart == 0) P.numberStart = P.p;
    }
    P.p++;
  }
  finish_command(&P);
  free(P.nums);
  if (P.err) {
    free(P.out);
    orc_set_error("Invalid path");
    *out = NULL;
    *n = 0;
    return ORC_ERR_PARSE;
  }
  *out = P.out;
  *n = P.no;
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* commandsToShapes  paths.nim:654-1057                                   */
/* ---------------------------------------------------------------------- */
static void add_segment(vbuf* shape, v2 at, v2 to) { /* 669-675 */
  v2 d = vsub(at, to);
  if (d.x != 0.0f || d.y != 0.0f) {
    if (shape->n == 0 || !veq(shape->p[shape->n - 1], at)) vb_push(shape, at);
    vb_push(shape, to);
  }
}

static v2 cubic_at(v2 at, v2 c1, v2 c2, v2 to, float t) { /* 679-686 */
  float t2 = t * t, t3 = t2 * t;
  v2 r = vmul(at, -t3 + 3.0f * t2 - 3.0f * t + 1.0f);
  r = vadd(r, vmul(c1, 3.0f * t3 - 6.0f * t2 + 3.0f * t));
  r = vadd(r, vmul(c2, -3.0f * t3 + 3.0f * t2));
  r = vadd(r, vmul(to, t3));
  return r;
}
static v2 cubic_deriv(v2 at, v2 c1, v2 c2, v2 to, float t) { /* 688-693 */
  float t2 = t * t;
  v2 r = vmul(at, -3.0f * t2 + 6.0f * t - 3.0f);
  r = vadd(r, vmul(c1, 9.0f * t2 - 12.0f * t + 3.0f));
  r = vadd(r, vmul(c2, -9.0f * t2 + 6.0f * t));
  r = vadd(r, vmul(to, 3.0f * t2));
  return r;
}

static int add_cubic(vbuf* shape, v2 at, v2 c1, v2 c2, v2 to, float errSq) { /* 677-723 */
  float t = 0.0f, step = 1.0f;
  v2 prev = at;
  v2 next = cubic_at(at, c1, c2, to, t + step);
  v2 halfway = cubic_at(at, c1, c2, to, t + step / 2.0f);
  for (;;) {
    if (step <= FLT_EPSILON) {
      orc_set_error("Unable to discretize cubic");
      return ORC_ERR_DISCRETIZE;
    }
    v2 midpoint = vdiv(vadd(prev, next), 2.0f);
    v2 lineTangent = vsub(midpoint, prev);
    v2 curveTangent = cubic_deriv(at, c1, c2, to, t + step / 2.0f);
    v2 curveTangentScaled = vmul(vnorm(curveTangent), vlen(lineTangent));
    float error = vlensq(vsub(midpoint, halfway));
    float errorTangent = vlensq(vsub(lineTangent, curveTangentScaled));
    if (error + errorTangent > errSq) {
      next = halfway;
      halfway = cubic_at(at, c1, c2, to, t + step / 4.0f);
      step /= 2.0f;
    } else {
      add_segment(shape, prev, next);
      t += step;
      if (t == 1.0f) break;
      prev = next;
      step = nim_minf(step * 2.0f, 1.0f - t);
      next = cubic_at(at, c1, c2, to, t + step);
      halfway = cubic_at(at, c1, c2, to, t + step / 2.0f);
    }
  }
  return ORC_OK;
}

static v2 quad_at(v2 at, v2 c, v2 to, float t) { /* 727-731 */
  float t2 = t * t;
  v2 r = vmul(at, t2 - 2.0f * t + 1.0f);
  r = vadd(r, vmul(c, -2.0f * t2 + 2.0f * t));
  r = vadd(r, vmul(to, t2));
  return r;
}

static int add_quadratic(vbuf* shape, v2 at, v2 c, v2 to, float errSq) { /* 725-762 */
  float t = 0.0f, step = 1.0f;
  v2 prev = at;
  v2 next = quad_at(at, c, to, t + step);
  v2 halfway = quad_at(at, c, to, t + step / 2.0f);
  int halfStepping = 0;
  for (;;) {
    if (step <= FLT_EPSILON) {
      orc_set_error("Unable to discretize quadratic");
      return ORC_ERR_DISCRETIZE;
    }
    v2 midpoint = vdiv(vadd(prev, next), 2.0f);
    float error = vlensq(vsub(midpoint, halfway));
    if (error > errSq) {
      next = halfway;
      halfway = quad_at(at, c, to, t + step / 4.0f);
      halfStepping = 1;
      step /= 2.0f;
    } else {
      add_segment(shape, prev, next);
      t += step;
      if (t == 1.0f) break;
      prev = next;
      if (halfStepping) step = nim_minf(step, 1.0f - t);
      else step = nim_minf(step * 2.0f, 1.0f - t);
      next = quad_at(at, c, to, t + step);
      halfway = quad_at(at, c, to, t + step / 2.0f);
    }
  }
  return ORC_OK;
}

/* addArc 764-884.  Uses libm float32 transcendentals; parity unpinned
 * (vmath rotate / Nim float-width rules). */
typedef struct { v2 radii; float rot[9]; v2 center; float theta, delta; } arcp;
static float svg_angle(v2 u, v2 v) {
  float dot = u.x * v.x + u.y * v.y;
  float len = vlen(u) * vlen(v);
  float q = dot / len;
  if (q < -1.0f) q = -1.0f;
  if (q > 1.0f) q = 1.0f;
  float r = acosf(q);
  if ((u.x * v.y - u.y * v.x) < 0.0f) r = -r;
  return r;
}
static arcp arc_params(v2 at, v2 radii, float rotation, int large, int sweep, v2 to) {
  radii = V(fabsf(radii.x), fabsf(radii.y));
  v2 radiiSq = V(radii.x * radii.x, radii.y * radii.y);
  /* `rotation / 180 * PI`: PI is float64, so the product is float64 and
   * narrowed to the float32 `radians`. */
  float radians = (float)((double)(rotation / 180.0f) * 3.141592653589793);
  v2 d = V((at.x - to.x) / 2.0f, (at.y - to.y) / 2.0f);
  float cr_ = cosf(radians), sr_ = sinf(radians);
  v2 p = V(cr_ * d.x + sr_ * d.y, -sr_ * d.x + cr_ * d.y);
  v2 pSq = V(p.x * p.x, p.y * p.y);
  float cr = pSq.x / radiiSq.x + pSq.y / radiiSq.y;
  if (cr > 1.0f) {
    radii = vmul(radii, sqrtf(cr));
    radiiSq = V(radii.x * radii.x, radii.y * radii.y);
  }
  float dq = radiiSq.x * pSq.y + radiiSq.y * pSq.x;
  float pq = (radiiSq.x * radiiSq.y - dq) / dq;
  float q = sqrtf(nim_maxf(0.0f, pq));
  if (large == sweep) q = -q;
  v2 cp = V(q * radii.x * p.y / radii.y, -q * radii.y * p.x / radii.x);
  v2 center = V(cr_ * cp.x - sr_ * cp.y + (at.x + to.x) / 2.0f,
                sr_ * cp.x + cr_ * cp.y + (at.y + to.y) / 2.0f);
  float theta = svg_angle(V(1, 0), V((p.x - cp.x) / radii.x, (p.y - cp.y) / radii.y));
  float delta = svg_angle(V((p.x - cp.x) / radii.x, (p.y - cp.y) / radii.y),
                          V((-p.x - cp.x) / radii.x, (-p.y - cp.y) / radii.y));
  /* float32 `mod` / compare against the float64 PI*2 widen to float64;
   * `+=` / `-=` narrow PI*2 to float32 (Nim's implicit conversions). */
  const double TWO_PI = 6.283185307179586;
  delta = (float)fmod((double)delta, TWO_PI);
  if (sweep && delta < 0.0f) delta += (float)TWO_PI;
  else if (!sweep && delta > 0.0f) delta -= (float)TWO_PI;
  while ((double)delta > TWO_PI) delta -= (float)TWO_PI;
  while ((double)delta < -TWO_PI) delta += (float)TWO_PI;
  arcp a;
  a.radii = radii;
  /* vmath rotate(angle), applied as Mat3 * Vec2, turns by +angle:
   * (cos*x - sin*y, sin*x + cos*y).  Parity unpinned (vmath). */
  float c = cosf(radians), s = sinf(radians);
  float rot[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
  memcpy(a.rot, rot, sizeof(rot));
  a.center = center;
  a.theta = theta;
  a.delta = delta;
  return a;
}
static v2 arc_at(const arcp* a, float ang) {
  v2 r = V(cosf(ang) * a->radii.x, sinf(ang) * a->radii.y);
  r = mat_apply(a->rot, r);
  return vadd(r, a->center);
}
static int add_arc(vbuf* shape, v2 at, v2 radii, float rotation, int large, int sweep,
                   v2 to, float errSq) {
  arcp arc = arc_params(at, radii, rotation, large, sweep, to);
  float t = 0.0f, step = 1.0f;
  v2 prev = at;
  while (t != 1.0f) {
    if (step <= FLT_EPSILON) {
      orc_set_error("Unable to discretize arc");
      return ORC_ERR_DISCRETIZE;
    }
    float aPrev = arc.theta + arc.delta * t;
    float a = arc.theta + arc.delta * (t + step);
    v2 next = arc_at(&arc, a);
    v2 halfway = arc_at(&arc, aPrev + (a - aPrev) / 2.0f);
    v2 midpoint = vdiv(vadd(prev, next), 2.0f);
    float error = vlensq(vsub(midpoint, halfway));
    if (error > errSq) {
      v2 quarterway = arc_at(&arc, aPrev + (a - aPrev) / 4.0f);
      v2 mid2 = vdiv(vadd(prev, halfway), 2.0f);
      float halfwayError = vlensq(vsub(mid2, quarterway));
      if (halfwayError < errSq) {
        add_segment(shape, prev, halfway);
        prev = halfway;
        t += step / 2.0f;
        step = nim_minf(step / 2.0f, 1.0f - t);
      } else {
        step = step / 4.0f;
      }
    } else {
      add_segment(shape, prev, next);
      prev = next;
      t += step;
      step = nim_minf(step * 2.0f, 1.0f - t);
    }
  }
  return ORC_OK;
}

static int commands_to_shapes(const float* c, int n, int closeSubpaths, float pixelScale,
                              shapes_t* out) {
  v2 start = V(0, 0), at = V(0, 0), prevCtrl = V(0, 0), prevCtrl2 = V(0, 0);
  int prevKind = ORC_CMD_MOVE;
  vbuf shape = {0};
  float errSq = powf(PIXEL_ERROR_MARGIN / pixelScale, 2.0f);
  int rc = ORC_OK;
  int i = 0;
  while (i < n && rc == ORC_OK) {
    int kind = (int)c[i];
    i++;
    int pc = param_count(kind);
    if (pc < 0 || i + pc > n) {
      orc_set_error("Invalid path command");
      rc = ORC_ERR_ARG;
      break;
    }
    switch (kind) {
      case ORC_CMD_MOVE:
        if (shape.n > 0) {
          if (closeSubpaths) add_segment(&shape, at, start);
          sh_add(out, &shape);
          shape.n = 0;
        }
        at = V(c[i], c[i + 1]);
        start = at;
        break;
      case ORC_CMD_LINE: {
        v2 to = V(c[i], c[i + 1]);
        add_segment(&shape, at, to);
        at = to;
      } break;
      case ORC_CMD_HLINE: {
        v2 to = V(c[i], at.y);
        add_segment(&shape, at, to);
        at = to;
      } break;
      case ORC_CMD_VLINE: {
        v2 to = V(at.x, c[i]);
        add_segment(&shape, at, to);
        at = to;
      } break;
      case ORC_CMD_CUBIC: {
        v2 c1 = V(c[i], c[i + 1]), c2 = V(c[i + 2], c[i + 3]), to = V(c[i + 4], c[i + 5]);
        rc = add_cubic(&shape, at, c1, c2, to, errSq);
        at = to;
        prevCtrl2 = c2;
      } break;
      case ORC_CMD_SCUBIC: {
        v2 c2 = V(c[i], c[i + 1]), to = V(c[i + 2], c[i + 3]);
        if (prevKind == ORC_CMD_CUBIC || prevKind == ORC_CMD_SCUBIC ||
            prevKind == ORC_CMD_RCUBIC || prevKind == ORC_CMD_RSCUBIC) {
          v2 c1 = vsub(vmul(at, 2.0f), prevCtrl2);
          rc = add_cubic(&shape, at, c1, c2, to, errSq);
        } else {
          rc = add_cubic(&shape, at, at, c2, to, errSq);
        }
        at = to;
        prevCtrl2 = c2;
      } break;
      case ORC_CMD_QUAD: {
        v2 cc = V(c[i], c[i + 1]), to = V(c[i + 2], c[i + 3]);
        rc = add_quadratic(&shape, at, cc, to, errSq);
        at = to;
        prevCtrl = cc;
      } break;
      case ORC_CMD_TQUAD: {
        v2 to = V(c[i], c[i + 1]);
        v2 cc = (prevKind == ORC_CMD_QUAD || prevKind == ORC_CMD_TQUAD ||
                 prevKind == ORC_CMD_RQUAD || prevKind == ORC_CMD_RTQUAD)
                    ? vsub(vmul(at, 2.0f), prevCtrl)
                    : at;
        rc = add_quadratic(&shape, at, cc, to, errSq);
        at = to;
        prevCtrl = cc;
      } break;
      case ORC_CMD_ARC: {
        v2 radii = V(c[i], c[i + 1]);
        v2 to = V(c[i + 5], c[i + 6]);
        rc = add_arc(&shape, at, radii, c[i + 2], c[i + 3] == 1.0f, c[i + 4] == 1.0f, to, errSq);
        at = to;
      } break;
      case ORC_CMD_RMOVE:
        if (shape.n > 0) {
          sh_add(out, &shape);
          shape.n = 0;
        }
        at.x += c[i];
        at.y += c[i + 1];
        start = at;
        break;
      case ORC_CMD_RLINE: {
        v2 to = V(at.x + c[i], at.y + c[i + 1]);
        add_segment(&shape, at, to);
        at = to;
      } break;
      case ORC_CMD_RHLINE: {
        v2 to = V(at.x + c[i], at.y);
        add_segment(&shape, at, to);
        at = to;
      } break;
      case ORC_CMD_RVLINE: {
        v2 to = V(at.x, at.y + c[i]);
        add_segment(&shape, at, to);
        at = to;
      } break;
      case ORC_CMD_RCUBIC: {
        v2 c1 = V(at.x + c[i], at.y + c[i + 1]);
        v2 c2 = V(at.x + c[i + 2], at.y + c[i + 3]);
        v2 to = V(at.x + c[i + 4], at.y + c[i + 5]);
        rc = add_cubic(&shape, at, c1, c2, to, errSq);
        at = to;
        prevCtrl2 = c2;
      } break;
      case ORC_CMD_RSCUBIC: {
        v2 c2 = V(at.x + c[i], at.y + c[i + 1]);
        v2 to = V(at.x + c[i + 2], at.y + c[i + 3]);
        v2 c1 = (prevKind == ORC_CMD_CUBIC || prevKind == ORC_CMD_SCUBIC ||
                 prevKind == ORC_CMD_RCUBIC || prevKind == ORC_CMD_RSCUBIC)
                    ? vsub(vmul(at, 2.0f), prevCtrl2)
                    : at;
        rc = add_cubic(&shape, at, c1, c2, to, errSq);
        at = to;
        prevCtrl2 = c2;
      } break;
      case ORC_CMD_RQUAD: {
        v2 cc = V(at.x + c[i], at.y + c[i + 1]);
        v2 to = V(at.x + c[i + 2], at.y + c[i + 3]);
        rc = add_quadratic(&shape, at, cc, to, errSq);
        at = to;
        prevCtrl = cc;
      } break;
      case ORC_CMD_RTQUAD: {
        v2 to = V(at.x + c[i], at.y + c[i + 1]);
        v2 cc = (prevKind == ORC_CMD_QUAD || prevKind == ORC_CMD_TQUAD ||
                 prevKind == ORC_CMD_RQUAD || prevKind == ORC_CMD_RTQUAD)
                    ? vsub(vmul(at, 2.0f), prevCtrl)
                    : at;
        rc = add_quadratic(&shape, at, cc, to, errSq);
        at = to;
        prevCtrl = cc;
      } break;
      case ORC_CMD_RARC: {
        v2 radii = V(c[i], c[i + 1]);
        v2 to = V(at.x + c[i + 5], at.y + c[i + 6]);
        rc = add_arc(&shape, at, radii, c[i + 2], c[i + 3] == 1.0f, c[i + 4] == 1.0f, to, errSq);
        at = to;
      } break;
      case ORC_CMD_CLOSE:
        if (!veq(at, start)) {
          add_segment(&shape, at, start);
          at = start;
        }
        if (shape.n > 0) {
          sh_add(out, &shape);
          shape.n = 0;
        }
        break;
    }
    i += pc;
    prevKind = kind;
  }
  if (rc == ORC_OK && shape.n > 0) {
    if (closeSubpaths) add_segment(&shape, at, start);
    sh_add(out, &shape);
  }
  free(shape.p);
  return rc;
}

static void transform_shapes(shapes_t* s, const float* m) { /* 1092-1096 */
  if (!m || mat_is_identity(m)) return;
  for (int i = 0; i < s->pts.n; i++) s->pts.p[i] = mat_apply(m, s->pts.p[i]);
}

/* ---------------------------------------------------------------------- */
/* segments, bounds, partitions                                           */
/* ---------------------------------------------------------------------- */
typedef struct { v2 at, to; int16_t w; } seg_t;
typedef struct { seg_t* s; int n, cap; } segbuf;

/* vmath quantize(v, n) = sign(v) * floor(abs(v) / n) * n.  Parity
 * unpinned (vmath); with n = 1/256 it equals trunc(v*256)/256. */
static inline float quantize(float v, float n) {
  float sg = v >= 0.0f ? 1.0f : -1.0f;
  return sg * floorf(fabsf(v) / n) * n;
}

static void shapes_to_segments(const shapes_t* sh, int skipHorizontal, segbuf* out) { /* 1059-1090 */
  const float q = 1.0f / 256.0f;
  for (int k = 0; k < sh->npoly; k++) {
    int a = sh->ofs[k], b = sh->ofs[k + 1];
    if (b <= a) continue;
    v2 vec1 = sh->pts.p[b - 1];
    vec1.y = quantize(vec1.y, q);
    for (int i = a; i < b; i++) {
      v2 vec2 = sh->pts.p[i];
      vec2.y = quantize(vec2.y, q);
      if (i == a && veq(vec1, vec2)) continue;
      seg_t s = {vec1, vec2, 1};
      vec1 = vec2;
      if (skipHorizontal && s.at.y == s.to.y) continue;
      if (s.at.y > s.to.y) {
        v2 t = s.at;
        s.at = s.to;
        s.to = t;
        s.w = -1;
      }
      if (out->n == out->cap) {
        out->cap = out->cap ? out->cap * 2 : 64;
        out->s = (seg_t*)realloc(out->s, sizeof(seg_t) * (size_t)out->cap);
      }
      out->s[out->n++] = s;
    }
  }
}

typedef struct { float x, y, w, h; } rect_t;

static rect_t compute_bounds(const segbuf* sg) { /* 1098-1117 */
  float xMin = INFINITY, xMax = -INFINITY, yMin = INFINITY, yMax = -INFINITY;
  for (int i = 0; i < sg->n; i++) {
    const seg_t* s = &sg->s[i];
    xMin = nim_minf(xMin, nim_minf(s->at.x, s->to.x));
    xMax = nim_maxf(xMax, nim_maxf(s->at.x, s->to.x));
    yMin = nim_minf(yMin, s->at.y);
    yMax = nim_maxf(yMax, s->to.y);
  }
  rect_t r = {0, 0, 0, 0};
  if (isnan(xMin) || isnan(xMax) || isnan(yMin) || isnan(yMax)) return r;
  r.x = xMin;
  r.y = yMin;
  r.w = xMax - xMin;
  r.h = yMax - yMin;
  return r;
}

/* No segment (or only NaN ones): computeBounds keeps its infinite initial
 * extremes, whose int conversions in fillShapes are undefined in C.  The
 * reference's tests fill such paths and expect nothing drawn and no error
 * (test_svg.nim: dashes.svg's horizontal lines with the default fill, the
 * tiger's lone moveto; test_fonts.nim: spans6's one-space span). */
static int no_bounds(rect_t r) { return r.x == INFINITY && r.w == -INFINITY; }

static rect_t snap_to_pixels(rect_t r) { /* common.nim:92-101 */
  float xMin = r.x, xMax = r.x + r.w, yMin = r.y, yMax = r.y + r.h;
  rect_t o;
  o.x = floorf(xMin);
  o.w = ceilf(xMax) - o.x;
  o.y = floorf(yMin);
  o.h = ceilf(yMax) - o.y;
  return o;
}

typedef struct {
  v2 at, to; /* possibly clipped to the partition */
  float m, b;
  int16_t w;
} entry_t;

typedef struct {
  entry_t* e;
  int n;
  int aa, two;
  int64_t top, bottom;
} part_t;

static entry_t init_entry(seg_t s) { /* 1127-1135 */
  entry_t e;
  e.at = s.at;
  e.to = s.to;
  e.w = s.w;
  e.m = 0.0f;
  e.b = 0.0f;
  float d = s.at.x - s.to.x;
  if (d == 0.0f) {
    e.b = s.at.x;
  } else {
    e.m = (s.at.y - s.to.y) / d;
    e.b = s.at.y - e.m * s.at.x;
  }
  return e;
}
static inline float solve_x(const entry_t* e, float y) { /* 1137-1141 */
  return e->m == 0.0f ? e->b : (y - e->b) / e->m;
}
static inline float solve_y(const entry_t* e, float x) { return e->m * x + e->b; } /* 1143 */
static inline float midpoint_x(v2 at, v2 to) { return (at.x + to.x) * 0.5f; }

static int seg_requires_aa(v2 at, v2 to) { /* 1149-1160 */
  return at.x != to.x || (at.x - truncf(at.x)) != 0.0f || (at.y - truncf(at.y)) != 0.0f ||
         (to.y - truncf(to.y)) != 0.0f;
}

/* bumpy intersects(Segment, Line) with the line y = ly through (0, ly) and
 * (1000, ly); the point is taken on the segment: u is the segment
 * parameter.  Parity unpinned (bumpy >= 1.1.3). */
static int clip_to_line(v2 sat, v2 sto, float ly, v2* at) {
  v2 s1 = V(1000.0f, 0.0f);
  v2 s2 = vsub(sto, sat);
  float denominator = (-s2.x * s1.y + s1.x * s2.y);
  float numerator = s1.x * (ly - sat.y) - s1.y * (0.0f - sat.x);
  float u = numerator / denominator;
  if (u >= 0.0f && u <= 1.0f) {
    *at = vadd(sat, vmul(s2, u));
    return 1;
  }
  return 0;
}

/* internal.nim:36-49 intersectsInside */
static int intersects_inside(v2 aat, v2 ato, v2 bat, v2 bto) {
  v2 s1 = vsub(ato, aat), s2 = vsub(bto, bat);
  float den = (-s2.x * s1.y + s1.x * s2.y);
  float s = (-s1.y * (aat.x - bat.x) + s1.x * (aat.y - bat.y)) / den;
  float t = (s2.x * (aat.y - bat.y) - s2.y * (aat.x - bat.x)) / den;
  return s > 0.0f && s < 1.0f && t > 0.0f && t < 1.0f;
}

static part_t* partition_segments(const segbuf* sg, int64_t top, int64_t height, int* np_out) { /* 1168-1262 */
  int64_t hd4 = height / 4; /* Nim div truncates */
  uint32_t maxP = (uint32_t)(hd4 > 1 ? hd4 : 1);
  uint32_t s2 = (uint32_t)(sg->n / 2 > 1 ? sg->n / 2 : 1);
  uint32_t nP = maxP < s2 ? maxP : s2;
  part_t* P = (part_t*)calloc(nP, sizeof(part_t));
  uint32_t startY = (uint32_t)top;
  uint32_t pH = (uint32_t)height / nP;
  P[0].top = top;
  P[0].bottom = top + (int64_t)pH;
  for (uint32_t i = 1; i < nP; i++) {
    P[i].top = P[i - 1].bottom;
    P[i].bottom = P[i - 1].bottom + (int64_t)pH;
  }
  P[nP - 1].bottom = top + height;

  if (nP == 1) {
    P[0].e = (entry_t*)malloc(sizeof(entry_t) * (size_t)(sg->n ? sg->n : 1));
    for (int i = 0; i < sg->n; i++) P[0].e[i] = init_entry(sg->s[i]);
    P[0].n = sg->n;
  } else {
    uint32_t* lo = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)sg->n);
    uint32_t* hi = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)sg->n);
    int* cnt = (int*)calloc(nP, sizeof(int));
    for (int i = 0; i < sg->n; i++) {
      uint32_t a = nim_f2u32(nim_maxf(0.0f, sg->s[i].at.y - (float)startY));
      uint32_t b = nim_f2u32(nim_maxf(0.0f, sg->s[i].to.y - (float)startY));
      a /= pH;
      b /= pH;
      if (a > nP - 1) a = nP - 1;
      if (b > nP - 1) b = nP - 1;
      lo[i] = a;
      hi[i] = b;
      for (uint32_t p = a; p <= b; p++) cnt[p]++;
    }
    for (uint32_t p = 0; p < nP; p++) {
      P[p].e = (entry_t*)malloc(sizeof(entry_t) * (size_t)(cnt[p] ? cnt[p] : 1));
      P[p].n = 0;
    }
    for (int i = 0; i < sg->n; i++) {
      entry_t e = init_entry(sg->s[i]);
      for (uint32_t p = lo[i]; p <= hi[i]; p++) P[p].e[P[p].n++] = e;
    }
    free(lo);
    free(hi);
    free(cnt);
  }

  for (uint32_t p = 0; p < nP; p++) {
    part_t* pa = &P[p];
    pa->aa = 0;
    for (int i = 0; i < pa->n; i++)
      if (seg_requires_aa(pa->e[i].at, pa->e[i].to)) { pa->aa = 1; break; }
    float ftop = (float)pa->top, fbot = (float)pa->bottom;
    for (int i = 0; i < pa->n; i++) {
      entry_t* e = &pa->e[i];
      if (e->at.y <= ftop && e->to.y >= fbot) {
        v2 at = V(0, 0);
        clip_to_line(e->at, e->to, ftop, &at);
        e->at = at;
        clip_to_line(e->at, e->to, fbot, &at);
        e->to = at;
      }
    }
    if (pa->n == 2) {
      entry_t* e0 = &pa->e[0];
      entry_t* e1 = &pa->e[1];
      if (!intersects_inside(e0->at, e0->to, e1->at, e1->to)) {
        if (e0->at.y <= ftop && e0->to.y >= fbot && e1->at.y <= ftop && e1->to.y >= fbot) {
          pa->two = 1;
          if (midpoint_x(e0->at, e0->to) > midpoint_x(e1->at, e1->to)) {
            entry_t t = *e0;
            *e0 = *e1;
            *e1 = t;
          }
        }
      }
    }
  }
  *np_out = (int)nP;
  return P;
}

/* ---------------------------------------------------------------------- */
/* scanline driver  paths.nim:1268-1912                                   */
/* ---------------------------------------------------------------------- */
typedef struct { int32_t at; int16_t w; } hit_t;

static inline int32_t fixed32(float f) { return nim_f2i32(f * 256.0f); } /* 1268 */
static inline int32_t fx_int(int32_t p) { return p / 256; }            /* 1271 */
static inline int32_t fx_trunc(int32_t p) { return (p / 256) * 256; }  /* 1274 */

static void sort_hits(hit_t* h, int n) { /* 1277-1286 insertion sort, stable */
  for (int i = 1; i < n; i++) {
    int j = i - 1, k = i;
    while (j >= 0 && h[j].at > h[k].at) {
      hit_t t = h[j + 1];
      h[j + 1] = h[j];
      h[j] = t;
      j--;
      k--;
    }
  }
}

static inline int should_fill(int rule, int64_t count) { /* 1288-1296 */
  return rule == ORC_NONZERO ? count != 0 : (count % 2) != 0;
}

typedef struct { int32_t a, b; int64_t count; } span_t;

/* walk 1298-1334: emits (prevAt, at, count) spans. */
static int walk(const hit_t* h, int n, int rule, span_t* out) {
  int i = 0, no = 0;
  int64_t count = 0;
  int32_t prevAt = 0;
  while (i < n) {
    int32_t at = h[i].at;
    int16_t w = h[i].w;
    if (at > 0) {
      if (should_fill(rule, count)) {
        if (i < n - 1) {
          int32_t nextAt = h[i + 1].at;
          int16_t nextW = h[i + 1].w;
          if (nextAt == at && w + nextW == 0) {
            i += 2;
            continue;
          }
          if (rule == ORC_NONZERO && count + w != 0) {
            count += w;
            i++;
            continue;
          }
        }
        out[no].a = prevAt;
        out[no].b = at;
        out[no].count = count;
        no++;
      }
      prevAt = at;
    }
    count += w;
    i++;
  }
  return no;
}

typedef struct {
  uint8_t* img;
  int W, H;
  orc_rgbx rgbx;
  int mode;
  orc_stats* st;
  int64_t cx0, cx1; /* column clip of non-mask writes */
} canvas_t;

/* Writes of a non-mask fill are confined to the path's pixel span widened by
 * ORC_WRITE_MARGIN columns on each side.  Only trapezoid slivers computed
 * from slope-intercept x can reach outside the span, and only by float
 * error; the clip gives such pathological writes a defined result (pixie
 * would write them) so that fills can be binned by their bounds. */
#define ORC_WRITE_MARGIN 16

static inline int px_clipped(canvas_t* c, int64_t x) {
  if (c->mode == ORC_MASK || (x >= c->cx0 && x < c->cx1)) return 0;
  if (c->st) c->st->dropped++;
  return 1;
}

static inline void px_set(canvas_t* c, int64_t x, int64_t y, orc_rgbx v) {
  if (px_clipped(c, x)) return;
  uint8_t* p = c->img + (y * c->W + x) * 4;
  p[0] = v.r;
  p[1] = v.g;
  p[2] = v.b;
  p[3] = v.a;
}
static inline orc_rgbx px_get(canvas_t* c, int64_t x, int64_t y) {
  uint8_t* p = c->img + (y * c->W + x) * 4;
  orc_rgbx v = {p[0], p[1], p[2], p[3]};
  return v;
}
static inline void px_blend(canvas_t* c, int64_t x, int64_t y, orc_rgbx src) {
  if (px_clipped(c, x)) return;
  px_set(c, x, y, orc_blend(c->mode, px_get(c, x, y), src));
  if (c->st) c->st->touched++;
}
/* clearUnsafe (1433-1440) on one row [x0, x1).  Negative lengths (which
 * make pixie's memset fault) are treated as empty. */
static void clear_row(canvas_t* c, int64_t x0, int64_t x1, int64_t y) {
  if (x0 == c->W || y == c->H) return;
  if (x0 < 0) x0 = 0;
  for (int64_t x = x0; x < x1 && x < c->W; x++) {
    orc_rgbx z = {0, 0, 0, 0};
    px_set(c, x, y, z);
    if (c->st) c->st->touched++;
  }
}

#ifdef ORC_SIMD_MODE
/* ---- AVX2-mode build: pixie's SIMD coverage path (simd/avx2.nim) -------
 * Built into liboracle_avx2.so only (the CPU baseline of pixie's SIMD
 * build; the scalar oracle above is the checker).  It follows the reference
 * where the SIMD path departs from the scalar one:
 *   A1 the coverage multiply of a vector block floors, c * cov / 255 via
 *      mulhi and 0x8081 (avx2.nim:418-434), where the scalar path rounds;
 *   A2 an Overwrite block that is neither all 0 nor all 255 writes every
 *      pixel, zeroing those of coverage 0 (avx2.nim:469-481);
 *   A3 a Mask block of coverage 255 is skipped only for an opaque paint, and
 *      multiplied by the paint's alpha otherwise (avx2.nim:707-724);
 *   A4 trapezoid edge pixels convert c * area with round-half-even
 *      (sse2.nim:6-11) instead of round(area * 255) then (c*x+127) div 255;
 *   A5 each line runs scalar until its pixel address is 32-byte aligned,
 *      for a 32-byte aligned image base (avx2.nim:443-447 and the like).
 * Vector blocks are 32 pixels, taken while i < len - 32, as in the
 * reference.  The reference's Mask loop tests i < len - 16 and can read
 * coverage past the line; this build stops at len - 32 instead. */
#include <immintrin.h>

/* floor(c * f / 255) for each channel of 8 pixels; f8 holds f << 8 in each
 * 16-bit lane */
static inline __m256i mul_div255_8(__m256i px, __m256i f8) {
  const __m256i oddMask = _mm256_set1_epi16((short)0xff00);
  const __m256i div255 = _mm256_set1_epi16((short)0x8081);
  __m256i even = _mm256_slli_epi16(px, 8);
  __m256i odd = _mm256_and_si256(px, oddMask);
  even = _mm256_mulhi_epu16(even, f8);
  odd = _mm256_mulhi_epu16(odd, f8);
  even = _mm256_srli_epi16(_mm256_mulhi_epu16(even, div255), 7);
  odd = _mm256_srli_epi16(_mm256_mulhi_epu16(odd, div255), 7);
  return _mm256_or_si256(even, _mm256_slli_epi16(odd, 8));
}
static inline __m256i lanes_of(__m256i v32) { /* per pixel: v in both 16-bit lanes, << 8 */
  return _mm256_or_si256(_mm256_slli_epi32(v32, 8), _mm256_slli_epi32(v32, 24));
}
static inline __m256i cov8(const uint8_t* cov) {
  return lanes_of(_mm256_cvtepu8_epi32(_mm_loadl_epi64((const __m128i*)cov)));
}
static inline __m256i blend_normal8(__m256i b, __m256i s) {
  const __m256i k = _mm256_sub_epi32(_mm256_set1_epi32(255), _mm256_srli_epi32(s, 24));
  return _mm256_add_epi8(s, mul_div255_8(b, lanes_of(k)));
}
static inline __m256i blend_mask8(__m256i b, __m256i s) {
  return mul_div255_8(b, lanes_of(_mm256_srli_epi32(s, 24)));
}

/* fillCoverage of one antialiased row for Overwrite / Normal / Mask
 * (blendLineCoverage*Avx2, avx2.nim:436-733).  Returns the pixels written. */
static int64_t coverage_line_simd(canvas_t* c, int64_t y, int64_t startX, const uint8_t* cov,
                                  int64_t len, orc_rgbx rgbx, int mode) {
  uint32_t* line = (uint32_t*)(c->img + (y * c->W + startX) * 4);
  const uint32_t rgbxU = (uint32_t)rgbx.r | ((uint32_t)rgbx.g << 8) | ((uint32_t)rgbx.b << 16) |
                         ((uint32_t)rgbx.a << 24);
  int64_t i = 0, touched = 0;
#define SCALAR_PX(k)                                                              \
  do {                                                                            \
    uint8_t cv_ = cov[k];                                                         \
    int64_t x_ = startX + (k);                                                    \
    if (mode == ORC_OVERWRITE) {                                                  \
      if (cv_ != 0) { px_set(c, x_, y, orc_mul_cov(rgbx, cv_)); touched++; }      \
    } else if (mode == ORC_MASK) {                                                \
      if (cv_ != 255) px_blend(c, x_, y, orc_mul_cov(rgbx, cv_));                 \
    } else if (cv_ != 0) {                                                        \
      px_blend(c, x_, y, orc_mul_cov(rgbx, cv_));                                 \
    }                                                                             \
  } while (0)
  while (i < len && ((y * c->W + startX + i) & 7) != 0) {
    SCALAR_PX(i);
    i++;
  }
  const __m256i rgbxVec = _mm256_set1_epi32((int)rgbxU);
  const __m256i zero = _mm256_setzero_si256(), all255 = _mm256_set1_epi8((char)255);
  while (i < len - 32) {
    const __m256i cv = _mm256_loadu_si256((const __m256i*)(cov + i));
    const int allZero = _mm256_movemask_epi8(_mm256_cmpeq_epi8(cv, zero)) == -1;
    const int allFull = _mm256_movemask_epi8(_mm256_cmpeq_epi8(cv, all255)) == -1;
    if (mode == ORC_OVERWRITE) {
      if (!allZero) {
        for (int j = 0; j < 32; j += 8) {
          const __m256i v = allFull ? rgbxVec : mul_div255_8(rgbxVec, cov8(cov + i + j));
          _mm256_storeu_si256((__m256i*)(line + i + j), v);
        }
        touched += 32;
      }
    } else if (mode == ORC_MASK) {
      if (allZero) {
        for (int j = 0; j < 32; j += 8) _mm256_storeu_si256((__m256i*)(line + i + j), zero);
      } else if (!(allFull && rgbx.a == 255)) {
        for (int j = 0; j < 32; j += 8) {
          const __m256i b = _mm256_loadu_si256((const __m256i*)(line + i + j));
          _mm256_storeu_si256((__m256i*)(line + i + j),
                              blend_mask8(b, mul_div255_8(rgbxVec, cov8(cov + i + j))));
        }
      }
      touched += 32;
    } else {
      if (allFull && rgbx.a == 255) {
        for (int j = 0; j < 32; j += 8) _mm256_storeu_si256((__m256i*)(line + i + j), rgbxVec);
        touched += 32;
      } else if (!allZero) {
        for (int j = 0; j < 32; j += 8) {
          const __m256i b = _mm256_loadu_si256((const __m256i*)(line + i + j));
          _mm256_storeu_si256((__m256i*)(line + i + j),
                              blend_normal8(b, mul_div255_8(rgbxVec, cov8(cov + i + j))));
        }
        touched += 32;
      }
    }
    i += 32;
  }
  for (; i < len; i++) SCALAR_PX(i);
#undef SCALAR_PX
  return touched;
}

/* applyOpacity(M128, float32) (sse2.nim:6-11): c * area per channel,
 * converted with round-half-even and saturated to 0..255 */
static orc_rgbx mul_area_simd(orc_rgbx c, float area) {
  float v[4] = {(float)c.r * area, (float)c.g * area, (float)c.b * area, (float)c.a * area};
  uint8_t o[4];
  for (int k = 0; k < 4; k++) {
    float r = rintf(v[k]);
    o[k] = !(r > 0.0f) ? 0 : (r > 255.0f ? 255 : (uint8_t)r);
  }
  orc_rgbx out = {o[0], o[1], o[2], o[3]};
  return out;
}
#define ORC_MUL_AREA(c, a) mul_area_simd((c), (a))
#else
#define ORC_MUL_AREA(c, a) orc_mul_opacity((c), (a))
#endif

/* fillHits 1540-1591 with walkInteger 1336-1348. */
static void fill_hits(canvas_t* c, int64_t startX, int64_t y, const hit_t* hits, int nh,
                      int rule, int maskClears, span_t* tmp) {
  int ns = walk(hits, nh, rule, tmp);
  if (c->mode == ORC_MASK) {
    int64_t filledTo = startX;
    for (int k = 0; k < ns; k++) {
      int64_t fs = fx_int(tmp[k].a), fl = (int64_t)fx_int(tmp[k].b) - fs;
      if (fl <= 0) continue;
      if (maskClears) {
        int64_t gap = fs - filledTo;
        if (gap > 0) clear_row(c, filledTo, fs, y);
      }
      if (c->rgbx.a != 255)
        for (int64_t x = fs; x < fs + fl; x++) px_blend(c, x, y, c->rgbx);
      filledTo = fs + fl;
    }
    if (maskClears) {
      clear_row(c, 0, startX, y);
      clear_row(c, filledTo, c->W, y);
    }
    return;
  }
  for (int k = 0; k < ns; k++) {
    int64_t fs = fx_int(tmp[k].a), fl = (int64_t)fx_int(tmp[k].b) - fs;
    if (fl <= 0) continue;
    for (int64_t x = fs; x < fs + fl; x++) {
      if (c->mode == ORC_OVERWRITE) {
        px_set(c, x, y, c->rgbx);
        if (c->st) c->st->touched++;
      } else {
        px_blend(c, x, y, c->rgbx);
      }
    }
  }
}

static int fill_shapes_impl(uint8_t* img, int W, int H, const shapes_t* sh, orc_rgbx rgbx,
                            int rule, int mode, orc_stats* st) {
  canvas_t cv = {img, W, H, rgbx, mode, st, 0, 0};
  segbuf sg = {0};
  shapes_to_segments(sh, 1, &sg);
  if (st) st->segments += sg.n;
  rect_t raw = compute_bounds(&sg);
  if (no_bounds(raw)) {
    free(sg.s);
    return ORC_OK;
  }
  rect_t bounds = snap_to_pixels(raw);
  int64_t startX = nim_f2i(bounds.x);
  if (startX < 0) startX = 0;
  int64_t startY = nim_f2i(bounds.y);
  if (startY < 0) startY = 0;
  int64_t pathWidth = 0;
  if (startX < W) {
    int64_t bw = nim_f2i(bounds.w);
    pathWidth = bw < W - startX ? bw : W - startX;
  }
  int64_t ph = nim_f2i(bounds.y + bounds.h);
  int64_t pathHeight = H < ph ? H : ph;
  if (pathWidth == 0) {
    free(sg.s);
    return ORC_OK;
  }
  if (pathWidth < 0) {
    free(sg.s);
    orc_set_error("Path int overflow detected");
    return ORC_ERR_INT_OVERFLOW;
  }

  cv.cx0 = startX - ORC_WRITE_MARGIN < 0 ? 0 : startX - ORC_WRITE_MARGIN;
  cv.cx1 = startX + pathWidth + ORC_WRITE_MARGIN > W ? W : startX + pathWidth + ORC_WRITE_MARGIN;
  int nP = 0;
  part_t* P = partition_segments(&sg, startY, pathHeight - startY, &nP);
  if (st) st->partitions += nP;
  int maxE = 0;
  for (int p = 0; p < nP; p++) maxE = P[p].n > maxE ? P[p].n : maxE;
  if (st && maxE > st->max_part) st->max_part = maxE;
  int cap = maxE > 2 ? maxE : 2;
  int* idx = (int*)malloc(sizeof(int) * (size_t)cap);
  v2* trapAt = (v2*)malloc(sizeof(v2) * (size_t)cap);
  v2* trapTo = (v2*)malloc(sizeof(v2) * (size_t)cap);
  hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * (size_t)cap);
  span_t* spans = (span_t*)malloc(sizeof(span_t) * (size_t)cap);
  uint8_t* cov = (uint8_t*)calloc((size_t)pathWidth, 1);

  int pi = 0;
  int64_t y = startY;
  while (y < pathHeight) {
    if (y >= P[pi].bottom) pi++;
    part_t* pa = &P[pi];
    int64_t pHeight = pa->bottom - pa->top;
    if (pHeight == 0) break; /* cannot happen: partitions are >= 1 row */

    if (pa->two && !pa->aa) { /* 1644-1668 */
      int64_t left = nim_f2i(pa->e[0].at.x), right = nim_f2i(pa->e[1].at.x);
      int64_t minX = left < 0 ? 0 : (left > W ? W : left);
      int64_t maxX = right < 0 ? 0 : (right > W ? W : right);
      for (int64_t r = 0; r < pHeight; r++) {
        hits[0].at = (int32_t)(minX * 256);
        hits[0].w = 1;
        hits[1].at = (int32_t)(maxX * 256);
        hits[1].w = -1;
        if (y + r < H) fill_hits(&cv, 0, y + r, hits, 2, ORC_NONZERO, 1, spans);
        if (st) st->rows_span2++;
      }
      y += pHeight;
      continue;
    }

    float scanTop = (float)y, scanBottom = (float)(y + 1);
    int all = 1, ne = 0;
    if (pa->two) {
      ne = 2;
      idx[0] = 0;
      idx[1] = 1;
    } else {
      for (int i = 0; i < pa->n; i++) {
        if (pa->e[i].to.y <= scanTop || pa->e[i].at.y >= scanBottom) continue;
        if (pa->e[i].at.y > scanTop || pa->e[i].to.y < scanBottom) all = 0;
        idx[ne++] = i;
      }
    }

    if (st && ne > st->max_sel) st->max_sel = ne;
    if (all && ne % 2 == 0) { /* trapezoid fast path, 1691-1872 */
      for (int i = 0; i < ne; i++) {
        int k = idx[i];
        trapAt[k] = V(solve_x(&pa->e[k], scanTop), scanTop);
        trapTo[k] = V(solve_x(&pa->e[k], scanBottom), scanBottom);
      }
      for (int i = 1; i < ne; i++) {
        int j = i - 1, k = i;
        while (j >= 0 && midpoint_x(trapAt[idx[j]], trapTo[idx[j]]) >
                             midpoint_x(trapAt[idx[k]], trapTo[idx[k]])) {
          int t = idx[j + 1];
          idx[j + 1] = idx[j];
          idx[j] = t;
          j--;
          k--;
        }
      }
      int noOverlap = 1;
      for (int i = 0; i < ne - 1; i++) {
        float leftMaxX = nim_maxf(trapAt[idx[i]].x, trapTo[idx[i]].x);
        float rightMinX = nim_minf(trapAt[idx[i + 1]].x, trapTo[idx[i + 1]].x);
        if (nim_f2i(ceilf(leftMaxX)) > nim_f2i(rightMinX)) {
          noOverlap = 0;
          break;
        }
      }
      if (noOverlap) {
        int simple = 1;
        int64_t wc = 0;
        for (int i = 0; i < ne; i += 2) {
          wc += pa->e[idx[i]].w;
          if (!should_fill(rule, wc)) { simple = 0; break; }
          wc += pa->e[idx[i + 1]].w;
          if (should_fill(rule, wc)) { simple = 0; break; }
        }
        if (simple) {
          int64_t filledTo = 0;
          for (int i = 0; i < ne; i += 2) {
            const entry_t* left = &pa->e[idx[i]];
            const entry_t* right = &pa->e[idx[i + 1]];
            v2 tlA = trapAt[idx[i]], tlT = trapTo[idx[i]];
            v2 trA = trapAt[idx[i + 1]], trT = trapTo[idx[i + 1]];
            float leftMaxX = nim_maxf(tlA.x, tlT.x);
            float rightMinX = nim_minf(trA.x, trT.x);
            int64_t leftCoverEnd = nim_f2i(ceilf(leftMaxX));
            int64_t rightCoverBegin = nim_f2i(truncf(rightMinX));
            { /* left sliver 1772-1809 */
              int inverted = tlA.x < tlT.x;
              float sliverStart = nim_minf(tlA.x, tlT.x);
              float rectStart = leftMaxX;
              float pen = sliverStart, prevPen = pen;
              float penY = inverted ? (float)y : (float)(y + 1);
              float prevPenY = penY;
              int64_t x0 = nim_f2i(sliverStart), x1 = nim_f2i(ceilf(rectStart));
              for (int64_t x = x0; x < x1; x++) {
                prevPen = pen;
                pen = (float)(x + 1);
                float rightRectArea = 0.0f;
                if (pen > rectStart) {
                  rightRectArea = pen - rectStart;
                  pen = rectStart;
                }
                prevPenY = penY;
                penY = solve_y(left, pen);
                if (x < 0 || x >= W) continue;
                float run = pen - prevPen;
                float triangleArea = 0.5f * run * fabsf(penY - prevPenY);
                float rectArea = inverted ? (prevPenY - (float)y) * run
                                          : ((float)(y + 1) - prevPenY) * run;
                float area = triangleArea + rectArea + rightRectArea;
                px_blend(&cv, x, y, ORC_MUL_AREA(rgbx, area));
              }
            }
            { /* right sliver 1811-1847 */
              int inverted = trA.x > trT.x;
              float rectEnd = rightMinX;
              float sliverEnd = nim_maxf(trA.x, trT.x);
              float pen = rectEnd, prevPen = pen;
              float penY = inverted ? (float)(y + 1) : (float)y;
              float prevPenY = penY;
              int64_t x0 = nim_f2i(rectEnd), x1 = nim_f2i(ceilf(sliverEnd));
              for (int64_t x = x0; x < x1; x++) {
                prevPen = pen;
                pen = (float)(x + 1);
                float a_ = fabsf(prevPen);
                float leftRectArea = a_ - floorf(a_); /* vmath fractional; unpinned */
                if (pen > sliverEnd) pen = sliverEnd;
                prevPenY = penY;
                penY = solve_y(right, pen);
                if (x < 0 || x >= W) continue;
                float run = pen - prevPen;
                float triangleArea = 0.5f * run * fabsf(penY - prevPenY);
                float rectArea = inverted ? (penY - (float)y) * run
                                          : ((float)(y + 1) - penY) * run;
                float area = leftRectArea + triangleArea + rectArea;
                px_blend(&cv, x, y, ORC_MUL_AREA(rgbx, area));
              }
            }
            int64_t fillBegin = leftCoverEnd < 0 ? 0 : (leftCoverEnd > W ? W : leftCoverEnd);
            int64_t fillEnd = rightCoverBegin < 0 ? 0 : (rightCoverBegin > W ? W : rightCoverBegin);
            hit_t h2[2];
            h2[0].at = fixed32((float)fillBegin);
            h2[0].w = 1;
            h2[1].at = fixed32((float)fillEnd);
            h2[1].w = -1;
            fill_hits(&cv, 0, y, h2, 2, ORC_NONZERO, 0, spans);
            if (mode == ORC_MASK) {
              int64_t clearTo = nim_f2i(nim_minf(tlA.x, tlT.x));
              clear_row(&cv, filledTo < W ? filledTo : W, clearTo < W ? clearTo : W, y);
            }
            filledTo = nim_f2i(ceilf(nim_maxf(trA.x, trT.x)));
          }
          if (mode == ORC_MASK) clear_row(&cv, filledTo < W ? filledTo : W, W, y);
          if (st) st->rows_trap++;
          y++;
          continue;
        }
      }
    }

    /* computeCoverage 1350-1431 */
    int aa = pa->aa;
    int quality = aa ? 5 : 1;
    uint8_t sampleCoverage = (uint8_t)(255 / quality);
    float offset = 1.0f / (float)quality;
    float initialOffset = offset / 2.0f + EPSILON;
    float yLine = (float)y + initialOffset - offset;
    int nh = 0;
    for (int m = 0; m < quality; m++) {
      yLine += offset;
      nh = 0;
      for (int i = 0; i < ne; i++) {
        const entry_t* e = &pa->e[idx[i]];
        if (e->at.y <= yLine && e->to.y >= yLine) {
          float x = e->m == 0.0f ? e->b : (yLine - e->b) / e->m;
          hits[nh].at = fixed32(nim_minf(x, (float)W));
          hits[nh].w = e->w;
          nh++;
        }
      }
      if (nh > 0) sort_hits(hits, nh);
      if (aa) {
        int ns = walk(hits, nh, rule, spans);
        for (int k = 0; k < ns; k++) {
          int32_t prevAt = spans[k].a, at = spans[k].b;
          int64_t fillStart = fx_int(prevAt);
          int pixelCrossed = fx_int(at) != fx_int(prevAt);
          int32_t leftCover = pixelCrossed ? fx_trunc(prevAt) + 256 - prevAt : at - prevAt;
          if (leftCover != 0) {
            fillStart++;
            int64_t ix = (int64_t)fx_int(prevAt) - startX;
            if (ix >= 0 && ix < pathWidth)
              cov[ix] = (uint8_t)(cov[ix] + (uint8_t)fx_int(leftCover * (int32_t)sampleCoverage));
            else if (st) st->dropped++;
          }
          if (pixelCrossed) {
            int32_t rightCover = at - fx_trunc(at);
            if (rightCover > 0) {
              int64_t ix = (int64_t)fx_int(at) - startX;
              if (ix >= 0 && ix < pathWidth)
                cov[ix] = (uint8_t)(cov[ix] + (uint8_t)fx_int(rightCover * (int32_t)sampleCoverage));
              else if (st) st->dropped++;
            }
          }
          int64_t fillLen = (int64_t)fx_int(at) - fillStart;
          for (int64_t j = fillStart; j < fillStart + fillLen; j++) {
            int64_t ix = j - startX;
            if (ix >= 0 && ix < pathWidth) cov[ix] = (uint8_t)(cov[ix] + sampleCoverage);
            else if (st) st->dropped++;
          }
        }
      }
    }

    if (aa) { /* fillCoverage 1479-1526 */
#ifdef ORC_SIMD_MODE
      if (mode == ORC_OVERWRITE || mode == ORC_NORMAL || mode == ORC_MASK) {
        int64_t t = coverage_line_simd(&cv, y, startX, cov, pathWidth, rgbx, mode);
        if (st) st->touched += t;
      } else
#endif
      for (int64_t i = 0; i < pathWidth; i++) {
        uint8_t cv_ = cov[i];
        int64_t x = startX + i;
        if (mode == ORC_OVERWRITE) {
          if (cv_ != 0) {
            px_set(&cv, x, y, orc_mul_cov(rgbx, cv_));
            if (st) st->touched++;
          }
        } else if (mode == ORC_MASK) {
          if (cv_ != 255) px_blend(&cv, x, y, orc_mul_cov(rgbx, cv_));
        } else {
          if (cv_ != 0) px_blend(&cv, x, y, orc_mul_cov(rgbx, cv_));
        }
      }
      if (mode == ORC_MASK) {
        clear_row(&cv, 0, startX, y);
        clear_row(&cv, startX + pathWidth, W, y);
      }
      memset(cov, 0, (size_t)pathWidth);
      if (st) st->rows_aa++;
    } else {
      fill_hits(&cv, startX, y, hits, nh, rule, 1, spans);
      if (st) st->rows_nonaa++;
    }
    y++;
  }

  if (mode == ORC_MASK) { /* 1910-1912 */
    for (int64_t r = 0; r < startY && r < H; r++) clear_row(&cv, 0, W, r);
    for (int64_t r = pathHeight < 0 ? 0 : pathHeight; r < H; r++) clear_row(&cv, 0, W, r);
  }

  for (int p = 0; p < nP; p++) free(P[p].e);
  free(P);
  free(idx);
  free(trapAt);
  free(trapTo);
  free(hits);
  free(spans);
  free(cov);
  free(sg.s);
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* strokeShapes  paths.nim:1914-2082                                      */
/* ---------------------------------------------------------------------- */
/* vmath angle(v) = arctan2(v.y, v.x); fixAngle wraps into [-PI, PI].
 * Parity unpinned (vmath). */
static float fix_angle(float a) {
  while (a > PI_F) a -= PI_F * 2.0f;
  while (a < -PI_F) a += PI_F * 2.0f;
  return a;
}
static float vangle(v2 v) { return atan2f(v.y, v.x); }

/* bumpy Line-Line intersection (miter joins).  Parity unpinned. */
static int line_intersects(v2 aa, v2 ab, v2 ba, v2 bb, v2* at) {
  v2 s1 = vsub(ab, aa), s2 = vsub(bb, ba);
  float den = (-s2.x * s1.y + s1.x * s2.y);
  if (den == 0.0f) return 0;
  float t = (s2.x * (aa.y - ba.y) - s2.y * (aa.x - ba.x)) / den;
  *at = vadd(aa, vmul(s1, t));
  return 1;
}

/* ellipse(cx, cy, rx, ry) paths.nim:608-619.  splineCircleK is a float64
 * constant, so the offset coordinates are computed in float64 and narrowed
 * to float32 at the call. */
static const double SPLINE_CIRCLE_K = 4.0 * (-1.0 + 1.4142135623730951) / 3.0;
static int ellipse_cmds(float cx, float cy, float rx, float ry, float* c) {
  double mx = SPLINE_CIRCLE_K * (double)rx, my = SPLINE_CIRCLE_K * (double)ry;
  int n = 0;
  c[n++] = ORC_CMD_MOVE; c[n++] = cx + rx; c[n++] = cy;
  c[n++] = ORC_CMD_CUBIC;
  c[n++] = cx + rx; c[n++] = (float)((double)cy + my);
  c[n++] = (float)((double)cx + mx); c[n++] = cy + ry;
  c[n++] = cx; c[n++] = cy + ry;
  c[n++] = ORC_CMD_CUBIC;
  c[n++] = (float)((double)cx - mx); c[n++] = cy + ry;
  c[n++] = cx - rx; c[n++] = (float)((double)cy + my);
  c[n++] = cx - rx; c[n++] = cy;
  c[n++] = ORC_CMD_CUBIC;
  c[n++] = cx - rx; c[n++] = (float)((double)cy - my);
  c[n++] = (float)((double)cx - mx); c[n++] = cy - ry;
  c[n++] = cx; c[n++] = cy - ry;
  c[n++] = ORC_CMD_CUBIC;
  c[n++] = (float)((double)cx + mx); c[n++] = cy - ry;
  c[n++] = cx + rx; c[n++] = (float)((double)cy - my);
  c[n++] = cx + rx; c[n++] = cy;
  c[n++] = ORC_CMD_CLOSE;
  return n;
}

typedef struct {
  float halfStroke, miterAngleLimit, pixelScale;
  int cap, join;
  shapes_t* out;
  int rc;
} stroker_t;

static void add_circle(stroker_t* S, v2 at) { /* makeCircle 1938-1941 */
  float c[40];
  int n = ellipse_cmds(at.x, at.y, S->halfStroke, S->halfStroke, c);
  shapes_t tmp;
  sh_init(&tmp);
  int rc = commands_to_shapes(c, n, 1, S->pixelScale, &tmp);
  if (rc != ORC_OK) S->rc = rc;
  else if (tmp.npoly > 0) {
    vbuf p = {tmp.pts.p + tmp.ofs[0], tmp.ofs[1] - tmp.ofs[0], 0};
    sh_add(S->out, &p);
  }
  sh_free(&tmp);
}

static void add_poly(stroker_t* S, const v2* pts, int n) {
  vbuf p = {(v2*)pts, n, 0};
  sh_add(S->out, &p);
}

static void add_rect(stroker_t* S, v2 at, v2 to) { /* makeRect 1943-1965 */
  float hs = S->halfStroke;
  v2 tangent = vnorm(vsub(to, at));
  v2 normal = V(tangent.y, tangent.x);
  v2 a = V(at.x + normal.x * hs, at.y - normal.y * hs);
  v2 b = V(to.x + normal.x * hs, to.y - normal.y * hs);
  v2 c = V(to.x - normal.x * hs, to.y + normal.y * hs);
  v2 d = V(at.x - normal.x * hs, at.y + normal.y * hs);
  v2 pts[5] = {a, b, c, d, a};
  add_poly(S, pts, 5);
}

static void add_join(stroker_t* S, v2 prevPos, v2 pos, v2 nextPos) { /* 1967-2013 */
  float minArea = PIXEL_ERROR_MARGIN / S->pixelScale;
  float hs = S->halfStroke;
  if (S->join == ORC_ROUND_JOIN) {
    float area = PI_F * hs * hs;
    if (area > minArea) add_circle(S, pos);
    return;
  }
  float angle = fix_angle(vangle(vsub(nextPos, pos)) - vangle(vsub(prevPos, pos)));
  if (fabsf(fabsf(angle) - PI_F) > EPSILON) {
    v2 a = vmul(vnorm(vsub(pos, prevPos)), hs);
    v2 b = vmul(vnorm(vsub(pos, nextPos)), hs);
    if (angle >= 0.0f) {
      a = V(-a.y, a.x);
      b = V(b.y, -b.x);
    } else {
      a = V(a.y, -a.x);
      b = V(-b.y, b.x);
    }
    int join = S->join;
    if (join == ORC_MITER && fabsf(angle) < S->miterAngleLimit) join = ORC_BEVEL;
    if (join == ORC_MITER) {
      v2 at;
      if (line_intersects(vadd(prevPos, a), vadd(pos, a), vadd(nextPos, b), vadd(pos, b), &at)) {
        float bl = vlensq(vsub(at, pos));
        float areaSq = 0.25f * (vlensq(a) * bl + vlensq(b) * bl);
        if (areaSq > minArea * minArea) {
          v2 pts[5] = {vadd(pos, a), at, vadd(pos, b), pos, vadd(pos, a)};
          add_poly(S, pts, 5);
        }
      }
    } else if (join == ORC_BEVEL) {
      float areaSq = 0.25f * vlensq(a) * vlensq(b);
      if (areaSq > minArea * minArea) {
        v2 pts[4] = {vadd(a, pos), vadd(b, pos), pos, vadd(a, pos)};
        add_poly(S, pts, 4);
      }
    }
  }
}

static int stroke_shapes(const shapes_t* in, float width, int cap, int join, float miterLimit,
                         const float* dashes_in, int ndashes_in, float pixelScale, shapes_t* out) {
  if (width <= 0.0f) return ORC_OK;
  stroker_t S;
  S.halfStroke = width / 2.0f;
  S.miterAngleLimit = asinf(1.0f / miterLimit) * 2.0f;
  S.pixelScale = pixelScale;
  S.cap = cap;
  S.join = join;
  S.out = out;
  S.rc = ORC_OK;
  int nd = ndashes_in;
  float* dashes = (float*)malloc(sizeof(float) * (size_t)(2 * nd + 1));
  for (int i = 0; i < nd; i++) dashes[i] = dashes_in[i];
  for (int k = 0; k < in->npoly && S.rc == ORC_OK; k++) {
    const v2* sh = in->pts.p + in->ofs[k];
    int n = in->ofs[k + 1] - in->ofs[k];
    if (n == 0) continue;
    if (!veq(sh[0], sh[n - 1])) {
      if (cap == ORC_ROUND_CAP) add_circle(&S, sh[0]);
      else if (cap == ORC_SQUARE) {
        v2 tangent = vnorm(vsub(sh[1], sh[0]));
        add_rect(&S, vsub(sh[0], vmul(tangent, S.halfStroke)), sh[0]);
      }
    }
    int ndk = nd;
    if (ndk % 2 != 0) {
      for (int i = 0; i < nd; i++) dashes[nd + i] = dashes[i];
      ndk = 2 * nd;
    }
    for (int i = 0; i < ndk; i++)
      if (dashes[i] <= 0.0f) {
        free(dashes);
        orc_set_error("Invalid line dash value");
        return ORC_ERR_DASH;
      }
    for (int i = 1; i < n; i++) {
      v2 pos = sh[i], prevPos = sh[i - 1];
      if (ndk > 0) {
        float distance = vlen(vsub(prevPos, pos));
        v2 dir = vnorm(vsub(pos, prevPos));
        v2 currPos = prevPos;
        int done = 0;
        while (!done) {
          for (int j = 0; j < ndk; j++) {
            float d = dashes[j];
            if (j % 2 == 0) {
              float dd = nim_minf(distance, d);
              add_rect(&S, currPos, vadd(currPos, vmul(dir, dd)));
            }
            currPos = vadd(currPos, vmul(dir, d));
            distance -= d;
            if (distance <= 0.0f) {
              done = 1;
              break;
            }
          }
        }
      } else {
        add_rect(&S, prevPos, pos);
      }
      if (i < n - 1) add_join(&S, prevPos, pos, sh[i + 1]);
    }
    if (veq(sh[0], sh[n - 1])) {
      if (n >= 2) add_join(&S, sh[n - 2], sh[n - 1], sh[1]);
    } else {
      if (cap == ORC_ROUND_CAP) add_circle(&S, sh[n - 1]);
      else if (cap == ORC_SQUARE) {
        v2 tangent = vnorm(vsub(sh[n - 1], sh[n - 2]));
        add_rect(&S, vadd(sh[n - 1], vmul(tangent, S.halfStroke)), sh[n - 1]);
      }
    }
  }
  free(dashes);
  return S.rc;
}

/* ---------------------------------------------------------------------- */
/* public entry points                                                    */
/* ---------------------------------------------------------------------- */
static const float IDENT[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

static void export_shapes(const shapes_t* s, float** xy, int** ofs, int* np) {
  *np = s->npoly;
  *xy = (float*)malloc(sizeof(float) * 2 * (size_t)(s->pts.n ? s->pts.n : 1));
  for (int i = 0; i < s->pts.n; i++) {
    (*xy)[2 * i] = s->pts.p[i].x;
    (*xy)[2 * i + 1] = s->pts.p[i].y;
  }
  *ofs = (int*)malloc(sizeof(int) * (size_t)(s->npoly + 1));
  for (int i = 0; i <= s->npoly; i++) (*ofs)[i] = s->ofs[i];
}

static void import_shapes(shapes_t* s, const float* xy, const int* ofs, int np) {
  sh_init(s);
  for (int k = 0; k < np; k++) {
    vbuf p = {0};
    for (int i = ofs[k]; i < ofs[k + 1]; i++) vb_push(&p, V(xy[2 * i], xy[2 * i + 1]));
    sh_add(s, &p);
    free(p.p);
  }
}

int orc_flatten(const float* cmds, int n, const float* mat9, int close, float** xy, int** ofs,
                int* np) {
  if (!mat9) mat9 = IDENT;
  shapes_t s;
  sh_init(&s);
  int rc = commands_to_shapes(cmds, n, close, pixel_scale(mat9), &s);
  if (rc == ORC_OK) {
    transform_shapes(&s, mat9);
    export_shapes(&s, xy, ofs, np);
  }
  sh_free(&s);
  return rc;
}

int orc_stroke_shapes(const float* cmds, int n, const float* mat9, float width, int cap,
                      int join, float miterLimit, const float* dashes, int nd, float** xy,
                      int** ofs, int* np) {
  if (!mat9) mat9 = IDENT;
  shapes_t s, o;
  sh_init(&s);
  sh_init(&o);
  float ps = pixel_scale(mat9);
  int rc = commands_to_shapes(cmds, n, 0, ps, &s);
  if (rc == ORC_OK) rc = stroke_shapes(&s, width, cap, join, miterLimit, dashes, nd, ps, &o);
  if (rc == ORC_OK) {
    transform_shapes(&o, mat9);
    export_shapes(&o, xy, ofs, np);
  }
  sh_free(&s);
  sh_free(&o);
  return rc;
}

int orc_fill_shapes(uint8_t* img, int w, int h, const float* xy, const int* ofs, int np,
                    orc_rgbx rgbx, int winding, int blend, orc_stats* st) {
  shapes_t s;
  import_shapes(&s, xy, ofs, np);
  int rc = fill_shapes_impl(img, w, h, &s, rgbx, winding, blend, st);
  sh_free(&s);
  return rc;
}

int orc_fill_path(uint8_t* img, int w, int h, const float* cmds, int n, const float* mat9,
                  orc_rgbx rgbx, int winding, int blend, orc_stats* st) {
  if (!mat9) mat9 = IDENT;
  /* paths.nim:2106-2113: solid paint is skipped when fully transparent
   * unless it overwrites. */
  if (!(rgbx.a > 0 || blend == ORC_OVERWRITE)) return ORC_OK;
  shapes_t s;
  sh_init(&s);
  int rc = commands_to_shapes(cmds, n, 1, pixel_scale(mat9), &s);
  if (rc == ORC_OK) {
    transform_shapes(&s, mat9);
    rc = fill_shapes_impl(img, w, h, &s, rgbx, winding, blend, st);
  }
  sh_free(&s);
  return rc;
}

int orc_stroke_path(uint8_t* img, int w, int h, const float* cmds, int n, const float* mat9,
                    orc_rgbx rgbx, float width, int cap, int join, float miterLimit,
                    const float* dashes, int nd, int blend, orc_stats* st) {
  if (!mat9) mat9 = IDENT;
  if (!(rgbx.a > 0 || blend == ORC_OVERWRITE)) return ORC_OK;
  shapes_t s, o;
  sh_init(&s);
  sh_init(&o);
  float ps = pixel_scale(mat9);
  int rc = commands_to_shapes(cmds, n, 0, ps, &s);
  if (rc == ORC_OK) rc = stroke_shapes(&s, width, cap, join, miterLimit, dashes, nd, ps, &o);
  if (rc == ORC_OK) {
    transform_shapes(&o, mat9);
    rc = fill_shapes_impl(img, w, h, &o, rgbx, ORC_NONZERO, blend, st);
  }
  sh_free(&s);
  sh_free(&o);
  return rc;
}

int orc_segments(const float* xy, const int* ofs, int np, float** segs, int16_t** wind,
                 int* nseg) {
  shapes_t s;
  import_shapes(&s, xy, ofs, np);
  segbuf sg = {0};
  shapes_to_segments(&s, 1, &sg);
  *nseg = sg.n;
  *segs = (float*)malloc(sizeof(float) * 4 * (size_t)(sg.n ? sg.n : 1));
  *wind = (int16_t*)malloc(sizeof(int16_t) * (size_t)(sg.n ? sg.n : 1));
  for (int i = 0; i < sg.n; i++) {
    (*segs)[4 * i] = sg.s[i].at.x;
    (*segs)[4 * i + 1] = sg.s[i].at.y;
    (*segs)[4 * i + 2] = sg.s[i].to.x;
    (*segs)[4 * i + 3] = sg.s[i].to.y;
    (*wind)[i] = sg.s[i].w;
  }
  free(sg.s);
  sh_free(&s);
  return ORC_OK;
}

int orc_partitions(const float* xy, const int* ofs, int np, int W, int H, int** parts,
                   int* nparts, float** entries, int* nentries, int* bounds) {
  shapes_t s;
  import_shapes(&s, xy, ofs, np);
  segbuf sg = {0};
  shapes_to_segments(&s, 1, &sg);
  rect_t raw = compute_bounds(&sg);
  if (no_bounds(raw)) raw.x = raw.y = raw.w = raw.h = 0.0f;
  rect_t b = snap_to_pixels(raw);
  int64_t startX = nim_f2i(b.x);
  if (startX < 0) startX = 0;
  int64_t startY = nim_f2i(b.y);
  if (startY < 0) startY = 0;
  int64_t pathWidth = 0;
  if (startX < W) {
    int64_t bw = nim_f2i(b.w);
    pathWidth = bw < W - startX ? bw : W - startX;
  }
  int64_t ph = nim_f2i(b.y + b.h);
  int64_t pathHeight = H < ph ? H : ph;
  bounds[0] = (int)startX;
  bounds[1] = (int)startY;
  bounds[2] = (int)(pathWidth < INT32_MIN ? INT32_MIN : pathWidth);
  bounds[3] = (int)pathHeight;
  *nparts = 0;
  *nentries = 0;
  *parts = NULL;
  *entries = NULL;
  if (pathWidth <= 0) {
    free(sg.s);
    sh_free(&s);
    return ORC_OK;
  }
  int nP = 0;
  part_t* P = partition_segments(&sg, startY, pathHeight - startY, &nP);
  int tot = 0;
  for (int p = 0; p < nP; p++) tot += P[p].n;
  *parts = (int*)malloc(sizeof(int) * 4 * (size_t)nP);
  *entries = (float*)malloc(sizeof(float) * 7 * (size_t)(tot ? tot : 1));
  int k = 0;
  for (int p = 0; p < nP; p++) {
    (*parts)[4 * p] = (int)P[p].top;
    (*parts)[4 * p + 1] = (int)P[p].bottom;
    (*parts)[4 * p + 2] = P[p].n;
    (*parts)[4 * p + 3] = (P[p].aa ? 1 : 0) | (P[p].two ? 2 : 0);
    for (int i = 0; i < P[p].n; i++, k++) {
      entry_t* e = &P[p].e[i];
      float* o = *entries + 7 * k;
      o[0] = e->at.y;
      o[1] = e->to.y;
      o[2] = e->m;
      o[3] = e->b;
      o[4] = e->at.x;
      o[5] = e->to.x;
      o[6] = (float)e->w;
    }
    free(P[p].e);
  }
  free(P);
  *nparts = nP;
  *nentries = tot;
  free(sg.s);
  sh_free(&s);
  return ORC_OK;
}

/* overlaps 2216-2241 + fillOverlaps 2243-2252.  The scanline-vs-segment
 * intersection here is bumpy intersects(Line, Segment), restated like
 * clip_to_line (parity unpinned). */
int orc_fill_overlaps(const float* cmds, int n, const float* mat9, float tx, float ty,
                      int winding, int* result) {
  if (!mat9) mat9 = IDENT;
  shapes_t s;
  sh_init(&s);
  int rc = commands_to_shapes(cmds, n, 1, pixel_scale(mat9), &s);
  *result = 0;
  if (rc != ORC_OK) {
    sh_free(&s);
    return rc;
  }
  transform_shapes(&s, mat9);
  segbuf sg = {0};
  shapes_to_segments(&s, 1, &sg);
  hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * (size_t)(sg.n ? sg.n : 1));
  int nh = 0;
  for (int i = 0; i < sg.n; i++) {
    seg_t* g = &sg.s[i];
    if (g->at.y <= ty && g->to.y >= ty) {
      v2 at;
      if (clip_to_line(g->at, g->to, ty, &at)) {
        if (!veq(g->to, at)) {
          hits[nh].at = fixed32(at.x);
          hits[nh].w = g->w;
          nh++;
        }
      }
    }
  }
  sort_hits(hits, nh);
  int32_t testX = fixed32(tx);
  int64_t count = 0;
  for (int i = 0; i < nh; i++) {
    if (hits[i].at > testX) {
      *result = should_fill(winding, count);
      break;
    }
    count += hits[i].w;
  }
  free(hits);
  free(sg.s);
  sh_free(&s);
  return ORC_OK;
}
