/*
 * oracle_paints.c -- gradient paints of pixie v6.0.0 and fillPath's route for
 * non-solid paints.  TEST INFRASTRUCTURE ONLY (see pixie_oracle.h).
 *
 * paints.nim:72-248 (gradientColor, fillGradientLinear / Radial / Angular)
 * and paths.nim:2115-2142 (mask = fillPath(white); fill = gradient;
 * fill.draw(mask, MaskBlend); image.draw(fill, blendMode)).  The vmath and
 * chroma helpers they call (Mat3 products and inverse, angle, fixAngle,
 * mix, Color -> ColorRGBX) are restated here in float32; those two
 * dependencies are not vendored, so the gradient masters
 * (tests/paths/gradient*.png) are what pins this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"
#include "pixie_oracle.h"

typedef struct { float x, y; } pv2;

static int orc_fail(int code, const char* msg) {
  orc_set_error(msg);
  return code;
}

static pv2 pv(float x, float y) { pv2 r = {x, y}; return r; }
static float plen(pv2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
static pv2 pnormalize(pv2 v) {
  float l = plen(v);
  return pv(v.x / l, v.y / l);
}
static float pangle(pv2 v) { return atan2f(v.y, v.x); }  /* vmath angle */
static float pfix_angle(float a) {                        /* vmath fixAngle */
  const float PI_F = 3.14159265358979323846f;
  while (a > PI_F) a -= PI_F * 2.0f;
  while (a < -PI_F) a += PI_F * 2.0f;
  return a;
}

/* Mat3 with m[3 * col + row], as the matrices of the C ABI */
static void m3_mul(const float* a, const float* b, float* o) {
  float r[9];
  for (int c = 0; c < 3; c++)
    for (int rr = 0; rr < 3; rr++)
      r[3 * c + rr] = a[0 * 3 + rr] * b[3 * c + 0] + a[1 * 3 + rr] * b[3 * c + 1] + a[2 * 3 + rr] * b[3 * c + 2];
  memcpy(o, r, sizeof(r));
}
static void m3_inverse(const float* m, float* o) {
  /* cofactor expansion, m[col][row] */
  const float a = m[0], b = m[3], c = m[6];
  const float d = m[1], e = m[4], f = m[7];
  const float g = m[2], h = m[5], i = m[8];
  const float A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const float det = a * A + b * B + c * C;
  const float inv = 1.0f / det;
  float r[9];
  r[0] = A * inv;
  r[1] = B * inv;
  r[2] = C * inv;
  r[3] = -(b * i - c * h) * inv;
  r[4] = (a * i - c * g) * inv;
  r[5] = -(a * h - b * g) * inv;
  r[6] = (b * f - c * e) * inv;
  r[7] = -(a * f - c * d) * inv;
  r[8] = (a * e - b * d) * inv;
  memcpy(o, r, sizeof(r));
}
static pv2 m3_apply(const float* m, pv2 p) {
  return pv(m[0] * p.x + m[3] * p.y + m[6], m[1] * p.x + m[4] * p.y + m[7]);
}

/* gradientColor (paints.nim:72-97): the last stop below t and the next one,
 * mixed linearly in straight float colour; alpha times the opacity. */
static orc_rgbx gradient_color(const orc_paint* p, float t) {
  int index = -1;
  for (int i = 0; i < p->nstops; i++) {
    if (p->stops[i].pos < t) index = i;
    if (p->stops[i].pos > t) break;
  }
  orc_stop c;
  if (index == -1) {
    c = p->stops[0];
  } else if (index + 1 >= p->nstops) {
    c = p->stops[index];
  } else {
    const orc_stop s1 = p->stops[index], s2 = p->stops[index + 1];
    const float v = (t - s1.pos) / (s2.pos - s1.pos);
    c.r = s1.r * (1.0f - v) + s2.r * v;
    c.g = s1.g * (1.0f - v) + s2.g * v;
    c.b = s1.b * (1.0f - v) + s2.b * v;
    c.a = s1.a * (1.0f - v) + s2.a * v;
  }
  return orc_rgbx_from_color(c.r, c.g, c.b, c.a * p->opacity);
}

/* fillGradient (paints.nim:99-248) with the paint's opacity (fillPath calls
 * it at opacity 1 and applies the opacity to the mask instead). */
int orc_fill_gradient(uint8_t* img, int w, int h, const orc_paint* p) {
  if (p->nstops == 0) return orc_fail(ORC_ERR_ARG, "Gradient must have at least 1 color stop");
  orc_paint q = *p;
  q.opacity = q.opacity < 0 ? 0 : (q.opacity > 1 ? 1 : q.opacity);
  if (q.opacity == 0) return ORC_OK;
  const float PI_F = 3.14159265358979323846f;
  if (p->kind == ORC_LINEAR_GRADIENT) {
    const pv2 at = pv(p->h[0], p->h[1]), to = pv(p->h[2], p->h[3]);
    const pv2 d = pv(to.x - at.x, to.y - at.y);
    const float det = d.x * d.x + d.y * d.y;
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        /* toLineSpace; the horizontal and vertical cases evaluate it at
         * (x, 0) and (0, y), which gives the same value */
        const pv2 pt = pv((float)x, (float)y);
        const float t = (d.y * (pt.y - at.y) + d.x * (pt.x - at.x)) / det;
        orc_rgbx c = gradient_color(&q, t);
        uint8_t* o = img + ((size_t)y * w + x) * 4;
        o[0] = c.r; o[1] = c.g; o[2] = c.b; o[3] = c.a;
      }
  } else if (p->kind == ORC_RADIAL_GRADIENT) {
    float m[9];
    orc_radial_matrix(p, m);
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const float t = plen(m3_apply(m, pv((float)x, (float)y)));
        orc_rgbx c = gradient_color(&q, t);
        uint8_t* o = img + ((size_t)y * w + x) * 4;
        o[0] = c.r; o[1] = c.g; o[2] = c.b; o[3] = c.a;
      }
  } else if (p->kind == ORC_ANGULAR_GRADIENT) {
    const pv2 center = pv(p->h[0], p->h[1]), edge = pv(p->h[2], p->h[3]);
    const float ga = pfix_angle(pangle(pnormalize(pv(edge.x - center.x, edge.y - center.y))));
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const float angle = pangle(pnormalize(pv((float)x - center.x, (float)y - center.y)));
        const float t = pfix_angle(angle + ga + PI_F / 2.0f) / 2.0f / PI_F + 0.5f;
        orc_rgbx c = gradient_color(&q, t);
        uint8_t* o = img + ((size_t)y * w + x) * 4;
        o[0] = c.r; o[1] = c.g; o[2] = c.b; o[3] = c.a;
      }
  } else {
    return orc_fail(ORC_ERR_ARG, "Paint must be a gradient");
  }
  return ORC_OK;
}

/* The radial gradient's matrix, (translate(center) * rotate(angle) *
 * scale(distanceX, distanceY)).inverse() (paints.nim:185-195). */
void orc_radial_matrix(const orc_paint* p, float* m) {
  const pv2 center = pv(p->h[0], p->h[1]), edge = pv(p->h[2], p->h[3]), skew = pv(p->h[4], p->h[5]);
  const float dx = plen(pv(center.x - edge.x, center.y - edge.y));
  const float dy = plen(pv(center.x - skew.x, center.y - skew.y));
  const float ga = pfix_angle(pangle(pnormalize(pv(center.x - edge.x, center.y - edge.y))));
  const float c = cosf(ga), s = sinf(ga);
  const float T[9] = {1, 0, 0, 0, 1, 0, center.x, center.y, 1};
  const float R[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
  const float S[9] = {dx, 0, 0, 0, dy, 0, 0, 0, 1};
  float TR[9], TRS[9];
  m3_mul(T, R, TR);
  m3_mul(TR, S, TRS);
  m3_inverse(TRS, m);
}

/* fillPath with a gradient paint (paths.nim:2115-2142). */
int orc_fill_path_paint(uint8_t* img, int w, int h, const float* cmds, int n, const float* mat9,
                        const orc_paint* p, int winding, orc_stats* st) {
  float opacity = p->opacity < 0 ? 0 : (p->opacity > 1 ? 1 : p->opacity);
  if (opacity == 0) return ORC_OK;
  const size_t npx = (size_t)w * h * 4;
  uint8_t* mask = (uint8_t*)calloc(npx, 1);
  uint8_t* fill = (uint8_t*)calloc(npx, 1);
  if (!mask || !fill) {
    free(mask);
    free(fill);
    return orc_fail(ORC_ERR_ARG, "out of memory");
  }
  orc_rgbx white = {255, 255, 255, 255};
  int rc = orc_fill_path(mask, w, h, cmds, n, mat9, white, winding, ORC_NORMAL, st);
  if (rc == ORC_OK) {
    orc_paint q = *p;
    q.opacity = 1.0f;
    rc = orc_fill_gradient(fill, w, h, &q);
  }
  if (rc == ORC_OK) {
    if (opacity != 1.0f) orc_apply_opacity(mask, w, h, opacity);
    orc_draw_rect(fill, w, h, mask, w, h, 0, 0, ORC_MASK);
    orc_draw_rect(img, w, h, fill, w, h, 0, 0, p->mode);
  }
  free(mask);
  free(fill);
  return rc;
}
