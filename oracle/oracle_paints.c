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
  if (p->kind == ORC_LINEAR_GRADIENT && p->nhandles != 2)
    return orc_fail(ORC_ERR_ARG, "Linear gradient requires 2 handles");
  if (p->kind == ORC_RADIAL_GRADIENT && p->nhandles != 3)
    return orc_fail(ORC_ERR_ARG, "Radial gradient requires 3 handles");
  if (p->kind == ORC_ANGULAR_GRADIENT && p->nhandles != 3)
    return orc_fail(ORC_ERR_ARG, "Angular gradient requires 2 handles"); /* sic: 3 are read */
  if (p->kind < ORC_LINEAR_GRADIENT || p->kind > ORC_ANGULAR_GRADIENT)
    return orc_fail(ORC_ERR_ARG, "Paint must be a gradient");
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
        float px, py;
        orc_m3_apply(m, (float)x, (float)y, &px, &py);
        const float t = plen(pv(px, py));
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
  orc_m3_mul(T, R, TR);
  orc_m3_mul(TR, S, TRS);
  orc_m3_inverse(TRS, m);
}

/* fillPath with a non-solid paint (paths.nim:2115-2142): image, tiled
 * image or gradient. */
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
    if (p->kind == ORC_IMAGE_PAINT)
      rc = orc_draw(fill, w, h, p->image, p->iw, p->ih, p->image_mat, ORC_NORMAL);
    else if (p->kind == ORC_TILED_IMAGE_PAINT)
      rc = orc_draw_tiled(fill, w, h, p->image, p->iw, p->ih, p->image_mat, ORC_NORMAL);
    else
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
