/*
 * oracle_pixels.c -- pixel and image operations of pixie v6.0.0, scalar
 * semantics.  TEST INFRASTRUCTURE ONLY (see pixie_oracle.h).
 *
 * Integer arithmetic follows Nim's types exactly: uint8 arithmetic wraps
 * modulo 256, uint32 modulo 2^32, `div` truncates.
 */
#include "pixie_oracle.h"
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static orc_rgbx mk(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  orc_rgbx c = {(uint8_t)r, (uint8_t)g, (uint8_t)b, (uint8_t)a};
  return c;
}

/* ---------------------------------------------------------------------- */
/* common.nim                                                             */
/* ---------------------------------------------------------------------- */

/* `*`(rgbx: ColorRGBX, opacity: uint8)  common.nim:79-90 */
orc_rgbx orc_mul_cov(orc_rgbx c, uint8_t cov) {
  if (cov == 0) return mk(0, 0, 0, 0);
  if (cov == 255) return c;
  return mk((c.r * (uint32_t)cov + 127) / 255, (c.g * (uint32_t)cov + 127) / 255,
            (c.b * (uint32_t)cov + 127) / 255, (c.a * (uint32_t)cov + 127) / 255);
}

/* `*`(color: ColorRGBX, opacity: float32)  common.nim:67-77.  x is
 * round(opacity*255) converted to uint32 the way the x86-64 C backend does
 * (through a 64-bit signed conversion), products wrap modulo 2^32. */
orc_rgbx orc_mul_opacity(orc_rgbx c, float opacity) {
  if (opacity == 0) return mk(0, 0, 0, 0);
  uint32_t x = nim_f2u32(roundf(opacity * 255.0f));
  return mk((c.r * x + 127) / 255, (c.g * x + 127) / 255, (c.b * x + 127) / 255,
            (c.a * x + 127) / 255);
}

/* ---------------------------------------------------------------------- */
/* chroma conversions (un-vendored dependency chroma >= 0.2.6; parity      */
/* unpinned except rgbx(ColorRGBA), which test_images.nim:204-228 pins).  */
/* ---------------------------------------------------------------------- */

/* ColorRGBA -> ColorRGBX premultiply: (c*a+127) div 255, identity at a=255.
 * Pinned by test_images.nim:204-228 (all 65,536 (r, a) pairs). */
orc_rgbx orc_rgbx_from_rgba(orc_rgbx s) {
  if (s.a == 255) return s;
  if (s.a == 0) return mk(0, 0, 0, 0);
  return mk((s.r * (uint32_t)s.a + 127) / 255, (s.g * (uint32_t)s.a + 127) / 255,
            (s.b * (uint32_t)s.a + 127) / 255, s.a);
}

/* ColorRGBX -> ColorRGBA un-premultiply.  chroma's rule as restated here:
 * multiplier = uint32((255 / a) * 255) (float32), channel =
 * (c * multiplier + 127) div 255, alpha 0 and 255 pass through.  Parity
 * unpinned (no reference test reaches it). */
orc_rgbx orc_rgba_from_rgbx(orc_rgbx c) {
  if (c.a == 0 || c.a == 255) return c;
  uint32_t mul = nim_f2u32((255.0f / (float)c.a) * 255.0f);
  return mk((c.r * mul + 127) / 255, (c.g * mul + 127) / 255,
            (c.b * mul + 127) / 255, c.a);
}

static uint8_t unit_to_u8(float v) {
  if (!(v > 0.0f)) v = 0.0f;
  if (v > 1.0f) v = 1.0f;
  return (uint8_t)roundf(v * 255.0f);
}

/* Color (float, straight) -> ColorRGBX: rgba() by round(c*255) then
 * premultiply.  Parity unpinned; the paint colours in the workloads are
 * handed over as RGBA8 already. */
orc_rgbx orc_rgbx_from_color(float r, float g, float b, float a) {
  orc_rgbx s = mk(unit_to_u8(r), unit_to_u8(g), unit_to_u8(b), unit_to_u8(a));
  return orc_rgbx_from_rgba(s);
}

/* internal.nim:68-84 straightAlphaTable / toStraightAlpha */
void orc_to_straight_alpha(uint8_t* d, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    uint8_t* p = d + 4 * i;
    uint32_t a = p[3];
    /* straightAlphaTable is a compile-time constant: Nim's VM evaluates
     * its float32 arithmetic in float64, so the table holds the float64
     * results (101 * 255 / 202 = 127.5000...01 rounds to 128, 39 * 255 /
     * 234 = 42.4999...9 to 42; float32 gives 127 and 43) */
    double mult = a > 0 ? (255.0 / (double)a) : 0.0;
    for (int k = 0; k < 3; k++) {
      double v = round((double)p[k] * mult);
      if (v > 255.0) v = 255.0;
      p[k] = (uint8_t)v;
    }
  }
}

/* internal.nim:86-96 toPremultipliedAlpha */
void orc_to_premultiplied_alpha(uint8_t* d, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    uint8_t* p = d + 4 * i;
    uint32_t a = p[3];
    if (a != 255) {
      p[0] = (uint8_t)((p[0] * a + 127) / 255);
      p[1] = (uint8_t)((p[1] * a + 127) / 255);
      p[2] = (uint8_t)((p[2] * a + 127) / 255);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* blends.nim                                                             */
/* ---------------------------------------------------------------------- */

/* blends.nim:18-39 alphaFix on straight ColorRGBA values. */
static orc_rgbx alpha_fix(orc_rgbx b, orc_rgbx s, orc_rgbx m) {
  uint32_t sa = s.a, ba = b.a;
  uint32_t t0 = sa * (255 - ba), t1 = sa * ba, t2 = (255 - sa) * ba;
  uint32_t r = t0 * s.r + t1 * m.r + t2 * b.r;
  uint32_t g = t0 * s.g + t1 * m.g + t2 * b.g;
  uint32_t bb = t0 * s.b + t1 * m.b + t2 * b.b;
  uint32_t a = sa + ba * (255 - sa) / 255;
  if (a == 0) return mk(0, 0, 0, 0);
  return mk(r / a / 255, g / a / 255, bb / a / 255, a);
}

/* blends.nim:41-43 */
static uint8_t blend_alpha(uint8_t b, uint8_t s) {
  return (uint8_t)(s + (uint8_t)(((uint32_t)b * (uint8_t)(255 - s)) / 255));
}

/* blends.nim:45-46 */
static uint8_t screen_u(uint32_t b, uint32_t s) {
  return (uint8_t)((int32_t)(b + s) - (int32_t)((b * s) / 255));
}

/* blends.nim:48-58 */
static uint8_t hard_light(uint32_t bc, uint32_t ba, uint32_t sc, uint32_t sa) {
  if (sc * 2 <= sa)
    return (uint8_t)((2 * sc * bc + sc * (255 - ba) + bc * (255 - sa)) / 255);
  return screen_u(bc, sc);
}

static orc_rgbx blend_normal(orc_rgbx b, orc_rgbx s) { /* 60-70 */
  if (b.a == 0 || s.a == 255) return s;
  if (s.a == 0) return b;
  uint32_t k = 255 - (uint32_t)s.a;
  orc_rgbx r;
  r.r = (uint8_t)(s.r + (uint8_t)((b.r * k) / 255));
  r.g = (uint8_t)(s.g + (uint8_t)((b.g * k) / 255));
  r.b = (uint8_t)(s.b + (uint8_t)((b.b * k) / 255));
  r.a = blend_alpha(b.a, s.a);
  return r;
}

static uint8_t darken_c(uint8_t bc, uint8_t ba, uint8_t sc, uint8_t sa) { /* 73-79 */
  uint32_t x = bc + ((uint32_t)(uint8_t)(255 - ba) * sc) / 255;
  uint32_t y = sc + ((uint32_t)(uint8_t)(255 - sa) * bc) / 255;
  return (uint8_t)(x < y ? x : y);
}

static uint8_t multiply_c(uint8_t bc, uint8_t ba, uint8_t sc, uint8_t sa) { /* 87-94 */
  return (uint8_t)(((uint32_t)(uint8_t)(255 - ba) * sc +
                    (uint32_t)(uint8_t)(255 - sa) * bc + (uint32_t)bc * sc) /
                   255);
}

static uint8_t lighten_c(uint8_t bc, uint8_t ba, uint8_t sc, uint8_t sa) { /* 129-135 */
  uint32_t x = bc + ((uint32_t)(uint8_t)(255 - ba) * sc) / 255;
  uint32_t y = sc + ((uint32_t)(uint8_t)(255 - sa) * bc) / 255;
  return (uint8_t)(x > y ? x : y);
}

static uint8_t color_burn_c(uint32_t b, uint32_t s) { /* 115-121 */
  if (b == 255) return 255;
  if (s == 0) return 0;
  uint32_t q = (255 * (255 - b)) / s;
  return (uint8_t)(255 - (uint8_t)(q < 255 ? q : 255));
}

static uint8_t color_dodge_c(uint32_t b, uint32_t s) { /* 162-168 */
  if (b == 0) return 0;
  if (s == 255) return 255;
  uint32_t q = (255 * b) / (255 - s);
  return (uint8_t)(q < 255 ? q : 255);
}

static uint8_t difference_c(uint8_t bc, uint8_t ba, uint8_t sc, uint8_t sa) { /* 191-199 */
  uint32_t x = (uint32_t)bc * sa, y = (uint32_t)sc * ba;
  uint32_t mn = x < y ? x : y;
  /* (bc + sc) is uint8 (wraps); the final .uint8 makes the wrap invisible. */
  return (uint8_t)((int32_t)(uint8_t)(bc + sc) - 2 * (int32_t)(mn / 255));
}

static uint8_t exclusion_c(uint32_t b, uint32_t s) { /* 207-209 */
  int32_t v = (int32_t)(b + s) - (int32_t)((2 * b * s) / 255);
  return (uint8_t)(v > 0 ? v : 0);
}

/* ---- float blend modes (chroma Color blends; W3C compositing-1, which
 * blends.nim:5-6 cites).  Parity unpinned. ---- */
typedef struct { float r, g, b, a; } fcol;

static fcol to_fcol(orc_rgbx c) {
  orc_rgbx s = orc_rgba_from_rgbx(c);
  fcol f = {(float)s.r / 255.0f, (float)s.g / 255.0f, (float)s.b / 255.0f,
            (float)s.a / 255.0f};
  return f;
}

static orc_rgbx from_fcol(fcol f) { return orc_rgbx_from_color(f.r, f.g, f.b, f.a); }

static fcol alpha_fix_f(fcol b, fcol s, fcol m) {
  fcol r;
  r.a = s.a + b.a * (1.0f - s.a);
  if (r.a == 0.0f) {
    fcol z = {0, 0, 0, 0};
    return z;
  }
  float t0 = s.a * (1.0f - b.a), t1 = s.a * b.a, t2 = (1.0f - s.a) * b.a;
  r.r = t0 * s.r + t1 * m.r + t2 * b.r;
  r.g = t0 * s.g + t1 * m.g + t2 * b.g;
  r.b = t0 * s.b + t1 * m.b + t2 * b.b;
  r.r /= r.a;
  r.g /= r.a;
  r.b /= r.a;
  return r;
}

static float soft_light_c(float b, float s) {
  if (s <= 0.5f) return b - (1.0f - 2.0f * s) * b * (1.0f - b);
  float d = (b <= 0.25f) ? ((16.0f * b - 12.0f) * b + 4.0f) * b : sqrtf(b);
  return b + (2.0f * s - 1.0f) * (d - b);
}

static float lum_f(fcol c) { return 0.3f * c.r + 0.59f * c.g + 0.11f * c.b; }

static fcol clip_color_f(fcol c) {
  float l = lum_f(c);
  float n = fminf(fminf(c.r, c.g), c.b);
  float x = fmaxf(fmaxf(c.r, c.g), c.b);
  /* the denominators can reach 0 through float32 rounding when every
   * channel sits at the luminosity; the spec's limit there is no change */
  if (n < 0.0f && l - n > 0.0f) {
    c.r = l + (c.r - l) * l / (l - n);
    c.g = l + (c.g - l) * l / (l - n);
    c.b = l + (c.b - l) * l / (l - n);
  }
  if (x > 1.0f && x - l > 0.0f) {
    c.r = l + (c.r - l) * (1.0f - l) / (x - l);
    c.g = l + (c.g - l) * (1.0f - l) / (x - l);
    c.b = l + (c.b - l) * (1.0f - l) / (x - l);
  }
  return c;
}

static fcol set_lum_f(fcol c, float l) {
  float d = l - lum_f(c);
  c.r += d;
  c.g += d;
  c.b += d;
  return clip_color_f(c);
}

static float sat_f(fcol c) {
  return fmaxf(fmaxf(c.r, c.g), c.b) - fminf(fminf(c.r, c.g), c.b);
}

static fcol set_sat_f(fcol c, float s) {
  /* W3C SetSat: scale mid channel, max -> s, min -> 0. */
  float* ch[3] = {&c.r, &c.g, &c.b};
  int imax = 0, imin = 0;
  for (int i = 1; i < 3; i++) {
    if (*ch[i] > *ch[imax]) imax = i;
    if (*ch[i] < *ch[imin]) imin = i;
  }
  if (imax == imin) {
    c.r = c.g = c.b = 0.0f;
    return c;
  }
  int imid = 3 - imax - imin;
  float cmax = *ch[imax], cmin = *ch[imin], cmid = *ch[imid];
  *ch[imid] = (cmid - cmin) * s / (cmax - cmin);
  *ch[imax] = s;
  *ch[imin] = 0.0f;
  return c;
}

static orc_rgbx blend_float(int mode, orc_rgbx bb, orc_rgbx ss) {
  fcol b = to_fcol(bb), s = to_fcol(ss), m = b;
  switch (mode) {
    case ORC_SOFT_LIGHT:
      m.r = soft_light_c(b.r, s.r);
      m.g = soft_light_c(b.g, s.g);
      m.b = soft_light_c(b.b, s.b);
      break;
    case ORC_HUE: m = set_lum_f(set_sat_f(s, sat_f(b)), lum_f(b)); break;
    case ORC_SATURATION: m = set_lum_f(set_sat_f(b, sat_f(s)), lum_f(b)); break;
    case ORC_COLOR: m = set_lum_f(s, lum_f(b)); break;
    case ORC_LUMINOSITY: m = set_lum_f(b, lum_f(s)); break;
    default: break;
  }
  return from_fcol(alpha_fix_f(b, s, m));
}

/* blender() dispatch, blends.nim:275-299 */
orc_rgbx orc_blend(int mode, orc_rgbx b, orc_rgbx s) {
  orc_rgbx r;
  switch (mode) {
    case ORC_NORMAL: return blend_normal(b, s);
    case ORC_DARKEN:
      r.r = darken_c(b.r, b.a, s.r, s.a);
      r.g = darken_c(b.g, b.a, s.g, s.a);
      r.b = darken_c(b.b, b.a, s.b, s.a);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_MULTIPLY:
      r.r = multiply_c(b.r, b.a, s.r, s.a);
      r.g = multiply_c(b.g, b.a, s.g, s.a);
      r.b = multiply_c(b.b, b.a, s.b, s.a);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_COLOR_BURN: {
      orc_rgbx bs = orc_rgba_from_rgbx(b), ss = orc_rgba_from_rgbx(s), m;
      m.r = color_burn_c(bs.r, ss.r);
      m.g = color_burn_c(bs.g, ss.g);
      m.b = color_burn_c(bs.b, ss.b);
      m.a = 0;
      return orc_rgbx_from_rgba(alpha_fix(bs, ss, m));
    }
    case ORC_LIGHTEN:
      r.r = lighten_c(b.r, b.a, s.r, s.a);
      r.g = lighten_c(b.g, b.a, s.g, s.a);
      r.b = lighten_c(b.b, b.a, s.b, s.a);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_SCREEN:
      r.r = screen_u(b.r, s.r);
      r.g = screen_u(b.g, s.g);
      r.b = screen_u(b.b, s.b);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_COLOR_DODGE: {
      orc_rgbx bs = orc_rgba_from_rgbx(b), ss = orc_rgba_from_rgbx(s), m;
      m.r = color_dodge_c(bs.r, ss.r);
      m.g = color_dodge_c(bs.g, ss.g);
      m.b = color_dodge_c(bs.b, ss.b);
      m.a = 0;
      return orc_rgbx_from_rgba(alpha_fix(bs, ss, m));
    }
    case ORC_OVERLAY: /* 175-179: hardLight with the roles swapped */
      r.r = hard_light(s.r, s.a, b.r, b.a);
      r.g = hard_light(s.g, s.a, b.g, b.a);
      r.b = hard_light(s.b, s.a, b.b, b.a);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_HARD_LIGHT: /* 184-188 */
      r.r = hard_light(b.r, b.a, s.r, s.a);
      r.g = hard_light(b.g, b.a, s.g, s.a);
      r.b = hard_light(b.b, b.a, s.b, s.a);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_DIFFERENCE:
      r.r = difference_c(b.r, b.a, s.r, s.a);
      r.g = difference_c(b.g, b.a, s.g, s.a);
      r.b = difference_c(b.b, b.a, s.b, s.a);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_EXCLUSION:
      r.r = exclusion_c(b.r, s.r);
      r.g = exclusion_c(b.g, s.g);
      r.b = exclusion_c(b.b, s.b);
      r.a = blend_alpha(b.a, s.a);
      return r;
    case ORC_SOFT_LIGHT:
    case ORC_HUE:
    case ORC_SATURATION:
    case ORC_COLOR:
    case ORC_LUMINOSITY: return blend_float(mode, b, s);
    case ORC_MASK: { /* 227-232 */
      uint32_t k = s.a;
      return mk((b.r * k) / 255, (b.g * k) / 255, (b.b * k) / 255, (b.a * k) / 255);
    }
    case ORC_OVERWRITE: return s;
    case ORC_SUBTRACT_MASK: { /* 234-239 */
      uint32_t a = ((uint32_t)b.a * (uint8_t)(255 - s.a)) / 255;
      return mk((b.r * a) / 255, (b.g * a) / 255, (b.b * a) / 255, a);
    }
    case ORC_EXCLUDE_MASK: { /* 241-246: note it scales the SOURCE colour */
      uint32_t mx = b.a > s.a ? b.a : s.a, mn = b.a < s.a ? b.a : s.a;
      uint32_t a = mx - mn;
      return mk((s.r * a) / 255, (s.g * a) / 255, (s.b * a) / 255, a);
    }
    default: return b;
  }
}

void orc_blend_pixels(int mode, const uint8_t* bd, const uint8_t* sc,
                      uint8_t* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    orc_rgbx b = {bd[4 * i], bd[4 * i + 1], bd[4 * i + 2], bd[4 * i + 3]};
    orc_rgbx s = {sc[4 * i], sc[4 * i + 1], sc[4 * i + 2], sc[4 * i + 3]};
    orc_rgbx r = orc_blend(mode, b, s);
    out[4 * i] = r.r;
    out[4 * i + 1] = r.g;
    out[4 * i + 2] = r.b;
    out[4 * i + 3] = r.a;
  }
}

/* ---------------------------------------------------------------------- */
/* internal.nim / images.nim                                              */
/* ---------------------------------------------------------------------- */

/* internal.nim:17-34.  s and the exp() argument are float32; the
 * 1/sqrt(2*PI*s^2) factor is float64 because `2 * PI` folds to a float64
 * constant, so each tap a is float64, summed in float64, stored to float32,
 * normalised in float64 and rounded from float32. */
void orc_gaussian_kernel(int radius, uint16_t* out) {
  int n = 2 * radius + 1;
  float* floats = (float*)malloc(sizeof(float) * (size_t)n);
  double total = 0.0;
  for (int step = -radius; step <= radius; step++) {
    float s = (float)radius / 2.2f;
    float s2 = s * s;
    float st = (float)step;
    double a = (1.0 / sqrt(6.283185307179586 * (double)s2)) *
               (double)expf((-1.0f * (st * st)) / (2.0f * s2));
    floats[step + radius] = (float)a;
    total += a;
  }
  for (int i = 0; i < n; i++) floats[i] = (float)((double)floats[i] / total);
  for (int i = 0; i < n; i++) out[i] = (uint16_t)roundf(floats[i] * 255.0f * 256.0f);
  free(floats);
}

/* images.nim:55-57 + internal.nim:55-66 */
void orc_fill(uint8_t* img, int w, int h, orc_rgbx c) {
  int64_t n = (int64_t)w * h;
  for (int64_t i = 0; i < n; i++) {
    img[4 * i] = c.r;
    img[4 * i + 1] = c.g;
    img[4 * i + 2] = c.b;
    img[4 * i + 3] = c.a;
  }
}

/* images.nim:261-277 */
void orc_apply_opacity(uint8_t* img, int w, int h, float opacity) {
  uint16_t o = (uint16_t)nim_f2u32(roundf(255.0f * opacity));
  if (o == 255) return;
  if (o == 0) {
    orc_rgbx z = {0, 0, 0, 0};
    orc_fill(img, w, h, z);
    return;
  }
  int64_t n = (int64_t)w * h * 4;
  for (int64_t i = 0; i < n; i++) img[i] = (uint8_t)((img[i] * (uint32_t)o) / 255);
}

/* images.nim:304-365.  Two passes; each accumulates uint32 sum(c*k) and
 * stores (sum div 256 div 255) as uint8.  The X pass output is kept
 * transposed as in pixie (the layout does not change the values). */
int orc_blur(uint8_t* img, int w, int h, float radius_f, orc_rgbx oob) {
  int64_t radius = nim_f2i(roundf(radius_f));
  if (radius == 0) return ORC_OK;
  if (radius < 0) {
    orc_set_error("Cannot apply negative blur");
    return ORC_ERR_NEG_BLUR;
  }
  int r = (int)radius;
  uint16_t* k = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(2 * r + 1));
  orc_gaussian_kernel(r, k);
  uint8_t* bx = (uint8_t*)malloc((size_t)w * h * 4); /* transposed: [x][y] */
  const uint8_t oc[4] = {oob.r, oob.g, oob.b, oob.a};
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      uint32_t v[4] = {0, 0, 0, 0};
      for (int xx = x - r; xx < (x + r < 0 ? x + r : 0); xx++)
        for (int c = 0; c < 4; c++) v[c] += oc[c] * (uint32_t)k[xx - x + r];
      int lo = x - r > 0 ? x - r : 0, hi = x + r < w - 1 ? x + r : w - 1;
      for (int xx = lo; xx <= hi; xx++)
        for (int c = 0; c < 4; c++)
          v[c] += img[((int64_t)y * w + xx) * 4 + c] * (uint32_t)k[xx - x + r];
      for (int xx = (x - r > w ? x - r : w); xx <= x + r; xx++)
        for (int c = 0; c < 4; c++) v[c] += oc[c] * (uint32_t)k[xx - x + r];
      for (int c = 0; c < 4; c++) bx[((int64_t)x * h + y) * 4 + c] = (uint8_t)(v[c] / 256 / 255);
    }
  }
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      uint32_t v[4] = {0, 0, 0, 0};
      for (int yy = y - r; yy < (y + r < 0 ? y + r : 0); yy++)
        for (int c = 0; c < 4; c++) v[c] += oc[c] * (uint32_t)k[yy - y + r];
      int lo = y - r > 0 ? y - r : 0, hi = y + r < h - 1 ? y + r : h - 1;
      for (int yy = lo; yy <= hi; yy++)
        for (int c = 0; c < 4; c++)
          v[c] += bx[((int64_t)x * h + yy) * 4 + c] * (uint32_t)k[yy - y + r];
      for (int yy = (y - r > h ? y - r : h); yy <= y + r; yy++)
        for (int c = 0; c < 4; c++) v[c] += oc[c] * (uint32_t)k[yy - y + r];
      for (int c = 0; c < 4; c++) img[((int64_t)y * w + x) * 4 + c] = (uint8_t)(v[c] / 256 / 255);
    }
  }
  free(bx);
  free(k);
  return ORC_OK;
}

/* images.nim:700-758.  Separable max (spread > 0) or min (spread < 0) of
 * alpha over a window clamped to the image; output rgbx(0,0,0,a). */
int orc_spread(uint8_t* img, int w, int h, float spread_f) {
  int64_t sp = nim_f2i(roundf(spread_f));
  if (sp == 0) return ORC_OK;
  uint8_t* sx = (uint8_t*)calloc((size_t)w * h, 1); /* alpha plane, [x][y] */
  int grow = sp > 0;
  int s = (int)(grow ? sp : -sp);
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      uint8_t v = grow ? 0 : 255;
      int lo = x - s > 0 ? x - s : 0, hi = x + s < w - 1 ? x + s : w - 1;
      for (int xx = lo; xx <= hi; xx++) {
        uint8_t a = img[((int64_t)y * w + xx) * 4 + 3];
        if (grow ? a > v : a < v) v = a;
        if (v == (grow ? 255 : 0)) break;
      }
      sx[(int64_t)x * h + y] = v;
    }
  }
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      uint8_t v = grow ? 0 : 255;
      int lo = y - s > 0 ? y - s : 0, hi = y + s < h - 1 ? y + s : h - 1;
      for (int yy = lo; yy <= hi; yy++) {
        uint8_t a = sx[(int64_t)x * h + yy];
        if (grow ? a > v : a < v) v = a;
        if (v == (grow ? 255 : 0)) break;
      }
      uint8_t* p = img + ((int64_t)y * w + x) * 4;
      p[0] = p[1] = p[2] = 0;
      p[3] = v;
    }
  }
  free(sx);
  return ORC_OK;
}

/* images.nim:468-529 blendRect */
void orc_draw_rect(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh,
                   int px, int py, int mode) {
  if (px >= aw || px + bw <= 0 || py >= ah || py + bh <= 0) {
    if (mode == ORC_MASK) memset(a, 0, (size_t)aw * ah * 4);
    return;
  }
  int xStart = -px > 0 ? -px : 0, yStart = -py > 0 ? -py : 0;
  int xEnd = bw < aw - px ? bw : aw - px, yEnd = bh < ah - py ? bh : ah - py;
  if (mode == ORC_MASK && yStart + py > 0) memset(a, 0, (size_t)(yStart + py) * aw * 4);
  for (int y = yStart; y < yEnd; y++) {
    uint8_t* arow = a + ((int64_t)(y + py) * aw) * 4;
    const uint8_t* brow = b + ((int64_t)y * bw) * 4;
    if (mode == ORC_MASK && xStart + px > 0) memset(arow, 0, (size_t)(xStart + px) * 4);
    for (int x = xStart; x < xEnd; x++) {
      uint8_t* ap = arow + (x + px) * 4;
      const uint8_t* bp = brow + x * 4;
      orc_rgbx bb = {ap[0], ap[1], ap[2], ap[3]};
      orc_rgbx ss = {bp[0], bp[1], bp[2], bp[3]};
      orc_rgbx r = orc_blend(mode, bb, ss);
      ap[0] = r.r;
      ap[1] = r.g;
      ap[2] = r.b;
      ap[3] = r.a;
    }
    if (mode == ORC_MASK && xEnd + px < aw)
      memset(arow + (xEnd + px) * 4, 0, (size_t)(aw - (xEnd + px)) * 4);
  }
  if (mode == ORC_MASK && yEnd + py < ah)
    memset(a + ((int64_t)(yEnd + py) * aw) * 4, 0, (size_t)(ah - (yEnd + py)) * aw * 4);
}

/* images.nim:760-776 shadow.  The offset copy is draw(image,
 * translate(offset), OverwriteBlend): the integer composite for whole-pixel
 * offsets, drawSmooth for fractional ones (images.nim:636-678). */
int orc_shadow(uint8_t* out, const uint8_t* img, int w, int h, float offx,
               float offy, float spread, float blur, orc_rgbx color) {
  size_t bytes = (size_t)w * h * 4;
  uint8_t* mask = (uint8_t*)malloc(bytes);
  if (offx == 0.0f && offy == 0.0f) {
    memcpy(mask, img, bytes);
  } else {
    memset(mask, 0, bytes);
    const float t[9] = {1, 0, 0, 0, 1, 0, offx, offy, 1};
    const int rc0 = orc_draw(mask, w, h, img, w, h, t, ORC_OVERWRITE);
    if (rc0 != ORC_OK) {
      free(mask);
      return rc0;
    }
  }
  int rc = orc_spread(mask, w, h, spread);
  if (rc == ORC_OK) rc = orc_blur(mask, w, h, blur, (orc_rgbx){0, 0, 0, 0});
  if (rc == ORC_OK) {
    orc_fill(out, w, h, color);
    orc_draw_rect(out, w, h, mask, w, h, 0, 0, ORC_MASK);
  }
  free(mask);
  return rc;
}

/* images.nim:138-166 */
float orc_diff_score(const uint8_t* m, const uint8_t* u, int w, int h) {
  int64_t score = 0, total = 0;
  for (int64_t i = 0; i < (int64_t)w * h; i++) {
    for (int c = 0; c < 4; c++) {
      int d = (int)m[4 * i + c] - (int)u[4 * i + c];
      score += d < 0 ? -d : d;
    }
    total += 255 * 4;
  }
  return 100.0f * (float)score / (float)total;
}
