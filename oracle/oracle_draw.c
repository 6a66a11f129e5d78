/*
 * oracle_draw.c -- transformed image draws of pixie v6.0.0.  TEST
 * INFRASTRUCTURE ONLY (see pixie_oracle.h).
 *
 *   mix(ColorRGBX)          common.nim:59-65
 *   minifyBy2 / magnifyBy2  images.nim:168-259 (the AVX2 minify loop,
 *                           avx2.nim:316-403, computes the same sums)
 *   getRgbaSmooth           images.nim:367-403
 *   drawCorrect (tiled)     images.nim:405-445, drawTiled 680-683
 *   drawSmooth              images.nim:531-634
 *   draw                    images.nim:636-678
 *
 * The vmath and bumpy helpers these call (Mat3 inverse and products,
 * Vec2 length, fractional, intersects(Line, Segment)) are restated in
 * float32; neither library is vendored, so the draw masters
 * (tests/images/masters/smooth*.png, minify_odd, rock_minified*, rotate*,
 * scaleHalf) are what pins this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"
#include "pixie_oracle.h"

static const float H = 0.5f; /* images.nim:5 */

static int fail(int code, const char* msg) {
  orc_set_error(msg);
  return code;
}

static orc_rgbx px_at(const uint8_t* img, int64_t i) {
  orc_rgbx c = {img[4 * i], img[4 * i + 1], img[4 * i + 2], img[4 * i + 3]};
  return c;
}
static void px_put(uint8_t* img, int64_t i, orc_rgbx c) {
  img[4 * i] = c.r;
  img[4 * i + 1] = c.g;
  img[4 * i + 2] = c.b;
  img[4 * i + 3] = c.a;
}
static int px_eq(orc_rgbx a, orc_rgbx b) {
  return a.r == b.r && a.g == b.g && a.b == b.b && a.a == b.a;
}

/* common.nim:59-65: integer weights round(t * 255) */
orc_rgbx orc_mix(orc_rgbx a, orc_rgbx b, float t) {
  const uint32_t x = nim_f2u32(roundf(t * 255.0f));
  orc_rgbx r;
  r.r = (uint8_t)((a.r * (255u - x) + b.r * x + 127u) / 255u);
  r.g = (uint8_t)((a.g * (255u - x) + b.g * x + 127u) / 255u);
  r.b = (uint8_t)((a.b * (255u - x) + b.b * x + 127u) / 255u);
  r.a = (uint8_t)((a.a * (255u - x) + b.a * x + 127u) / 255u);
  return r;
}

/* one minifyBy2 step (images.nim:176-228) into dst of ((w+1)/2, (h+1)/2) */
static void minify_once(const uint8_t* src, int w, int h, uint8_t* dst) {
  const int ew = w / 2, eh = h / 2, rw = (w & 1) ? ew + 1 : ew;
  for (int y = 0; y < eh; y++) {
    const int64_t top = (int64_t)w * (2 * y), bot = (int64_t)w * (2 * y + 1);
    for (int x = 0; x < ew; x++) {
      const orc_rgbx a = px_at(src, top + 2 * x), b = px_at(src, top + 2 * x + 1);
      const orc_rgbx c = px_at(src, bot + 2 * x + 1), d = px_at(src, bot + 2 * x);
      orc_rgbx m;
      m.r = (uint8_t)(((uint32_t)a.r + b.r + c.r + d.r + 2) / 4);
      m.g = (uint8_t)(((uint32_t)a.g + b.g + c.g + d.g + 2) / 4);
      m.b = (uint8_t)(((uint32_t)a.b + b.b + c.b + d.b + 2) / 4);
      m.a = (uint8_t)(((uint32_t)a.a + b.a + c.a + d.a + 2) / 4);
      px_put(dst, (int64_t)rw * y + x, m);
    }
    if (w & 1) {
      const orc_rgbx m = orc_mix(px_at(src, top + w - 1), px_at(src, bot + w - 1), 0.5f);
      px_put(dst, (int64_t)rw * y + rw - 1, orc_mul_opacity(m, 0.5f));
    }
  }
  if (h & 1) {
    const int64_t last = (int64_t)w * (h - 1);
    for (int x = 0; x < ew; x++) {
      const orc_rgbx m = orc_mix(px_at(src, last + 2 * x), px_at(src, last + 2 * x + 1), 0.5f);
      px_put(dst, (int64_t)rw * eh + x, orc_mul_opacity(m, 0.5f));
    }
    if (w & 1) px_put(dst, (int64_t)rw * eh + rw - 1, orc_mul_opacity(px_at(src, last + w - 1), 0.25f));
  }
}

void orc_minify_size(int w, int h, int power, int* ow, int* oh) {
  for (int i = 0; i < power; i++) {
    w = (w + 1) / 2;
    h = (h + 1) / 2;
  }
  *ow = w;
  *oh = h;
}

/* minifyBy2(power) into dst, whose size orc_minify_size gives */
int orc_minify_by2(const uint8_t* src, int w, int h, int power, uint8_t* dst) {
  if (power < 0) return fail(ORC_ERR_ARG, "Cannot minifyBy2 with negative power");
  if (power == 0) {
    memcpy(dst, src, (size_t)w * h * 4);
    return ORC_OK;
  }
  const uint8_t* cur = src;
  uint8_t* tmp = NULL;
  int cw = w, ch = h;
  for (int i = 0; i < power; i++) {
    const int nw = (cw + 1) / 2, nh = (ch + 1) / 2;
    uint8_t* out = (i == power - 1) ? dst : (uint8_t*)malloc((size_t)nw * nh * 4);
    if (!out) {
      free(tmp);
      return fail(ORC_ERR_NOMEM, "out of memory");
    }
    minify_once(cur, cw, ch, out);
    free(tmp);
    tmp = (out == dst) ? NULL : out;
    cur = out;
    cw = nw;
    ch = nh;
  }
  return ORC_OK;
}

/* magnifyBy2(power) (images.nim:230-259): each pixel repeated 2^power times
 * in x and y, into dst of (w << power, h << power) */
int orc_magnify_by2(const uint8_t* src, int w, int h, int power, uint8_t* dst) {
  if (power < 0) return fail(ORC_ERR_ARG, "Cannot magnifyBy2 with negative power");
  const int s = 1 << power, ow = w * s;
  for (int y = 0; y < h * s; y++)
    for (int x = 0; x < ow; x++) px_put(dst, (int64_t)ow * y + x, px_at(src, (int64_t)w * (y / s) + x / s));
  return ORC_OK;
}

/* images.nim:367-403.  Outside the image reads transparent; wrapped reads
 * use Nim's mod (the sign of the dividend), and pixie then indexes the data
 * at w * (y0 mod h) + (x0 mod w) without a bounds check: a flat index inside
 * the buffer reads that pixel, one outside it (undefined in pixie) reads
 * transparent here. */
static orc_rgbx fetch(const uint8_t* img, int w, int h, int64_t x, int64_t y, int wrapped) {
  const orc_rgbx zero = {0, 0, 0, 0};
  if (wrapped) {
    const int64_t i = (int64_t)w * (y % h) + (x % w);
    if (i < 0 || i >= (int64_t)w * h) return zero;
    return px_at(img, i);
  }
  if (x < 0 || x >= w || y < 0 || y >= h) return zero;
  return px_at(img, (int64_t)w * y + x);
}

orc_rgbx orc_get_rgba_smooth(const uint8_t* img, int w, int h, float x, float y, int wrapped) {
  const float fx = floorf(x), fy = floorf(y);
  const int64_t x0 = nim_f2i(fx), y0 = nim_f2i(fy);
  const int64_t x1 = x0 + 1, y1 = y0 + 1;
  const float xf = x - fx, yf = y - fy;
  const orc_rgbx x0y0 = fetch(img, w, h, x0, y0, wrapped), x1y0 = fetch(img, w, h, x1, y0, wrapped);
  const orc_rgbx x0y1 = fetch(img, w, h, x0, y1, wrapped), x1y1 = fetch(img, w, h, x1, y1, wrapped);
  orc_rgbx top = x0y0, bottom = x0y1;
  if (xf > 0.0f && !px_eq(x0y0, x1y0)) top = orc_mix(x0y0, x1y0, xf);
  if (xf > 0.0f && !px_eq(x0y1, x1y1)) bottom = orc_mix(x0y1, x1y1, xf);
  if (yf != 0.0f && !px_eq(top, bottom)) return orc_mix(top, bottom, yf);
  return top;
}

static float vlen(float x, float y) { return sqrtf(x * x + y * y); }

/* the movement vectors of a draw: p = inv * (h, h), dx, dy (images.nim:
 * 409-413, 538-542, 641-645) */
static void movement(const float* inv, float* p, float* dx, float* dy) {
  float ax, ay, bx, by;
  orc_m3_apply(inv, 0 + H, 0 + H, &p[0], &p[1]);
  orc_m3_apply(inv, 1 + H, 0 + H, &ax, &ay);
  orc_m3_apply(inv, 0 + H, 1 + H, &bx, &by);
  dx[0] = ax - p[0];
  dx[1] = ay - p[1];
  dy[0] = bx - p[0];
  dy[1] = by - p[1];
}

/* The minify / magnify steps a draw takes before sampling.  Returns the
 * power (> 0 minify, < 0 magnify) through *steps, with the image and the
 * adjusted vectors.  pixie loops forever when filterBy2 is 0 or not finite
 * (a degenerate transform); that raises here. */
static int filter_steps(float* p, float* dx, float* dy, int* steps) {
  float f = nim_maxf(vlen(dx[0], dx[1]), vlen(dy[0], dy[1]));
  if (!(f > 0.0f) || isinf(f)) return fail(ORC_ERR_ARG, "degenerate draw transform");
  int n = 0;
  while (f >= 2.0f) {
    for (int k = 0; k < 2; k++) {
      p[k] /= 2;
      dx[k] /= 2;
      dy[k] /= 2;
    }
    f /= 2;
    n++;
  }
  while (f <= 0.5f) {
    for (int k = 0; k < 2; k++) {
      p[k] *= 2;
      dx[k] *= 2;
      dy[k] *= 2;
    }
    f *= 2;
    n--;
  }
  *steps = n;
  return ORC_OK;
}

/* b minified (steps > 0) or magnified (steps < 0); *out is b itself for 0 */
static int resample(const uint8_t* b, int bw, int bh, int steps, uint8_t** out, int* ow, int* oh) {
  if (steps == 0) {
    *out = (uint8_t*)b;
    *ow = bw;
    *oh = bh;
    return ORC_OK;
  }
  if (steps > 0) {
    orc_minify_size(bw, bh, steps, ow, oh);
    *out = (uint8_t*)malloc((size_t)*ow * *oh * 4);
    if (!*out) return fail(ORC_ERR_NOMEM, "out of memory");
    return orc_minify_by2(b, bw, bh, steps, *out);
  }
  const int s = -steps;
  if (s > 20 || ((int64_t)bw << s) > 1 << 30 || ((int64_t)bh << s) > 1 << 30 ||
      ((int64_t)bw << s) * ((int64_t)bh << s) > ((int64_t)1 << 31))
    return fail(ORC_ERR_NOMEM, "magnified draw source too large");
  *ow = bw << s;
  *oh = bh << s;
  *out = (uint8_t*)malloc((size_t)*ow * *oh * 4);
  if (!*out) return fail(ORC_ERR_NOMEM, "out of memory");
  return orc_magnify_by2(b, bw, bh, s, *out);
}

/* drawCorrect (images.nim:405-445) */
int orc_draw_correct(uint8_t* a, int aw, int ah, const uint8_t* b0, int bw0, int bh0,
                     const float* mat9, int mode, int tiled) {
  float inv[9], p[2], dx[2], dy[2];
  orc_m3_inverse(mat9, inv);
  movement(inv, p, dx, dy);
  int steps, rc;
  if ((rc = filter_steps(p, dx, dy, &steps)) != ORC_OK) return rc;
  /* inverseTransform = scale(1/2) * inv per minify, scale(2) * inv per
   * magnify: exact powers of two on every entry of the first two rows */
  const float k = ldexpf(1.0f, -steps);
  for (int c = 0; c < 3; c++) {
    inv[3 * c + 0] *= k;
    inv[3 * c + 1] *= k;
  }
  uint8_t* b;
  int bw, bh;
  if ((rc = resample(b0, bw0, bh0, steps, &b, &bw, &bh)) != ORC_OK) return rc;
  for (int y = 0; y < ah; y++)
    for (int x = 0; x < aw; x++) {
      float sx, sy;
      orc_m3_apply(inv, (float)x + H, (float)y + H, &sx, &sy);
      const orc_rgbx s = orc_get_rgba_smooth(b, bw, bh, sx - H, sy - H, tiled);
      const int64_t i = (int64_t)aw * y + x;
      px_put(a, i, orc_blend(mode, px_at(a, i), s));
    }
  if (b != b0) free(b);
  return ORC_OK;
}

/* bumpy intersects(Line, Segment): the line through la, lb; u is the
 * parameter along the segment */
static int line_segment(float lax, float lay, float lbx, float lby, const float* s, float* at) {
  const float s1x = lbx - lax, s1y = lby - lay;
  const float s2x = s[2] - s[0], s2y = s[3] - s[1];
  const float den = (-s2x * s1y + s1x * s2y);
  const float num = s1x * (lay - s[1]) - s1y * (lax - s[0]);
  const float u = num / den;
  if (u >= 0.0f && u <= 1.0f) {
    at[0] = s[0] + u * s2x;
    at[1] = s[1] + u * s2y;
    return 1;
  }
  return 0;
}

static void zero_px(uint8_t* a, int64_t from, int64_t n) {
  if (n > 0) memset(a + 4 * from, 0, (size_t)n * 4);
}

/* drawSmooth (images.nim:531-634) */
static int draw_smooth(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh,
                       const float* m, int mode) {
  float c[4][2];
  orc_m3_apply(m, 0, 0, &c[0][0], &c[0][1]);
  orc_m3_apply(m, (float)bw, 0, &c[1][0], &c[1][1]);
  orc_m3_apply(m, (float)bw, (float)bh, &c[2][0], &c[2][1]);
  orc_m3_apply(m, 0, (float)bh, &c[3][0], &c[3][1]);
  float per[4][4];
  for (int i = 0; i < 4; i++) {
    per[i][0] = c[i][0];
    per[i][1] = c[i][1];
    per[i][2] = c[(i + 1) & 3][0];
    per[i][3] = c[(i + 1) & 3][1];
  }
  float inv[9], p[2], dx[2], dy[2];
  orc_m3_inverse(m, inv);
  movement(inv, p, dx, dy);
  int64_t yStart = ah, yEnd = 0;
  for (int i = 0; i < 4; i++) {
    const int64_t f = nim_f2i(floorf(per[i][1])), cl = nim_f2i(ceilf(per[i][1]));
    if (f < yStart) yStart = f;
    if (cl > yEnd) yEnd = cl;
  }
  yStart = yStart < 0 ? 0 : (yStart > ah ? ah : yStart);
  yEnd = yEnd < 0 ? 0 : (yEnd > ah ? ah : yEnd);
  if (mode == ORC_MASK && yStart > 0) zero_px(a, 0, yStart * aw);
  orc_rgbx* line = (orc_rgbx*)malloc(sizeof(orc_rgbx) * (size_t)(aw > 0 ? aw : 1));
  if (!line) return fail(ORC_ERR_NOMEM, "out of memory");
  for (int64_t y = yStart; y < yEnd; y++) {
    float xMin = (float)aw, xMax = 0.0f;
    for (int yo = 0; yo < 2; yo++) {
      const float ly = (float)y + (float)yo;
      for (int i = 0; i < 4; i++) {
        float at[2];
        if (line_segment(-1000.0f, ly, 1000.0f, ly, per[i], at) &&
            !(per[i][2] == at[0] && per[i][3] == at[1])) {
          xMin = nim_minf(xMin, at[0]);
          xMax = nim_maxf(xMax, at[0]);
        }
      }
    }
    int64_t xStart = nim_f2i(floorf(xMin)), xEnd = nim_f2i(ceilf(xMax));
    xStart = xStart < 0 ? 0 : (xStart > aw ? aw : xStart);
    xEnd = xEnd < 0 ? 0 : (xEnd > aw ? aw : xEnd);
    if (xEnd - xStart == 0) continue;
    /* srcPos = p + dx * xStart + dy * y, then - h, then += dx per pixel */
    float sx = p[0] + dx[0] * (float)xStart + dy[0] * (float)y;
    float sy = p[1] + dx[1] * (float)xStart + dy[1] * (float)y;
    sx = sx - H;
    sy = sy - H;
    for (int64_t x = xStart; x < xEnd; x++) {
      line[x] = orc_get_rgba_smooth(b, bw, bh, sx, sy, 0);
      sx += dx[0];
      sy += dx[1];
    }
    const int64_t row = (int64_t)aw * y;
    if (mode == ORC_OVERWRITE) {
      for (int64_t x = xStart; x < xEnd; x++) px_put(a, row + x, line[x]);
    } else {
      if (mode == ORC_MASK) zero_px(a, row, xStart);
      for (int64_t x = xStart; x < xEnd; x++) px_put(a, row + x, orc_blend(mode, px_at(a, row + x), line[x]));
      if (mode == ORC_MASK) zero_px(a, row + xEnd, aw - xEnd);
    }
  }
  if (mode == ORC_MASK && ah - yEnd > 0) zero_px(a, yEnd * aw, (ah - yEnd) * aw);
  free(line);
  return ORC_OK;
}

/* draw (images.nim:636-678) */
int orc_draw(uint8_t* a, int aw, int ah, const uint8_t* b0, int bw0, int bh0, const float* mat9,
             int mode) {
  float t[9], inv[9], p[2], dx[2], dy[2];
  memcpy(t, mat9, sizeof(t));
  orc_m3_inverse(t, inv);
  movement(inv, p, dx, dy);
  int steps, rc;
  if ((rc = filter_steps(p, dx, dy, &steps)) != ORC_OK) return rc;
  /* transform * scale(2) per minify, * scale(1/2) per magnify: the first
   * two columns scale by a power of two */
  const float k = ldexpf(1.0f, steps);
  for (int r = 0; r < 3; r++) {
    t[r] *= k;
    t[3 + r] *= k;
  }
  uint8_t* b;
  int bw, bh;
  if ((rc = resample(b0, bw0, bh0, steps, &b, &bw, &bh)) != ORC_OK) return rc;
  const int hasRotationOrScaling = !(dx[0] == 1.0f && dx[1] == 0.0f && dy[0] == 0.0f && dy[1] == 1.0f);
  /* vmath fractional(v) = |v| - floor(|v|); only its test against 0 is used */
  const float tx = t[6], ty = t[7];
  const int smooth = !(vlen(dx[0], dx[1]) == 1.0f && vlen(dy[0], dy[1]) == 1.0f &&
                       (fabsf(tx) - floorf(fabsf(tx))) == 0.0f &&
                       (fabsf(ty) - floorf(fabsf(ty))) == 0.0f);
  if (hasRotationOrScaling || smooth) {
    rc = draw_smooth(a, aw, ah, b, bw, bh, t, mode);
  } else {
    /* ivec2(transform[2, 0].int32, ...): offsets past the int32 range put
     * the image entirely off the canvas */
    const float lim = 2147483520.0f;
    if (fabsf(tx) > lim || fabsf(ty) > lim) {
      if (mode == ORC_MASK) memset(a, 0, (size_t)aw * ah * 4);
    } else {
      orc_draw_rect(a, aw, ah, b, bw, bh, (int)tx, (int)ty, mode);
    }
  }
  if (b != b0) free(b);
  return rc;
}

int orc_draw_tiled(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh, const float* mat9,
                   int mode) {
  return orc_draw_correct(a, aw, ah, b, bw, bh, mat9, mode, 1);
}
