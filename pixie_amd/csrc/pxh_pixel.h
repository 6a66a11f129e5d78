// pxh_pixel.h -- premultiplied RGBA8 pixel operations (blends.nim,
// common.nim:67-90), shared by the HIP kernels and host code.
//
// Pixels are packed uint32 (r low byte).  Integer modes are exact restatements
// of blends.nim with Nim's wrapping uint8/uint32 arithmetic.  ColorBurn /
// ColorDodge go through chroma's straight-alpha conversion and the five
// non-separable / soft-light modes through chroma's float32 Color blends
// (W3C compositing-1); those conversions are un-vendored in the reference and
// restated here (no reference fixture pins them: parity unpinned).
#pragma once

#include "pxh_common.h"

#if !(defined(__HIPCC__) || defined(__HIP__))
#include <math.h>
#endif

namespace pxh {

// `*`(ColorRGBX, uint8) common.nim:79-90, generalised to any uint32 factor x
// (the float-opacity product common.nim:67-77 computes x = round(o*255) and
// the same (c*x+127) div 255, wrapping modulo 2^32).
PXH_HD uint32_t mul_cov(uint32_t c, uint32_t x) {
  if (x == 0) return 0;
  if (x == 255) return c;
  uint32_t r = (ch_r(c) * x + 127u) / 255u;
  uint32_t g = (ch_g(c) * x + 127u) / 255u;
  uint32_t b = (ch_b(c) * x + 127u) / 255u;
  uint32_t a = (ch_a(c) * x + 127u) / 255u;
  return pack(r, g, b, a);
}

// Coverage factor of a trapezoid edge pixel: `*`(ColorRGBX, float32) takes
// round(area*255) through a 64-bit conversion; area == 0 gives transparent.
PXH_HD uint32_t area_factor(float area) {
  if (area == 0.0f) return 0;
  uint32_t x = nim_f2u32(roundf(area * 255.0f));
  // x == 0 and x == 255 give the same results through mul_cov's shortcuts.
  return x;
}

PXH_HD uint32_t blend_alpha(uint32_t ba, uint32_t sa) {  // blends.nim:41-43
  return (sa + ((ba * ((255u - sa) & 255u)) / 255u)) & 255u;
}

PXH_HD uint32_t screen_c(uint32_t b, uint32_t s) {  // blends.nim:45-46
  return (uint32_t)((int32_t)(b + s) - (int32_t)((b * s) / 255u)) & 255u;
}

PXH_HD uint32_t hard_light_c(uint32_t bc, uint32_t ba, uint32_t sc, uint32_t sa) {  // 48-58
  if (sc * 2u <= sa) return ((2u * sc * bc + sc * (255u - ba) + bc * (255u - sa)) / 255u) & 255u;
  return screen_c(bc, sc);
}

// n / d for n < 2^24 and quotients below 2^22 (d >= 1).  On the device: the
// float quotient through the hardware reciprocal (relative error far below
// 2^-20, so the truncation is off by at most one), then one remainder
// correction, which makes it exact.
PXH_HD uint32_t udiv_small(uint32_t n, uint32_t d) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t q = (uint32_t)((float)n * __builtin_amdgcn_rcpf((float)d));
  const int32_t r = (int32_t)(n - q * d);
  if (r < 0) q--;
  else if (r >= (int32_t)d) q++;
  return q;
#else
  return n / d;
#endif
}

// chroma ColorRGBX -> ColorRGBA (un-premultiply).  The factor lies in
// [256, 65025], where truncating to uint32 is the 64-bit conversion's result.
PXH_HD uint32_t unpremultiply(uint32_t c) {
  uint32_t a = ch_a(c);
  if (a == 0 || a == 255) return c;
  uint32_t mul = (uint32_t)((255.0f / (float)a) * 255.0f);
  return pack((ch_r(c) * mul + 127u) / 255u, (ch_g(c) * mul + 127u) / 255u,
              (ch_b(c) * mul + 127u) / 255u, a);
}

// chroma ColorRGBA -> ColorRGBX (premultiply), pinned by test_images.nim:204-228.
PXH_HD uint32_t premultiply(uint32_t s) {
  uint32_t a = ch_a(s);
  if (a == 255) return s;
  if (a == 0) return 0;
  return pack((ch_r(s) * a + 127u) / 255u, (ch_g(s) * a + 127u) / 255u,
              (ch_b(s) * a + 127u) / 255u, a);
}

// blends.nim:18-39 alphaFix on straight colours.
PXH_HD uint32_t alpha_fix(uint32_t b, uint32_t s, uint32_t m) {
  uint32_t sa = ch_a(s), ba = ch_a(b);
  uint32_t t0 = sa * (255u - ba), t1 = sa * ba, t2 = (255u - sa) * ba;
  uint32_t r = t0 * ch_r(s) + t1 * ch_r(m) + t2 * ch_r(b);
  uint32_t g = t0 * ch_g(s) + t1 * ch_g(m) + t2 * ch_g(b);
  uint32_t bb = t0 * ch_b(s) + t1 * ch_b(m) + t2 * ch_b(b);
  uint32_t a = sa + ba * (255u - sa) / 255u;
  if (a == 0) return 0;
  // r, g, bb < 255 * 255 * (a + 1) < 2^24, so each r / a < 2^17
  return pack(udiv_small(r, a) / 255u, udiv_small(g, a) / 255u, udiv_small(bb, a) / 255u, a);
}

// ---- float32 colour modes (chroma Color, straight alpha) ----
struct FColor { float r, g, b, a; };

// k / 255.0f for an integer k in [0, 255], rounded like the division: the
// product with the rounded reciprocal, corrected by one fma step.  Checked
// equal to the division for all 256 inputs.
PXH_HD float byte_div255(uint32_t k) {
  const float x = (float)k, inv = 1.0f / 255.0f;
  const float q = x * inv;
  return fmaf(fmaf(-q, 255.0f, x), inv, q);
}

PXH_HD FColor to_fcolor(uint32_t c) {
  uint32_t s = unpremultiply(c);
  FColor f = {byte_div255(ch_r(s)), byte_div255(ch_g(s)), byte_div255(ch_b(s)),
              byte_div255(ch_a(s))};
  return f;
}
PXH_HD uint32_t unit_u8(float v) {
  if (!(v > 0.0f)) v = 0.0f;
  if (v > 1.0f) v = 1.0f;
  return (uint32_t)roundf(v * 255.0f);
}
PXH_HD uint32_t from_fcolor(FColor f) {
  return premultiply(pack(unit_u8(f.r), unit_u8(f.g), unit_u8(f.b), unit_u8(f.a)));
}
PXH_HD FColor alpha_fix_f(FColor b, FColor s, FColor m) {
  FColor r;
  r.a = s.a + b.a * (1.0f - s.a);
  if (r.a == 0.0f) {
    FColor z = {0.0f, 0.0f, 0.0f, 0.0f};
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
PXH_HD float soft_light_f(float b, float s) {
  if (s <= 0.5f) return b - (1.0f - 2.0f * s) * b * (1.0f - b);
  float d = (b <= 0.25f) ? ((16.0f * b - 12.0f) * b + 4.0f) * b : sqrtf(b);
  return b + (2.0f * s - 1.0f) * (d - b);
}
PXH_HD float lum_f(FColor c) { return 0.3f * c.r + 0.59f * c.g + 0.11f * c.b; }
PXH_HD FColor clip_color_f(FColor c) {
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
PXH_HD FColor set_lum_f(FColor c, float l) {
  float d = l - lum_f(c);
  c.r += d;
  c.g += d;
  c.b += d;
  return clip_color_f(c);
}
PXH_HD float sat_f(FColor c) { return fmaxf(fmaxf(c.r, c.g), c.b) - fminf(fminf(c.r, c.g), c.b); }
PXH_HD FColor set_sat_f(FColor c, float s) {
  float v[3] = {c.r, c.g, c.b};
  int imax = 0, imin = 0;
  for (int i = 1; i < 3; i++) {
    if (v[i] > v[imax]) imax = i;
    if (v[i] < v[imin]) imin = i;
  }
  if (imax == imin) {
    c.r = c.g = c.b = 0.0f;
    return c;
  }
  int imid = 3 - imax - imin;
  float cmax = v[imax], cmin = v[imin], cmid = v[imid];
  float o[3];
  o[imid] = (cmid - cmin) * s / (cmax - cmin);
  o[imax] = s;
  o[imin] = 0.0f;
  c.r = o[0];
  c.g = o[1];
  c.b = o[2];
  return c;
}
PXH_HD uint32_t blend_float(int32_t mode, uint32_t bb, uint32_t ss) {
  FColor b = to_fcolor(bb), s = to_fcolor(ss), m = b;
  if (mode == kSoftLight) {
    m.r = soft_light_f(b.r, s.r);
    m.g = soft_light_f(b.g, s.g);
    m.b = soft_light_f(b.b, s.b);
  } else if (mode == kHue) {
    m = set_lum_f(set_sat_f(s, sat_f(b)), lum_f(b));
  } else if (mode == kSaturation) {
    m = set_lum_f(set_sat_f(b, sat_f(s)), lum_f(b));
  } else if (mode == kColor) {
    m = set_lum_f(s, lum_f(b));
  } else {  // kLuminosity
    m = set_lum_f(b, lum_f(s));
  }
  return from_fcolor(alpha_fix_f(b, s, m));
}

// blender(mode)(backdrop, source), blends.nim:60-299.
PXH_HD uint32_t blend(int32_t mode, uint32_t b, uint32_t s) {
  uint32_t br = ch_r(b), bg = ch_g(b), bb = ch_b(b), ba = ch_a(b);
  uint32_t sr = ch_r(s), sg = ch_g(s), sb = ch_b(s), sa = ch_a(s);
  switch (mode) {
    case kNormal: {  // 60-70
      if (ba == 0 || sa == 255) return s;
      if (sa == 0) return b;
      uint32_t k = 255u - sa;
      return pack(sr + (br * k) / 255u, sg + (bg * k) / 255u, sb + (bb * k) / 255u,
                  blend_alpha(ba, sa));
    }
    case kDarken:
    case kLighten: {  // 72-84, 128-140
      uint32_t ia = 255u - ba, is = 255u - sa;
      uint32_t xr = br + (ia * sr) / 255u, yr = sr + (is * br) / 255u;
      uint32_t xg = bg + (ia * sg) / 255u, yg = sg + (is * bg) / 255u;
      uint32_t xb = bb + (ia * sb) / 255u, yb = sb + (is * bb) / 255u;
      if (mode == kDarken)
        return pack(xr < yr ? xr : yr, xg < yg ? xg : yg, xb < yb ? xb : yb, blend_alpha(ba, sa));
      return pack(xr > yr ? xr : yr, xg > yg ? xg : yg, xb > yb ? xb : yb, blend_alpha(ba, sa));
    }
    case kMultiply: {  // 86-99
      uint32_t ia = 255u - ba, is = 255u - sa;
      return pack((ia * sr + is * br + br * sr) / 255u, (ia * sg + is * bg + bg * sg) / 255u,
                  (ia * sb + is * bb + bb * sb) / 255u, blend_alpha(ba, sa));
    }
    case kColorBurn:
    case kColorDodge: {  // 111-126, 158-173
      uint32_t bs = unpremultiply(b), ssr = unpremultiply(s);
      uint32_t m = 0;
      for (int k = 0; k < 3; k++) {
        uint32_t x = (bs >> (8 * k)) & 255u, y = (ssr >> (8 * k)) & 255u, v;
        if (mode == kColorBurn) {
          if (x == 255) v = 255;
          else if (y == 0) v = 0;
          else {
            uint32_t q = udiv_small(255u * (255u - x), y);
            v = (255u - (q < 255u ? q : 255u)) & 255u;
          }
        } else {
          if (x == 0) v = 0;
          else if (y == 255) v = 255;
          else {
            uint32_t q = udiv_small(255u * x, 255u - y);
            v = q < 255u ? q : 255u;
          }
        }
        m |= v << (8 * k);
      }
      return premultiply(alpha_fix(bs, ssr, m));
    }
    case kScreen:  // 142-146
      return pack(screen_c(br, sr), screen_c(bg, sg), screen_c(bb, sb), blend_alpha(ba, sa));
    case kOverlay:  // 175-179
      return pack(hard_light_c(sr, sa, br, ba), hard_light_c(sg, sa, bg, ba),
                  hard_light_c(sb, sa, bb, ba), blend_alpha(ba, sa));
    case kHardLight:  // 184-188
      return pack(hard_light_c(br, ba, sr, sa), hard_light_c(bg, ba, sg, sa),
                  hard_light_c(bb, ba, sb, sa), blend_alpha(ba, sa));
    case kDifference: {  // 190-204
      uint32_t o = 0;
      for (int k = 0; k < 3; k++) {
        uint32_t x = (b >> (8 * k)) & 255u, y = (s >> (8 * k)) & 255u;
        uint32_t p = x * sa, q = y * ba;
        uint32_t mn = p < q ? p : q;
        o |= ((uint32_t)((int32_t)((x + y) & 255u) - 2 * (int32_t)(mn / 255u)) & 255u) << (8 * k);
      }
      return o | (blend_alpha(ba, sa) << 24);
    }
    case kExclusion: {  // 206-213
      uint32_t o = 0;
      for (int k = 0; k < 3; k++) {
        uint32_t x = (b >> (8 * k)) & 255u, y = (s >> (8 * k)) & 255u;
        int32_t v = (int32_t)(x + y) - (int32_t)((2u * x * y) / 255u);
        o |= ((uint32_t)(v > 0 ? v : 0) & 255u) << (8 * k);
      }
      return o | (blend_alpha(ba, sa) << 24);
    }
    case kSoftLight:
    case kHue:
    case kSaturation:
    case kColor:
    case kLuminosity:
      return blend_float(mode, b, s);
    case kMask:  // 227-232
      return pack((br * sa) / 255u, (bg * sa) / 255u, (bb * sa) / 255u, (ba * sa) / 255u);
    case kOverwrite:
      return s;
    case kSubtractMask: {  // 234-239
      uint32_t a = (ba * (255u - sa)) / 255u;
      return pack((br * a) / 255u, (bg * a) / 255u, (bb * a) / 255u, a);
    }
    case kExcludeMask: {  // 241-246: scales the source colour
      uint32_t a = (ba > sa ? ba : sa) - (ba < sa ? ba : sa);
      return pack((sr * a) / 255u, (sg * a) / 255u, (sb * a) / 255u, a);
    }
    default:
      return b;
  }
}

}  // namespace pxh
