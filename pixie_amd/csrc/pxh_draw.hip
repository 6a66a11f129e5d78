// pxh_draw.hip -- transformed image draws on the device.
//
//   minify_kernel        minifyBy2, one power (images.nim:168-228)
//   magnify_kernel       magnifyBy2 (images.nim:230-259), for the integer
//                        composite after a magnification
//   smooth_*_kernel      drawSmooth (images.nim:531-634): bilinear samples
//                        (getRgbaSmooth, images.nim:367-403) along each
//                        row's span of the transformed image, blended with
//                        any of the 20 modes
//   draw_correct_kernel  drawCorrect (images.nim:405-445), the tiled draw:
//                        every pixel sampled independently, wrapped
//
// Sources magnified by 2^mag are read in place (pixel (x, y) of the
// magnified image is pixel (x >> mag, y >> mag)), so magnifyBy2 is never
// materialised for sampling.  Float math follows pixie's float32 operation
// order (no contraction; see the Makefile).  Memory-bound gather / blend
// work: no MFMA.
#include <hip/hip_runtime.h>

#include "pxh_draw.h"
#include "pxh_pixel.h"

namespace pxh {

// mix(ColorRGBX, ColorRGBX, t), common.nim:59-65
__device__ __forceinline__ uint32_t mix_rgbx(uint32_t a, uint32_t b, float t) {
  const uint32_t x = nim_f2u32(roundf(t * 255.0f));
  const uint32_t y = 255u - x;
  return pack((ch_r(a) * y + ch_r(b) * x + 127u) / 255u, (ch_g(a) * y + ch_g(b) * x + 127u) / 255u,
              (ch_b(a) * y + ch_b(b) * x + 127u) / 255u, (ch_a(a) * y + ch_a(b) * x + 127u) / 255u);
}

// `*`(ColorRGBX, float32) for the constant opacities minifyBy2 uses
// (common.nim:67-77): x = round(0.5 * 255) = 128, round(0.25 * 255) = 64
__device__ __forceinline__ uint32_t mul_x(uint32_t c, uint32_t x) {
  return pack((ch_r(c) * x + 127u) / 255u, (ch_g(c) * x + 127u) / 255u, (ch_b(c) * x + 127u) / 255u,
              (ch_a(c) * x + 127u) / 255u);
}

// ---------------------------------------------------------------------------
// minifyBy2, one power: one thread per destination pixel
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) minify_kernel(MinifyArgs a) {
  const int32_t rw = (a.w + 1) / 2;
  const int32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= rw) return;
  const int32_t ew = a.w / 2, eh = a.h / 2;
  const uint32_t* top = a.src + (size_t)a.w * (2 * y);
  uint32_t out;
  if (y < eh) {
    const uint32_t* bot = top + a.w;
    if (x < ew) {
      const uint32_t p = top[2 * x], q = top[2 * x + 1], r = bot[2 * x + 1], s = bot[2 * x];
      out = pack((ch_r(p) + ch_r(q) + ch_r(r) + ch_r(s) + 2u) / 4u,
                 (ch_g(p) + ch_g(q) + ch_g(r) + ch_g(s) + 2u) / 4u,
                 (ch_b(p) + ch_b(q) + ch_b(r) + ch_b(s) + 2u) / 4u,
                 (ch_a(p) + ch_a(q) + ch_a(r) + ch_a(s) + 2u) / 4u);
    } else {  // odd width: the last column
      out = mul_x(mix_rgbx(top[a.w - 1], bot[a.w - 1], 0.5f), 128u);
    }
  } else {  // odd height: the last row
    out = x < ew ? mul_x(mix_rgbx(top[2 * x], top[2 * x + 1], 0.5f), 128u) : mul_x(top[a.w - 1], 64u);
  }
  a.dst[(size_t)rw * y + x] = out;
}

int launch_minify(const MinifyArgs& a, hipStream_t s) {
  const int32_t rw = (a.w + 1) / 2, rh = (a.h + 1) / 2;
  hipLaunchKernelGGL(minify_kernel, dim3((rw + 255) / 256, rh), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

__global__ void __launch_bounds__(256) magnify_kernel(MagnifyArgs a) {
  const int32_t ow = a.w << a.power;
  const int32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= ow) return;
  a.dst[(size_t)ow * y + x] = a.src[(size_t)a.w * (y >> a.power) + (x >> a.power)];
}

int launch_magnify(const MagnifyArgs& a, hipStream_t s) {
  const int32_t ow = a.w << a.power, oh = a.h << a.power;
  hipLaunchKernelGGL(magnify_kernel, dim3((ow + 255) / 256, oh), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// ---------------------------------------------------------------------------
// getRgbaSmooth over a source magnified by 2^mag in place
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t src_at(const DrawSrc& s, int64_t x, int64_t y) {
  return s.px[(size_t)(y >> s.mag) * s.w + (size_t)(x >> s.mag)];
}

// the bilinear mix of getRgbaSmooth once its four pixels are loaded
__device__ __forceinline__ uint32_t bilinear(uint32_t p00, uint32_t p10, uint32_t p01, uint32_t p11,
                                             float xf, float yf) {
  uint32_t top = p00, bottom = p01;
  if (xf > 0.0f && p00 != p10) top = mix_rgbx(p00, p10, xf);
  if (xf > 0.0f && p01 != p11) bottom = mix_rgbx(p01, p11, xf);
  if (yf != 0.0f && top != bottom) return mix_rgbx(top, bottom, yf);
  return top;
}

// floor(v) as an int32 pixel coordinate, or -2 when it lies outside
// [-2, 2^30): the virtual source is at most 2^30 on a side (the host checks),
// so such a coordinate and its neighbour read transparent, as pixie's
// out-of-bounds reads do (NaN included)
__device__ __forceinline__ int32_t pixel_coord(float fv) {
  return (fv >= -2.0f && fv < 1073741824.0f) ? (int32_t)fv : -2;
}

// getRgbaSmooth (images.nim:367-403), unwrapped.  The four reads are
// branch-free (an out-of-bounds read loads pixel 0 and is discarded) so that
// they issue together.
__device__ __forceinline__ uint32_t smooth_sample(const DrawSrc& s, float x, float y) {
  const float fx = floorf(x), fy = floorf(y);
  const float xf = x - fx, yf = y - fy;
  const int32_t x0 = pixel_coord(fx), y0 = pixel_coord(fy);
  const uint32_t W = (uint32_t)s.w << s.mag, H = (uint32_t)s.h << s.mag;
  const bool vx0 = (uint32_t)x0 < W, vx1 = (uint32_t)(x0 + 1) < W;
  const bool vy0 = (uint32_t)y0 < H, vy1 = (uint32_t)(y0 + 1) < H;
  const uint32_t cx0 = vx0 ? (uint32_t)x0 >> s.mag : 0u, cx1 = vx1 ? (uint32_t)(x0 + 1) >> s.mag : 0u;
  const size_t r0 = vy0 ? (size_t)((uint32_t)y0 >> s.mag) * s.w : 0, r1 = vy1 ? (size_t)((uint32_t)(y0 + 1) >> s.mag) * s.w : 0;
  uint32_t p00 = s.px[r0 + cx0], p10 = s.px[r0 + cx1], p01 = s.px[r1 + cx0], p11 = s.px[r1 + cx1];
  p00 = (vx0 && vy0) ? p00 : 0u;
  p10 = (vx1 && vy0) ? p10 : 0u;
  p01 = (vx0 && vy1) ? p01 : 0u;
  p11 = (vx1 && vy1) ? p11 : 0u;
  uint32_t top = p00, bottom = p01;
  if (xf > 0.0f && p00 != p10) top = mix_rgbx(p00, p10, xf);
  if (xf > 0.0f && p01 != p11) bottom = mix_rgbx(p01, p11, xf);
  if (yf != 0.0f && top != bottom) return mix_rgbx(top, bottom, yf);
  return top;
}

// Nim's `mod` (the sign of the dividend) of v and v + 1 by n > 0, with one
// division: 32-bit when v fits, else 64-bit
__device__ __forceinline__ void mod_pair(int64_t v, int64_t n, int64_t* m0, int64_t* m1) {
  int64_t r;
  if (v >= -(1ll << 31) && v < (1ll << 31) && n < (1ll << 31))
    r = (int64_t)((int32_t)v % (int32_t)n);
  else
    r = v % n;
  *m0 = r;
  if (v >= 0)
    *m1 = (r + 1 == n) ? 0 : r + 1;
  else
    *m1 = (r == 0) ? 1 - n : r + 1;  // v + 1 <= 0
}

// the wrapped read of getRgbaSmooth (images.nim:379-383): pixie indexes the
// data with the flat index w * (y mod h) + (x mod w) unchecked; inside the
// buffer that is a pixel (usually (x mod w, y mod h) itself), outside it the
// read is undefined in pixie and transparent here.  Returns the stored
// pixel's index and whether the read is inside the buffer; no load.
__device__ __forceinline__ size_t wrap_index(const DrawSrc& s, int64_t xm, int64_t ym, bool* valid) {
  if (xm >= 0 && ym >= 0) {
    *valid = true;
    return (size_t)(ym >> s.mag) * s.w + (size_t)(xm >> s.mag);
  }
  const int64_t W = (int64_t)s.w << s.mag, H = (int64_t)s.h << s.mag;
  const int64_t i = W * ym + xm;
  *valid = i >= 0 && i < W * H;
  if (!*valid) return 0;
  return (size_t)((i / W) >> s.mag) * s.w + (size_t)((i % W) >> s.mag);
}

__device__ uint32_t wrapped_sample(const DrawSrc& s, float x, float y) {
  const float fx = floorf(x), fy = floorf(y);
  const int64_t x0 = nim_f2i(fx), y0 = nim_f2i(fy);
  const float xf = x - fx, yf = y - fy;
  const int64_t W = (int64_t)s.w << s.mag, H = (int64_t)s.h << s.mag;
  int64_t xa, xb, ya, yb;
  mod_pair(x0, W, &xa, &xb);
  mod_pair(y0, H, &ya, &yb);
  bool v00, v10, v01, v11;
  const size_t i00 = wrap_index(s, xa, ya, &v00), i10 = wrap_index(s, xb, ya, &v10);
  const size_t i01 = wrap_index(s, xa, yb, &v01), i11 = wrap_index(s, xb, yb, &v11);
  uint32_t p00 = s.px[i00], p10 = s.px[i10], p01 = s.px[i01], p11 = s.px[i11];
  p00 = v00 ? p00 : 0u;
  p10 = v10 ? p10 : 0u;
  p01 = v01 ? p01 : 0u;
  p11 = v11 ? p11 : 0u;
  return bilinear(p00, p10, p01, p11, xf, yf);
}

// ---------------------------------------------------------------------------
// drawSmooth in three steps.  Each row's sample positions are the running sum
// srcPos += dx from its first pixel, which float32 only reproduces when summed
// in order, so:
//   smooth_rows_kernel    one thread per row: the row's span from the
//                         perimeter, then the running sum along it, storing
//                         every kSmoothRun-th position (the start of each run)
//   smooth_runs_kernel    one thread per run of kSmoothRun pixels: continues
//                         the sum from the run's stored start, samples and
//                         blends
//   smooth_clear_kernel   MaskBlend only: clears the rows outside [yStart,
//                         yEnd) and each row outside its span, skipping rows
//                         whose span is empty (images.nim:560-561, 598-626)
// ---------------------------------------------------------------------------

// bumpy intersects(Line, Segment) for the horizontal line through
// (-1000, ly) and (1000, ly); u is the segment parameter
__device__ __forceinline__ bool scan_hit(const float* seg, float ly, float* atx, float* aty) {
  const float lax = -1000.0f, lay = ly, lbx = 1000.0f, lby = ly;
  const float s1x = lbx - lax, s1y = lby - lay;
  const float s2x = seg[2] - seg[0], s2y = seg[3] - seg[1];
  const float den = (-s2x * s1y + s1x * s2y);
  const float num = s1x * (lay - seg[1]) - s1y * (lax - seg[0]);
  const float u = num / den;
  if (u >= 0.0f && u <= 1.0f) {
    *atx = seg[0] + u * s2x;
    *aty = seg[1] + u * s2y;
    return true;
  }
  return false;
}

__global__ void __launch_bounds__(64) smooth_rows_kernel(SmoothArgs a) {
  const int32_t y = blockIdx.x * 64 + threadIdx.x;
  if (y >= a.ah) return;
  int32_t s = 0, e = 0;
  if (y >= a.yStart && y < a.yEnd) {
    float xMin = (float)a.aw, xMax = 0.0f;
    for (int yo = 0; yo < 2; yo++) {
      const float ly = (float)y + (float)yo;
      for (int i = 0; i < 4; i++) {
        float ax, ay;
        if (scan_hit(a.per[i], ly, &ax, &ay) && !(a.per[i][2] == ax && a.per[i][3] == ay)) {
          xMin = nim_minf(xMin, ax);
          xMax = nim_maxf(xMax, ax);
        }
      }
    }
    int64_t x0 = nim_f2i(floorf(xMin)), x1 = nim_f2i(ceilf(xMax));
    x0 = x0 < 0 ? 0 : (x0 > a.aw ? a.aw : x0);
    x1 = x1 < 0 ? 0 : (x1 > a.aw ? a.aw : x1);
    s = (int32_t)x0;
    e = (int32_t)x1;
  } else if (a.mode == kMask) {
    s = a.aw;  // cleared whole: a span that ends before it starts
    e = 0;
  }
  a.span[y] = make_int2(s, e);
  // srcPos = p + dx * xStart + dy * y, then - h (images.nim:593-594)
  float sx = a.p[0] + a.dx[0] * (float)s + a.dy[0] * (float)y;
  float sy = a.p[1] + a.dx[1] * (float)s + a.dy[1] * (float)y;
  sx = sx - 0.5f;
  sy = sy - 0.5f;
  float2* starts = a.starts + (size_t)y * a.nRuns;
  const int32_t len = e - s;
  const float ddx = a.dx[0], ddy = a.dx[1];
  int32_t j = 0;
  for (; j + kSmoothRun <= len; j += kSmoothRun) {
    starts[j / kSmoothRun] = make_float2(sx, sy);
#pragma unroll
    for (int k = 0; k < kSmoothRun; k++) {
      sx += ddx;
      sy += ddy;
    }
  }
  if (j < len) starts[j / kSmoothRun] = make_float2(sx, sy);
}

__device__ __noinline__ uint32_t blend_call(int32_t mode, uint32_t b, uint32_t s) { return blend(mode, b, s); }

// A wave takes 64 consecutive runs of a row (64 * kSmoothRun pixels).  Each
// lane continues the running sum along its own run and leaves the positions
// in LDS; the wave then works through the pixels in order, lane l taking
// pixel 64 j + l at step j, so that neighbouring lanes read neighbouring
// source and destination pixels.  Three straight-line phases let the loads
// issue together: every source read (pixels past the span read clamped
// in-bounds addresses and are discarded), the bilinear mixes, then the
// destination reads, blends and writes.
__global__ void __launch_bounds__(64) smooth_runs_kernel(SmoothArgs a) {
  __shared__ float2 pos[64 * kSmoothRun];
  const int32_t y = blockIdx.y;
  const int32_t lane = threadIdx.x;
  const int2 sp = a.span[y];
  const int32_t xw = sp.x + blockIdx.x * 64 * kSmoothRun;  // the wave's first pixel
  if (xw >= sp.y) return;                                  // uniform over the block
  const int32_t run = blockIdx.x * 64 + lane;
  float2 p = run < a.nRuns ? a.starts[(size_t)y * a.nRuns + run] : make_float2(0.0f, 0.0f);
#pragma unroll
  for (int k = 0; k < kSmoothRun; k++) {
    pos[lane * kSmoothRun + k] = p;
    p.x += a.dx[0];
    p.y += a.dx[1];
  }
  __syncthreads();
  const DrawSrc src = a.src;
  const uint32_t W = (uint32_t)src.w << src.mag, H = (uint32_t)src.h << src.mag;
  uint32_t q[kSmoothRun][4];
  float xf[kSmoothRun], yf[kSmoothRun];
#pragma unroll
  for (int j = 0; j < kSmoothRun; j++) {
    const float2 v = pos[j * 64 + lane];
    const float fx = floorf(v.x), fy = floorf(v.y);
    xf[j] = v.x - fx;
    yf[j] = v.y - fy;
    const int32_t px0 = pixel_coord(fx), py0 = pixel_coord(fy);
    const bool vx0 = (uint32_t)px0 < W, vx1 = (uint32_t)(px0 + 1) < W;
    const bool vy0 = (uint32_t)py0 < H, vy1 = (uint32_t)(py0 + 1) < H;
    const uint32_t c0 = vx0 ? (uint32_t)px0 >> src.mag : 0u, c1 = vx1 ? (uint32_t)(px0 + 1) >> src.mag : 0u;
    const size_t r0 = vy0 ? (size_t)((uint32_t)py0 >> src.mag) * src.w : 0;
    const size_t r1 = vy1 ? (size_t)((uint32_t)(py0 + 1) >> src.mag) * src.w : 0;
    const uint32_t a00 = src.px[r0 + c0], a10 = src.px[r0 + c1];
    const uint32_t a01 = src.px[r1 + c0], a11 = src.px[r1 + c1];
    q[j][0] = (vx0 && vy0) ? a00 : 0u;
    q[j][1] = (vx1 && vy0) ? a10 : 0u;
    q[j][2] = (vx0 && vy1) ? a01 : 0u;
    q[j][3] = (vx1 && vy1) ? a11 : 0u;
  }
  uint32_t smp[kSmoothRun];
#pragma unroll
  for (int j = 0; j < kSmoothRun; j++) smp[j] = bilinear(q[j][0], q[j][1], q[j][2], q[j][3], xf[j], yf[j]);
  uint32_t* row = a.dst + (size_t)y * a.aw;
  const int32_t lim = sp.y;  // pixels at or past the span's end are not touched
  if (a.mode == kOverwrite) {
#pragma unroll
    for (int j = 0; j < kSmoothRun; j++)
      if (xw + j * 64 + lane < lim) row[xw + j * 64 + lane] = smp[j];
    return;
  }
  uint32_t d[kSmoothRun];
#pragma unroll
  for (int j = 0; j < kSmoothRun; j++) {
    const int32_t x = xw + j * 64 + lane;
    d[j] = row[x < lim ? x : xw];
  }
  if (a.mode == kNormal) {
#pragma unroll
    for (int j = 0; j < kSmoothRun; j++) d[j] = blend(kNormal, d[j], smp[j]);
  } else {
#pragma unroll
    for (int j = 0; j < kSmoothRun; j++) d[j] = blend_call(a.mode, d[j], smp[j]);
  }
#pragma unroll
  for (int j = 0; j < kSmoothRun; j++)
    if (xw + j * 64 + lane < lim) row[xw + j * 64 + lane] = d[j];
}

__global__ void __launch_bounds__(256) smooth_clear_kernel(SmoothArgs a) {
  const int32_t y = blockIdx.y;
  const int32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= a.aw) return;
  const int2 sp = a.span[y];
  if (sp.x == sp.y) return;  // an empty span: the row is skipped whole
  if (x < sp.x || x >= sp.y) a.dst[(size_t)y * a.aw + x] = 0u;
}

int launch_draw_smooth(const SmoothArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(smooth_rows_kernel, dim3((a.ah + 63) / 64), dim3(64), 0, s, a);
  hipLaunchKernelGGL(smooth_runs_kernel, dim3((a.nRuns + 63) / 64, a.ah), dim3(64), 0, s, a);
  if (a.mode == kMask)
    hipLaunchKernelGGL(smooth_clear_kernel, dim3((a.aw + 255) / 256, a.ah), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// ---------------------------------------------------------------------------
// drawCorrect: every destination pixel, sampled at inverse * (x + h, y + h)
// - h, wrapped for drawTiled
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) draw_correct_kernel(CorrectArgs a) {
  const int32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= a.aw) return;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const float sx = a.inv[0] * px + a.inv[3] * py + a.inv[6];
  const float sy = a.inv[1] * px + a.inv[4] * py + a.inv[7];
  const uint32_t s = a.wrapped ? wrapped_sample(a.src, sx - 0.5f, sy - 0.5f)
                               : smooth_sample(a.src, sx - 0.5f, sy - 0.5f);
  uint32_t* d = a.dst + (size_t)y * a.aw + x;
  *d = a.mode == kNormal ? blend(kNormal, *d, s) : blend_call(a.mode, *d, s);
}

int launch_draw_correct(const CorrectArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(draw_correct_kernel, dim3((a.aw + 255) / 256, a.ah), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

}  // namespace pxh
