// pxh_draw.hip -- transformed image draws on the device.
//
//   minify_kernel        minifyBy2, one power (images.nim:168-228)
//   magnify_kernel       magnifyBy2 (images.nim:230-259), for the integer
//                        composite after a magnification
//   draw_smooth_kernel   drawSmooth (images.nim:531-634): bilinear samples
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
__device__ __forceinline__ uint32_t src_fetch(const DrawSrc& s, int64_t x, int64_t y, bool wrapped) {
  const int64_t W = (int64_t)s.w << s.mag, H = (int64_t)s.h << s.mag;
  if (wrapped) {
    // Nim's mod keeps the dividend's sign; pixie indexes the data with the
    // flat index w * (y mod h) + (x mod w) unchecked: inside the buffer it
    // reads that pixel, outside it reads transparent here
    const int64_t i = W * (y % H) + (x % W);
    if (i < 0 || i >= W * H) return 0u;
    x = i % W;
    y = i / W;
  } else if (x < 0 || x >= W || y < 0 || y >= H) {
    return 0u;
  }
  return s.px[(size_t)(y >> s.mag) * s.w + (size_t)(x >> s.mag)];
}

__device__ uint32_t smooth_sample(const DrawSrc& s, float x, float y, bool wrapped) {
  const float fx = floorf(x), fy = floorf(y);
  const int64_t x0 = nim_f2i(fx), y0 = nim_f2i(fy);
  const float xf = x - fx, yf = y - fy;
  const uint32_t p00 = src_fetch(s, x0, y0, wrapped), p10 = src_fetch(s, x0 + 1, y0, wrapped);
  const uint32_t p01 = src_fetch(s, x0, y0 + 1, wrapped), p11 = src_fetch(s, x0 + 1, y0 + 1, wrapped);
  uint32_t top = p00, bottom = p01;
  if (xf > 0.0f && p00 != p10) top = mix_rgbx(p00, p10, xf);
  if (xf > 0.0f && p01 != p11) bottom = mix_rgbx(p01, p11, xf);
  if (yf != 0.0f && top != bottom) return mix_rgbx(top, bottom, yf);
  return top;
}

// ---------------------------------------------------------------------------
// drawSmooth.  A workgroup owns kSmoothRows rows.  Each row's sample
// positions are the running sum srcPos += dx from its first pixel, which
// float32 only reproduces when summed in order: in wave 0, lane 2r sums the
// x component of row r and lane 2r + 1 the y component, kSmoothChunk
// positions at a time into LDS; then all 256 threads sample and blend those
// positions.  Rows outside [yStart, yEnd) and columns outside a row's span
// are left alone, or cleared for MaskBlend (images.nim:560-561, 612-626).
// ---------------------------------------------------------------------------
constexpr int kSmoothRows = 32;
constexpr int kSmoothChunk = 64;

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

__global__ void __launch_bounds__(256) draw_smooth_kernel(SmoothArgs a) {
  __shared__ float2 pos[kSmoothRows][kSmoothChunk];
  __shared__ int32_t xs[kSmoothRows], xe[kSmoothRows];
  __shared__ int32_t nChunks;
  const int32_t tid = threadIdx.x;
  const int32_t row0 = blockIdx.x * kSmoothRows;
  const bool mask = a.mode == kMask;
  // per-row span: xs = xe marks a row that is left alone
  if (tid < kSmoothRows) {
    const int32_t y = row0 + tid;
    int32_t s = 0, e = 0;
    if (y < a.ah && y >= a.yStart && y < a.yEnd) {
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
    } else if (y < a.ah && mask) {
      s = a.aw;  // a MaskBlend row outside [yStart, yEnd): cleared whole
      e = 0;
    }
    xs[tid] = s;
    xe[tid] = e;
  }
  __syncthreads();
  if (tid == 0) {
    int32_t n = 0;
    for (int r = 0; r < kSmoothRows; r++) {
      const int32_t len = xe[r] - xs[r];
      if (len > 0) n = max(n, (len + kSmoothChunk - 1) / kSmoothChunk);
    }
    nChunks = n;
  }
  // MaskBlend clears each row outside its span; a row whose span is empty
  // (xEnd == xStart) is skipped whole, as pixie's `continue` does
  if (mask) {
    for (int r = 0; r < kSmoothRows; r++) {
      const int32_t y = row0 + r;
      if (y >= a.ah || xe[r] == xs[r]) continue;
      uint32_t* row = a.dst + (size_t)y * a.aw;
      for (int32_t x = tid; x < a.aw; x += 256)
        if (x < xs[r] || x >= xe[r]) row[x] = 0u;
    }
  }
  __syncthreads();
  // the running sums: lane 2r / 2r + 1 of wave 0 owns row r's x / y
  float run = 0.0f, step = 0.0f;
  const int32_t sr = tid >> 1;
  if (tid < 2 * kSmoothRows) {
    const int32_t y = row0 + sr;
    const int c = tid & 1;
    // srcPos = p + dx * xStart + dy * y, then - h (images.nim:593-594)
    run = a.p[c] + a.dx[c] * (float)xs[sr] + a.dy[c] * (float)y;
    run = run - 0.5f;
    step = a.dx[c];
  }
  for (int32_t ch = 0; ch < nChunks; ch++) {
    if (tid < 2 * kSmoothRows) {
      const int32_t base = ch * kSmoothChunk;
      const int32_t len = xe[sr] - xs[sr];
      float* dstp = (tid & 1) ? &pos[sr][0].y : &pos[sr][0].x;
      for (int j = 0; j < kSmoothChunk; j++) {
        if (base + j < len) {
          dstp[2 * j] = run;
          run += step;
        }
      }
    }
    __syncthreads();
    for (int k = tid; k < kSmoothRows * kSmoothChunk; k += 256) {
      const int32_t r = k / kSmoothChunk, j = k % kSmoothChunk;
      const int32_t y = row0 + r;
      const int32_t x = xs[r] + ch * kSmoothChunk + j;
      if (y >= a.ah || x >= xe[r]) continue;
      const float2 sp = pos[r][j];
      const uint32_t s = smooth_sample(a.src, sp.x, sp.y, false);
      uint32_t* d = a.dst + (size_t)y * a.aw + x;
      *d = a.mode == kOverwrite ? s : blend(a.mode, *d, s);
    }
    __syncthreads();
  }
}

int launch_draw_smooth(const SmoothArgs& a, hipStream_t s) {
  const int grid = (a.ah + kSmoothRows - 1) / kSmoothRows;
  hipLaunchKernelGGL(draw_smooth_kernel, dim3(grid), dim3(256), 0, s, a);
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
  const uint32_t s = smooth_sample(a.src, sx - 0.5f, sy - 0.5f, a.wrapped != 0);
  uint32_t* d = a.dst + (size_t)y * a.aw + x;
  *d = blend(a.mode, *d, s);
}

int launch_draw_correct(const CorrectArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(draw_correct_kernel, dim3((a.aw + 255) / 256, a.ah), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

}  // namespace pxh
