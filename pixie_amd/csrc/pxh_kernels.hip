// pxh_kernels.hip -- HIP kernels for gfx950 (MI355X) and their launchers.
//
//   fill_rows_kernel   one fillShapes call (paths.nim:1593-1912): a wave per
//                      (canvas row, column window) runs the row program and the fused
//                      coverage -> paint -> blend -> write-back epilogue
//   composite_kernel   blendRect (images.nim:468-529), all 20 blend modes
//   blur_x / blur_y    blur (images.nim:304-365), LDS-staged halo tiles,
//                      integer-exact
//   spread_x / _y      spread (images.nim:700-758)
//   opacity_kernel     applyOpacity (images.nim:261-277)
//
// Every kernel is memory-bound integer / float32 work: no MFMA.
#include <hip/hip_runtime.h>

#include "pxh_kernels.h"
#include "pxh_raster.h"

namespace pxh {

// ---------------------------------------------------------------------------
// fill: one wave (64 threads) per workgroup, grid-stride over (row, window)
// units
// ---------------------------------------------------------------------------
template <bool kA8>
__global__ void __launch_bounds__(64) fill_rows_kernel(FillArgs a) {
  extern __shared__ __align__(16) char smem[];
  const FillInfo& F = *a.F;
  const size_t fixedBytes = row_lds_fixed_bytes(a.ww);
  char* arr = a.capInLds ? smem + fixedBytes : a.gscratch + (size_t)blockIdx.x * row_cap_bytes(a.cap);
  RowLds L = row_lds_carve(smem, arr, a.cap, a.ww);
  for (int i = threadIdx.x; i < a.ww + 4; i += kWave) L.D[i] = 0;
  wave_sync();
  const bool mask = F.mode == kMask;
  const int32_t W = F.canvasW;
  const int32_t lo = mask ? 0 : F.xlo, hi = mask ? W : F.xhi;
  const int32_t x0 = lo & ~3;
  const int32_t nWin = hi > x0 ? (hi - x0 + a.ww - 1) / a.ww : 0;
  const int64_t nUnits = (int64_t)a.nRows * nWin;
  uint32_t ntouch = 0;
  // one work unit per (row, column window), grid-stride
  for (int64_t u = blockIdx.x; u < nUnits; u += gridDim.x) {
    const int32_t y = a.rowBase + (int32_t)(u / nWin);
    const int32_t wx = x0 + (int32_t)(u % nWin) * a.ww;
    RowCtx C;
    C.F = a.F;
    C.parts = a.parts;
    C.E = a.E;
    C.EW = a.EW;
    C.y = y;
    C.wx0 = wx;
    C.wx1 = wx + a.ww < hi ? wx + a.ww : hi;
    C.cx0 = lo;
    C.cx1 = hi;
    int32_t kind = row_program(C, L);
    if (kind < 0) {
      if (threadIdx.x == 0) atomicOr(a.err, 1);
      break;
    }
    if (kind == kRowNone) continue;
    if (kA8)
      ntouch += row_epilogue_a8(C, L, kind, (uint8_t*)a.canvas, W, F.rgbx, F.mode);
    else
      ntouch += row_epilogue(C, L, kind, (uint32_t*)a.canvas, W, F.rgbx, F.mode);
  }
  if (a.touched) {
    // wave reduction, one atomic per workgroup
    for (int o = 32; o > 0; o >>= 1) ntouch += __shfl_xor(ntouch, o, kWave);
    if (threadIdx.x == 0 && ntouch) atomicAdd(a.touched, (unsigned long long)ntouch);
  }
}

int launch_fill_rows(const FillArgs& a, int grid, size_t ldsBytes, bool a8, hipStream_t s) {
  if (a8)
    hipLaunchKernelGGL(fill_rows_kernel<true>, dim3(grid), dim3(64), ldsBytes, s, a);
  else
    hipLaunchKernelGGL(fill_rows_kernel<false>, dim3(grid), dim3(64), ldsBytes, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int set_fill_lds_limit(size_t bytes) {
  if (hipFuncSetAttribute((const void*)fill_rows_kernel<false>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return kErrDevice;
  if (hipFuncSetAttribute((const void*)fill_rows_kernel<true>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return kErrDevice;
  return 0;
}

// ---------------------------------------------------------------------------
// composite (blendRect): each thread handles 4 consecutive destination pixels
// ---------------------------------------------------------------------------
template <bool kSrcA8>
__global__ void __launch_bounds__(256) composite_kernel(CompositeArgs a) {
  const int32_t y = blockIdx.y;  // destination row
  const int32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (x0 >= a.aw) return;
  uint32_t* arow = a.dst + (size_t)y * a.aw;
  const bool mask = a.mode == kMask;
  const int32_t sy = y - a.py;
  const bool rowIn = sy >= a.yStart && sy < a.yEnd;
  const bool vec = ((a.aw & 3) == 0);
  uint32_t px[4];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int32_t x = x0 + k;
    px[k] = 0;
    if (x >= a.aw) continue;
    int32_t sx = x - a.px;
    bool in = rowIn && sx >= a.xStart && sx < a.xEnd;
    if (in || mask) any = true;
  }
  if (!any) return;
  if (vec) {
    uint4 v = *(const uint4*)(arow + x0);
    px[0] = v.x; px[1] = v.y; px[2] = v.z; px[3] = v.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) px[k] = (x0 + k < a.aw) ? arow[x0 + k] : 0u;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int32_t x = x0 + k;
    if (x >= a.aw) continue;
    int32_t sx = x - a.px;
    bool in = rowIn && sx >= a.xStart && sx < a.xEnd;
    if (in) {
      uint32_t s;
      if (kSrcA8) {
        uint32_t v = ((const uint8_t*)a.src)[(size_t)sy * a.bw + sx];
        s = v * 0x01010101u;
      } else {
        s = ((const uint32_t*)a.src)[(size_t)sy * a.bw + sx];
      }
      px[k] = blend(a.mode, px[k], s);
    } else if (mask) {
      px[k] = 0;
    }
  }
  if (vec) {
    *(uint4*)(arow + x0) = make_uint4(px[0], px[1], px[2], px[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (x0 + k < a.aw) arow[x0 + k] = px[k];
  }
}

int launch_composite(const CompositeArgs& a, bool srcA8, hipStream_t s) {
  dim3 grid((a.aw + 1023) / 1024, a.ah);
  if (srcA8)
    hipLaunchKernelGGL(composite_kernel<true>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(composite_kernel<false>, grid, dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// ---------------------------------------------------------------------------
// blur: separable gaussian on premultiplied RGBA8 (or an alpha plane).
// Each pass computes sum(c*k) in uint32 and stores (sum div 256 div 255).
// The LUT is symmetric, so taps are paired: k[r+j]*(c[-j] + c[+j]).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void acc_rgba(uint32_t* acc, uint32_t p, uint32_t k) {
  acc[0] += (p & 255u) * k;
  acc[1] += ((p >> 8) & 255u) * k;
  acc[2] += ((p >> 16) & 255u) * k;
  acc[3] += (p >> 24) * k;
}
__device__ __forceinline__ void acc_pair(uint32_t* acc, uint32_t p, uint32_t q, uint32_t k) {
  uint32_t rb = (p & 0x00FF00FFu) + (q & 0x00FF00FFu);
  uint32_t ga = ((p >> 8) & 0x00FF00FFu) + ((q >> 8) & 0x00FF00FFu);
  acc[0] += (rb & 0xFFFFu) * k;
  acc[2] += (rb >> 16) * k;
  acc[1] += (ga & 0xFFFFu) * k;
  acc[3] += (ga >> 16) * k;
}
__device__ __forceinline__ uint32_t acc_out(const uint32_t* acc) {
  return pack(acc[0] / 256u / 255u, acc[1] / 256u / 255u, acc[2] / 256u / 255u, acc[3] / 256u / 255u);
}

// X pass over rows: a block produces 256 outputs of one row.  LDS holds the
// row segment [x0 - r, x0 + 256 + r) with out-of-bounds samples replaced by
// the oob colour.
__global__ void __launch_bounds__(256) blur_x_kernel(BlurArgs a) {
  extern __shared__ uint32_t srow[];
  const int32_t y = blockIdx.y;
  const int32_t x0 = blockIdx.x * 256;
  const int32_t r = a.radius;
  const int32_t n = 256 + 2 * r;
  const uint32_t* in = a.src + (size_t)y * a.w;
  for (int i = threadIdx.x; i < n; i += 256) {
    int32_t x = x0 - r + i;
    srow[i] = (x >= 0 && x < a.w) ? in[x] : a.oob;
  }
  __syncthreads();
  const int32_t x = x0 + threadIdx.x;
  if (x >= a.w) return;
  const int32_t c = threadIdx.x + r;
  uint32_t acc[4] = {0, 0, 0, 0};
  acc_rgba(acc, srow[c], a.lut[r]);
  for (int32_t j = 1; j <= r; j++) acc_pair(acc, srow[c - j], srow[c + j], a.lut[r + j]);
  a.dst[(size_t)y * a.w + x] = acc_out(acc);
}

// Y pass: a block produces a 64-column x kTileRows tile; LDS holds the
// columns' [y0 - r, y0 + kTileRows + r) rows.
constexpr int kBlurTileRows = 64;
__global__ void __launch_bounds__(256) blur_y_kernel(BlurArgs a) {
  extern __shared__ uint32_t stile[];
  const int32_t x0 = blockIdx.x * 64, y0 = blockIdx.y * kBlurTileRows;
  const int32_t r = a.radius;
  const int32_t rows = kBlurTileRows + 2 * r;
  const int32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  const int32_t x = x0 + tx;
  for (int32_t i = ty; i < rows; i += 4) {
    int32_t yy = y0 - r + i;
    uint32_t v = a.oob;
    if (yy >= 0 && yy < a.h && x < a.w) v = a.src[(size_t)yy * a.w + x];
    stile[i * 64 + tx] = v;
  }
  __syncthreads();
  if (x >= a.w) return;
  for (int32_t o = ty; o < kBlurTileRows; o += 4) {
    int32_t y = y0 + o;
    if (y >= a.h) break;
    const int32_t c = o + r;
    uint32_t acc[4] = {0, 0, 0, 0};
    acc_rgba(acc, stile[c * 64 + tx], a.lut[r]);
    for (int32_t j = 1; j <= r; j++)
      acc_pair(acc, stile[(c - j) * 64 + tx], stile[(c + j) * 64 + tx], a.lut[r + j]);
    uint32_t out = acc_out(acc);
    if (a.shadowColor) out = blend(kMask, a.color, out);  // shadow: fill(color).draw(mask, MaskBlend)
    a.dst[(size_t)y * a.w + x] = out;
  }
}

// Writes a Gaussian table passed by value into device memory (the
// allocation-free blur keeps its table in static device scratch).
__global__ void __launch_bounds__(256) lut_store_kernel(LutTaps t, uint32_t* dst) {
  for (int i = threadIdx.x; i < t.n; i += 256) dst[i] = t.v[i];
}

int launch_lut_store(const LutTaps& t, uint32_t* dst, hipStream_t s) {
  hipLaunchKernelGGL(lut_store_kernel, dim3(1), dim3(256), 0, s, t, dst);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

size_t blur_x_lds(int r) { return (size_t)(256 + 2 * r) * 4; }
size_t blur_y_lds(int r) { return (size_t)64 * (kBlurTileRows + 2 * r) * 4; }

int launch_blur(const BlurArgs& ax, const BlurArgs& ay, hipStream_t s) {
  dim3 gx((ax.w + 255) / 256, ax.h);
  hipLaunchKernelGGL(blur_x_kernel, gx, dim3(256), blur_x_lds(ax.radius), s, ax);
  dim3 gy((ay.w + 63) / 64, (ay.h + kBlurTileRows - 1) / kBlurTileRows);
  hipLaunchKernelGGL(blur_y_kernel, gy, dim3(256), blur_y_lds(ay.radius), s, ay);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

int set_blur_lds_limit(size_t bytes) {
  if (hipFuncSetAttribute((const void*)blur_x_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)bytes) != hipSuccess)
    return kErrDevice;
  if (hipFuncSetAttribute((const void*)blur_y_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)bytes) != hipSuccess)
    return kErrDevice;
  return 0;
}

// ---------------------------------------------------------------------------
// spread (images.nim:700-758): X pass writes an alpha plane, Y pass writes
// rgbx(0,0,0,a).  The window is clamped to the image (no oob padding).
// When shift != (0,0) the X pass reads the source translated (shadow's
// OverwriteBlend draw of the offset image, images.nim:765-769).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) spread_x_kernel(SpreadArgs a) {
  const int32_t y = blockIdx.y;
  const int32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= a.w) return;
  const int32_t s = a.s;
  const bool grow = a.grow;
  int32_t lo = x - s > 0 ? x - s : 0, hi = x + s < a.w - 1 ? x + s : a.w - 1;
  uint32_t v = grow ? 0u : 255u;
  const int32_t sy = y - a.dy;
  for (int32_t xx = lo; xx <= hi; xx++) {
    int32_t sx = xx - a.dx;
    uint32_t al = 0;
    if (sy >= 0 && sy < a.h && sx >= 0 && sx < a.w) al = a.src[(size_t)sy * a.w + sx] >> 24;
    if (grow ? al > v : al < v) v = al;
  }
  a.plane[(size_t)y * a.w + x] = (uint8_t)v;
}

__global__ void __launch_bounds__(256) spread_y_kernel(SpreadArgs a) {
  const int32_t y = blockIdx.y;
  const int32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= a.w) return;
  const int32_t s = a.s;
  const bool grow = a.grow;
  int32_t lo = y - s > 0 ? y - s : 0, hi = y + s < a.h - 1 ? y + s : a.h - 1;
  uint32_t v = grow ? 0u : 255u;
  for (int32_t yy = lo; yy <= hi; yy++) {
    uint32_t al = a.plane[(size_t)yy * a.w + x];
    if (grow ? al > v : al < v) v = al;
  }
  a.dst[(size_t)y * a.w + x] = v << 24;
}

int launch_spread(const SpreadArgs& a, hipStream_t s) {
  dim3 g((a.w + 255) / 256, a.h);
  hipLaunchKernelGGL(spread_x_kernel, g, dim3(256), 0, s, a);
  hipLaunchKernelGGL(spread_y_kernel, g, dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// Offset copy for shadow without spread: dst = draw(src, translate(dx,dy),
// OverwriteBlend) onto a transparent image.
__global__ void __launch_bounds__(256) shift_kernel(SpreadArgs a) {
  const int32_t y = blockIdx.y;
  const int32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= a.w) return;
  int32_t sx = x - a.dx, sy = y - a.dy;
  uint32_t v = 0;
  if (sx >= 0 && sx < a.w && sy >= 0 && sy < a.h) v = a.src[(size_t)sy * a.w + sx];
  a.dst[(size_t)y * a.w + x] = v;
}

int launch_shift(const SpreadArgs& a, hipStream_t s) {
  dim3 g((a.w + 255) / 256, a.h);
  hipLaunchKernelGGL(shift_kernel, g, dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// ---------------------------------------------------------------------------
// applyOpacity (images.nim:261-277), o = round(255*opacity) in (0, 255)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) opacity_kernel(uint32_t* img, int64_t n, uint32_t o) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t p = img[i];
  img[i] = pack((ch_r(p) * o) / 255u, (ch_g(p) * o) / 255u, (ch_b(p) * o) / 255u, (ch_a(p) * o) / 255u);
}

int launch_opacity(uint32_t* img, int64_t n, uint32_t o, hipStream_t s) {
  hipLaunchKernelGGL(opacity_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, img, n, o);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// ---------------------------------------------------------------------------
// blend test kernel: out[i] = blend(mode, b[i], s[i])
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) blend_pixels_kernel(const uint32_t* b, const uint32_t* src,
                                                           uint32_t* out, int64_t n, int32_t mode) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = blend(mode, b[i], src[i]);
}

int launch_blend_pixels(const uint32_t* b, const uint32_t* s_, uint32_t* out, int64_t n, int32_t mode,
                        hipStream_t s) {
  hipLaunchKernelGGL(blend_pixels_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, s_,
                     out, n, mode);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

}  // namespace pxh
