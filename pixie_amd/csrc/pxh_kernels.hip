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
  // the wide path stages its keys through AX and, when they are in LDS, the
  // arrays that follow it
  L.wide = (uint64_t*)L.AX;
  L.wideCap = (int)(((size_t)a.ww * 4 + (a.capInLds ? row_cap_bytes(a.cap) : 0)) / 8);
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
// gradient paints (paints.nim:72-248) through fillPath's mask route
// (paths.nim:2115-2142), one pass: each pixel takes the path's coverage from
// the A8 mask, the gradient colour at its integer coordinates, masks the
// colour with the coverage as fill.draw(mask, MaskBlend) does, and blends it
// as image.draw(fill, blendMode) does.  Float math is pixie's float32 order.
// ---------------------------------------------------------------------------
// gradientColor (paints.nim:72-97), then Color -> ColorRGBX (round to 8 bits,
// premultiply with (c * a + 127) div 255)
__device__ uint32_t gradient_color(const PaintArgs& a, float t) {
  int index = -1;
  for (int i = 0; i < a.nstops; i++) {
    const float pos = a.stops[5 * i + 4];
    if (pos < t) index = i;
    if (pos > t) break;
  }
  float c[4];
  if (index == -1 || index + 1 >= a.nstops) {
    const int k = index == -1 ? 0 : index;
    for (int j = 0; j < 4; j++) c[j] = a.stops[5 * k + j];
  } else {
    const float* s1 = a.stops + 5 * index;
    const float* s2 = a.stops + 5 * (index + 1);
    const float v = (t - s1[4]) / (s2[4] - s1[4]);
    for (int j = 0; j < 4; j++) c[j] = s1[j] * (1.0f - v) + s2[j] * v;
  }
  const FColor f = {c[0], c[1], c[2], c[3]};
  return from_fcolor(f);
}

__device__ __forceinline__ float fix_angle_p(float x) {
  const float PI = 3.14159265358979323846f;
  while (x > PI) x -= PI * 2.0f;
  while (x < -PI) x += PI * 2.0f;
  return x;
}

// the gradient parameter t of pixel (x, y) (paints.nim:108-113, 199-209,
// 224-236)
__device__ float gradient_t(const PaintArgs& a, int32_t x, int32_t y) {
  const float fx = (float)x, fy = (float)y;
  if (a.kind == 3) {  // linear: toLineSpace
    const float dx = a.hx[2] - a.hx[0], dy = a.hx[3] - a.hx[1];
    const float det = dx * dx + dy * dy;
    return (dy * (fy - a.hx[1]) + dx * (fx - a.hx[0])) / det;
  }
  if (a.kind == 4) {  // radial: |inverse gradient matrix * xy|
    const float px = a.m[0] * fx + a.m[3] * fy + a.m[6];
    const float py = a.m[1] * fx + a.m[4] * fy + a.m[7];
    return sqrtf(px * px + py * py);
  }
  const float PI = 3.14159265358979323846f;  // angular
  const float vx = fx - a.hx[0], vy = fy - a.hx[1];
  const float l = sqrtf(vx * vx + vy * vy);
  const float angle = atan2f(vy / l, vx / l);
  return fix_angle_p(angle + a.ga + PI / 2.0f) / 2.0f / PI + 0.5f;
}

__global__ void __launch_bounds__(256) paint_kernel(PaintArgs a) {
  const int32_t y = blockIdx.y;
  const int32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= a.w) return;
  const size_t i = (size_t)y * a.w + x;
  uint32_t m = a.mask[i];
  if (a.useOpacity) m = m * a.o255 / 255u;  // applyOpacity on the mask (images.nim:261-277)
  // image paints read the image drawn onto a transparent canvas
  const uint32_t g = a.kind <= 2 ? a.fill[i] : gradient_color(a, gradient_t(a, x, y));
  // fill.draw(mask, MaskBlend): every channel times the mask's alpha
  const uint32_t src = pack(ch_r(g) * m / 255u, ch_g(g) * m / 255u, ch_b(g) * m / 255u, ch_a(g) * m / 255u);
  // a transparent source leaves the backdrop of these modes bit for bit
  const int md = a.mode;
  const bool keepOnZero = md == kNormal || md == kDarken || md == kMultiply || md == kLighten ||
                          md == kScreen || md == kOverlay || md == kHardLight || md == kDifference ||
                          md == kExclusion;
  if (src == 0u && keepOnZero) return;
  a.dst[i] = blend(md, a.dst[i], src);
}

int launch_paint(const PaintArgs& a, hipStream_t s) {
  dim3 g((a.w + 255) / 256, a.h);
  hipLaunchKernelGGL(paint_kernel, g, dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
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
// blur: separable gaussian on premultiplied RGBA8 (images.nim:304-365).
//
// Each pass computes, per channel, sum(c * k) over 2r + 1 taps in uint32 and
// stores (sum div 256 div 255).  The sums are formed with v_dot4_u32_u8
// (four byte products per instruction): the pixels are byte-transposed into
// channel planes (byte i of a plane dword = one channel of pixel i), and each
// 16-bit tap k = 256 * kh + kl is split into its bytes, so
//   sum(c * k) = 256 * sum(c * kh) + sum(c * kl)
// exactly (every partial sum stays below 2^32).  A thread computes four
// consecutive outputs (m = 0..3); output m of a thread whose first input
// dword is d0 weighs byte i of dword d0 + d by the tap 4d + i - m - a, where
// a = R - r and R is r rounded up to a multiple of 4, so one plane dword
// feeds four outputs through four shifted weight tables (zero outside the
// window).  The table is built once per radius by blur_table_kernel.
// ---------------------------------------------------------------------------

// Weight dwords of one blur: wtab[(d * 4 + m) * 2 + h], h = 0 the low bytes
// of the taps, h = 1 the high bytes (the eight dwords of input dword d are
// contiguous, so a wave reads them as two broadcast 16-byte LDS reads).
int blur_table_dwords(int r) { return (2 * r + 10) / 4; }

__global__ void __launch_bounds__(256) blur_table_kernel(const uint32_t* lut, int r, int nd,
                                                         uint32_t* wtab) {
  const int a = ((r + 3) & ~3) - r;
  for (int e = threadIdx.x; e < 4 * nd; e += 256) {
    const int d = e / 4, m = e % 4;
    uint32_t lo = 0, hi = 0;
    for (int i = 0; i < 4; i++) {
      const int j = 4 * d + i - m - a;
      const uint32_t k = (j >= 0 && j <= 2 * r) ? lut[j] : 0u;
      lo |= (k & 255u) << (8 * i);
      hi |= (k >> 8) << (8 * i);
    }
    wtab[2 * e] = lo;
    wtab[2 * e + 1] = hi;
  }
}

int launch_blur_table(const uint32_t* lut, int r, uint32_t* wtab, hipStream_t s) {
  hipLaunchKernelGGL(blur_table_kernel, dim3(1), dim3(256), 0, s, lut, r, blur_table_dwords(r), wtab);
  return hipGetLastError() == hipSuccess ? 0 : kErrDevice;
}

// Four RGBA8 pixels -> four channel planes (plane c, byte i = channel c of
// pixel i), eight byte permutes.
__device__ __forceinline__ uint4 planes4(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3) {
  const uint32_t t0 = __builtin_amdgcn_perm(p1, p0, 0x05010400u);  // r0 r1 g0 g1
  const uint32_t t1 = __builtin_amdgcn_perm(p1, p0, 0x07030602u);  // b0 b1 a0 a1
  const uint32_t t2 = __builtin_amdgcn_perm(p3, p2, 0x05010400u);  // r2 r3 g2 g3
  const uint32_t t3 = __builtin_amdgcn_perm(p3, p2, 0x07030602u);  // b2 b3 a2 a3
  return make_uint4(__builtin_amdgcn_perm(t2, t0, 0x05040100u), __builtin_amdgcn_perm(t2, t0, 0x07060302u),
                    __builtin_amdgcn_perm(t3, t1, 0x05040100u), __builtin_amdgcn_perm(t3, t1, 0x07060302u));
}

// The four outputs' channel sums of one thread.
struct BlurAcc {
  uint32_t lo[4][4], hi[4][4];  // [m][channel]
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int c = 0; c < 4; c++) lo[m][c] = hi[m][c] = 0;
  }
  // v: the four channel planes of one input dword; w: its eight weight
  // dwords (LDS, the same for every lane)
  __device__ __forceinline__ void add(const uint4 v, const uint4* w) {
    const uint4 w01 = w[0], w23 = w[1];
    const uint32_t ws[8] = {w01.x, w01.y, w01.z, w01.w, w23.x, w23.y, w23.z, w23.w};
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const uint32_t wl = ws[2 * m], wh = ws[2 * m + 1];
      lo[m][0] = __builtin_amdgcn_udot4(v.x, wl, lo[m][0], false);
      hi[m][0] = __builtin_amdgcn_udot4(v.x, wh, hi[m][0], false);
      lo[m][1] = __builtin_amdgcn_udot4(v.y, wl, lo[m][1], false);
      hi[m][1] = __builtin_amdgcn_udot4(v.y, wh, hi[m][1], false);
      lo[m][2] = __builtin_amdgcn_udot4(v.z, wl, lo[m][2], false);
      hi[m][2] = __builtin_amdgcn_udot4(v.z, wh, hi[m][2], false);
      lo[m][3] = __builtin_amdgcn_udot4(v.w, wl, lo[m][3], false);
      hi[m][3] = __builtin_amdgcn_udot4(v.w, wh, hi[m][3], false);
    }
  }
  __device__ __forceinline__ uint32_t out(int m) const {
    uint32_t c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = (lo[m][k] + (hi[m][k] << 8)) / 256u / 255u;
    return pack(c[0], c[1], c[2], c[3]);
  }
};

// X pass: a block produces 1024 outputs of one row, four per thread.  LDS
// holds the row segment from x0 - R in channel planes, out-of-bounds pixels
// replaced by the oob colour.
__global__ void __launch_bounds__(256) blur_x_kernel(BlurArgs a) {
  extern __shared__ uint4 splane[];
  const int32_t y = blockIdx.y;
  const int32_t x0 = blockIdx.x * 1024;
  const int nd = a.nd;
  uint4* sw = splane + 256 + nd;  // the weight table, 2 uint4 per input dword
  for (int i = threadIdx.x; i < 2 * nd; i += 256) sw[i] = ((const uint4*)a.wtab)[i];
  const int32_t s0 = x0 - ((a.radius + 3) & ~3);
  const uint32_t* in = a.src + (size_t)y * a.w;
  const bool vec = (a.w & 3) == 0;
  for (int q = threadIdx.x; q < 256 + nd; q += 256) {
    const int32_t xs = s0 + 4 * q;
    uint32_t p[4];
    if (vec && xs >= 0 && xs + 3 < a.w) {
      const uint4 v = *(const uint4*)(in + xs);
      p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w;
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) p[i] = (xs + i >= 0 && xs + i < a.w) ? in[xs + i] : a.oob;
    }
    splane[q] = planes4(p[0], p[1], p[2], p[3]);
  }
  __syncthreads();
  const int t = threadIdx.x;
  const int32_t xo = x0 + 4 * t;
  if (xo >= a.w) return;
  BlurAcc acc;
  acc.zero();
  for (int d = 0; d < nd; d++) acc.add(splane[t + d], sw + 2 * d);
  uint32_t* outp = a.dst + (size_t)y * a.w;
  if (vec && xo + 3 < a.w) {
    *(uint4*)(outp + xo) = make_uint4(acc.out(0), acc.out(1), acc.out(2), acc.out(3));
  } else {
#pragma unroll
    for (int m = 0; m < 4; m++)
      if (xo + m < a.w) outp[xo + m] = acc.out(m);
  }
}

// Y pass: a block produces a 64-column x 64-row tile; thread (column, g)
// computes the row quads g, g + 4, g + 8, g + 12.  LDS holds the columns'
// rows from y0 - R, four rows to a plane dword.
constexpr int kBlurTileRows = 64;
__global__ void __launch_bounds__(256) blur_y_kernel(BlurArgs a) {
  extern __shared__ uint4 stile[];
  const int32_t x0 = blockIdx.x * 64, y0 = blockIdx.y * kBlurTileRows;
  const int nd = a.nd;
  const int32_t s0 = y0 - ((a.radius + 3) & ~3);
  const int nq = kBlurTileRows / 4 - 1 + nd;  // row quads the tile reads
  uint4* sw = stile + nq * 64;                 // the weight table
  for (int i = threadIdx.x; i < 2 * nd; i += 256) sw[i] = ((const uint4*)a.wtab)[i];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int32_t x = x0 + tx;
  for (int q = ty; q < nq; q += 4) {
    const int32_t ys = s0 + 4 * q;
    uint32_t p[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int32_t yy = ys + i;
      p[i] = (yy >= 0 && yy < a.h && x < a.w) ? a.src[(size_t)yy * a.w + x] : a.oob;
    }
    stile[q * 64 + tx] = planes4(p[0], p[1], p[2], p[3]);
  }
  __syncthreads();
  if (x >= a.w) return;
  for (int Q = ty; Q < kBlurTileRows / 4; Q += 4) {
    const int32_t yq = y0 + 4 * Q;
    if (yq >= a.h) break;
    BlurAcc acc;
    acc.zero();
    for (int d = 0; d < nd; d++) acc.add(stile[(Q + d) * 64 + tx], sw + 2 * d);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      if (yq + m >= a.h) break;
      uint32_t out = acc.out(m);
      if (a.shadowColor) out = blend(kMask, a.color, out);  // shadow: fill(color).draw(mask, MaskBlend)
      a.dst[(size_t)(yq + m) * a.w + x] = out;
    }
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

size_t blur_x_lds(int r) { return (size_t)(256 + 3 * blur_table_dwords(r)) * 16; }
size_t blur_y_lds(int r) {
  return (size_t)(kBlurTileRows / 4 - 1 + blur_table_dwords(r)) * 64 * 16 + (size_t)blur_table_dwords(r) * 32;
}

int launch_blur(const BlurArgs& ax, const BlurArgs& ay, hipStream_t s) {
  dim3 gx((ax.w + 1023) / 1024, ax.h);
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
