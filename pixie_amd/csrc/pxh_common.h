// pxh_common.h -- types and scalar helpers shared by host code and HIP kernels.
//
// Numeric conventions follow pixie v6.0.0 compiled by Nim's C backend on
// x86-64: float32 IEEE arithmetic in source order (everything here is built
// with -ffp-contract=off and correctly rounded divide/sqrt), float -> int
// conversions are C casts as gcc emits them (cvttss2si: NaN and out-of-range
// values become the "integer indefinite" value), unsigned products wrap.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define PXH_HD __host__ __device__ __forceinline__
#else
#define PXH_HD inline
#endif

namespace pxh {

// BlendMode ordinals, common.nim:6-29.
enum BlendMode : int32_t {
  kNormal = 0, kDarken, kMultiply, kColorBurn, kLighten, kScreen, kColorDodge,
  kOverlay, kSoftLight, kHardLight, kDifference, kExclusion, kHue, kSaturation,
  kColor, kLuminosity, kMask, kOverwrite, kSubtractMask, kExcludeMask,
  kNumBlendModes
};

// WindingRule, paths.nim:5-8.
enum WindingRule : int32_t { kNonZero = 0, kEvenOdd = 1 };

// Error codes returned through the C-ABI (0 = ok).  The messages mirror the
// PixieError texts raised by the reference.
enum Status : int32_t {
  kOk = 0,
  kErrDiscretize = 1,   // "Unable to discretize cubic/quadratic/arc"
  kErrIntOverflow = 2,  // "Path int overflow detected"
  kErrDash = 3,         // "Invalid line dash value"
  kErrNegBlur = 4,      // "Cannot apply negative blur"
  kErrParse = 5,        // "Invalid path ..."
  kErrArg = 6,          // bad argument (size, mode, handle)
  kErrNoMem = 7,        // device or host allocation failed
  kErrDevice = 8,       // HIP runtime error
  kErrCapacity = 9,     // a fixed-capacity buffer would overflow
};

// float -> int64 as gcc's cvttss2si (64-bit destination).
PXH_HD int64_t nim_f2i(float x) {
  if (!(x >= -9.2233720368547758e18f && x < 9.2233720368547758e18f)) return INT64_MIN;
  return (int64_t)x;
}
// float -> uint32 via the 64-bit conversion.
PXH_HD uint32_t nim_f2u32(float x) { return (uint32_t)(uint64_t)nim_f2i(x); }
// float -> int32 (Fixed32 conversion), 32-bit destination.
PXH_HD int32_t nim_f2i32(float x) {
  if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
  return (int32_t)x;
}
// Nim's NaN-aware float min / max (system.nim).
PXH_HD float nim_minf(float x, float y) { return (x <= y || y != y) ? x : y; }
PXH_HD float nim_maxf(float x, float y) { return (y <= x || y != y) ? x : y; }

// vmath Mat3 in float32, m[3 * col + row] (a[i, j] = m[3 * i + j]): a point
// transformed as a[0, j] * x + a[1, j] * y + a[2, j], and the product a * b
// with each entry summed over k in order.
PXH_HD void m3_apply(const float* m, float x, float y, float* ox, float* oy) {
  *ox = m[0] * x + m[3] * y + m[6];
  *oy = m[1] * x + m[4] * y + m[7];
}
PXH_HD void m3_mul(const float* a, const float* b, float* o) {
  float r[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r[3 * i + j] = b[3 * i + 0] * a[0 * 3 + j] + b[3 * i + 1] * a[1 * 3 + j] + b[3 * i + 2] * a[2 * 3 + j];
  for (int i = 0; i < 9; i++) o[i] = r[i];
}

// 24.8 fixed point, paths.nim:1268-1275.
PXH_HD int32_t fixed32(float f) { return nim_f2i32(f * 256.0f); }
PXH_HD int32_t fx_int(int32_t p) { return p / 256; }
PXH_HD int32_t fx_trunc(int32_t p) { return (p / 256) * 256; }

PXH_HD bool should_fill(int32_t rule, int64_t count) {  // paths.nim:1288-1296
  return rule == kNonZero ? count != 0 : (count % 2) != 0;
}

// Pixel packing: r in the low byte, as pixie's ColorRGBX {r,g,b,a} in memory.
PXH_HD uint32_t pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return (r & 255u) | ((g & 255u) << 8) | ((b & 255u) << 16) | ((a & 255u) << 24);
}
PXH_HD uint32_t ch_r(uint32_t c) { return c & 255u; }
PXH_HD uint32_t ch_g(uint32_t c) { return (c >> 8) & 255u; }
PXH_HD uint32_t ch_b(uint32_t c) { return (c >> 16) & 255u; }
PXH_HD uint32_t ch_a(uint32_t c) { return c >> 24; }

// ---------------------------------------------------------------------------
// Device data layout of a fill (one fillShapes call)
// ---------------------------------------------------------------------------

// Per-fill header: the locals fillShapes derives before its scanline loop
// (paths.nim:1602-1629), plus the paint.
struct FillInfo {
  int32_t startX, startY, pathWidth, pathHeight;  // paths.nim:1606-1613
  int32_t nParts, bandH;      // partitionSegments: band i = [startY+i*bandH, ..)
  int32_t partBase;           // first PartInfo of this fill
  uint32_t rgbx;              // premultiplied paint colour
  int32_t mode, rule;         // BlendMode, WindingRule
  int32_t xlo, xhi;           // column range the fill may write (see below)
  int32_t status;             // kOk, or the error the geometry raised
  int32_t canvasW, canvasH;   // target size
  int32_t pad0;
};

// Writes of a non-mask fill are confined to [startX - kWriteMargin,
// startX + pathWidth + kWriteMargin) so that fills can be binned to tiles by
// their bounds.  Coverage writes are already confined to the path span; only
// trapezoid slivers, computed from slope-intercept x, can stray and only by
// float error.  The oracle applies the same clip.
constexpr int32_t kWriteMargin = 16;

// Partition (band) header, paths.nim:36-39.
struct PartInfo {
  int32_t entryBase;  // first entry
  int32_t count;      // entries in the band
  int32_t flags;      // bit0: requiresAntiAliasing, bit1: twoNonintersectingSpanningSegments
  int32_t pad;
};

// Partition entry, paths.nim:31-34: (at.y, to.y) after clipping to the band,
// slope-intercept (m, b) of the unclipped segment.  Winding is kept in a
// separate int8 array.
struct Entry {
  float aty, toy, m, b;
};

}  // namespace pxh
