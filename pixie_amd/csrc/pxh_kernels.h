// pxh_kernels.h -- argument blocks and launchers of the HIP kernels
// (pxh_kernels.hip).  Host code includes this; device pointers only.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pxh_common.h"

namespace pxh {

struct FillArgs {
  const FillInfo* F;     // one fill
  const PartInfo* parts; // its bands
  const Entry* E;        // its entries
  const int8_t* EW;      // entry windings
  void* canvas;          // RGBA8 (uint32 per pixel) or A8 target
  int32_t rowBase, nRows;
  int32_t cap, ww;       // row scratch capacity, window width
  int32_t capInLds;      // cap arrays in LDS (else gscratch)
  char* gscratch;        // per-workgroup cap arrays when not in LDS
  int32_t* err;          // set to 1 when a band exceeds cap
  unsigned long long* touched;  // pixels written (N_touch), may be null
};

struct CompositeArgs {
  uint32_t* dst;
  const void* src;       // RGBA8 or A8 source
  int32_t aw, ah, bw, bh;
  int32_t px, py;        // integer translation of src
  int32_t xStart, xEnd, yStart, yEnd;  // src rect that lands inside dst
  int32_t mode;
};

struct BlurArgs {
  const uint32_t* src;
  uint32_t* dst;
  const uint32_t* wtab;  // weight table of blur_table_kernel (device)
  int32_t w, h, radius;
  uint32_t oob;
  int32_t shadowColor;   // Y pass only: write blendMask(color, result)
  uint32_t color;
  int32_t nd;            // blur_table_dwords(radius)
};

// Gaussian taps passed by value (kernel arguments stay under 4 KB).
constexpr int kMaxLutTaps = 1024;
struct LutTaps {
  uint16_t v[kMaxLutTaps];
  int32_t n;
};

struct SpreadArgs {
  const uint32_t* src;
  uint8_t* plane;        // alpha plane scratch (w*h)
  uint32_t* dst;
  int32_t w, h, s, grow;
  int32_t dx, dy;        // read src translated by (dx, dy)
};

// Non-solid paint over the whole canvas (fillPath's route, paths.nim:
// 2115-2142): dst = blend(mode, dst, blendMask(fill, applyOpacity(mask))),
// the mask being the coverage of a white fill and fill the gradient (made
// per pixel) or the paint's image drawn onto a transparent canvas.
constexpr int kMaxStops = 32;
struct PaintArgs {
  uint32_t* dst;
  const uint8_t* mask;      // A8 coverage of the path
  int32_t w, h;
  const uint32_t* fill;     // image paints (kinds 1, 2): the image drawn on transparent
  int32_t kind, mode;       // PaintKind ordinal (1..5), blend mode
  int32_t useOpacity;       // opacity != 1: mask *= o255 / 255
  uint32_t o255;            // round(255 * opacity)
  float hx[6];              // gradient handles
  float m[9];               // radial: the inverse gradient matrix
  float ga;                 // angular: fixAngle(angle(normalize(edge - center)))
  int32_t nstops;
  float stops[kMaxStops * 5];  // r, g, b, a (straight), position
};
int launch_paint(const PaintArgs& a, hipStream_t s);

int launch_fill_rows(const FillArgs& a, int grid, size_t ldsBytes, bool a8, hipStream_t s);
int set_fill_lds_limit(size_t bytes);
int launch_composite(const CompositeArgs& a, bool srcA8, hipStream_t s);
int launch_blur(const BlurArgs& ax, const BlurArgs& ay, hipStream_t s);
int set_blur_lds_limit(size_t bytes);
int launch_lut_store(const LutTaps& t, uint32_t* dst, hipStream_t s);
// Weight table of a blur of radius r (2 * 4 * blur_table_dwords(r) dwords)
// built on the device from its 2r + 1 taps.
int blur_table_dwords(int r);
int launch_blur_table(const uint32_t* lut, int r, uint32_t* wtab, hipStream_t s);
size_t blur_x_lds(int r);
size_t blur_y_lds(int r);
int launch_spread(const SpreadArgs& a, hipStream_t s);
int launch_shift(const SpreadArgs& a, hipStream_t s);
int launch_opacity(uint32_t* img, int64_t n, uint32_t o, hipStream_t s);
int launch_blend_pixels(const uint32_t* b, const uint32_t* s_, uint32_t* out, int64_t n,
                        int32_t mode, hipStream_t s);

}  // namespace pxh
