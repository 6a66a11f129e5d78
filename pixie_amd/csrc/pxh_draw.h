// pxh_draw.h -- argument blocks and launchers of the transformed-draw
// kernels (pxh_draw.hip).  Host code includes this; device pointers only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pxh_common.h"

namespace pxh {

// A draw source: an RGBA8 image read as if magnified by 2^mag.
struct DrawSrc {
  const uint32_t* px;
  int32_t w, h;   // stored size
  int32_t mag;    // magnifyBy2 power applied on the fly (0 = none)
};

struct MinifyArgs {
  const uint32_t* src;
  int32_t w, h;   // source size; dst is ((w + 1) / 2, (h + 1) / 2)
  uint32_t* dst;
};

struct MagnifyArgs {
  const uint32_t* src;
  int32_t w, h, power;  // dst is (w << power, h << power)
  uint32_t* dst;
};

// drawSmooth (images.nim:531-634) of src onto dst.  Rows are cut into runs
// of kSmoothRun pixels; the running sample position at each run's start is
// kept in `starts` (ah x nRuns).
constexpr int kSmoothRun = 16;
struct SmoothArgs {
  uint32_t* dst;
  int32_t aw, ah;
  DrawSrc src;
  float per[4][4];            // perimeter segments (at.x, at.y, to.x, to.y)
  float p[2], dx[2], dy[2];   // movement vectors of the inverse transform
  int32_t yStart, yEnd;       // clamped row range of the perimeter
  int32_t mode;
  int2* span;                 // per row: [xStart, xEnd) (scratch, ah entries)
  float2* starts;             // per row and run: the start position (scratch)
  int32_t nRuns;              // ceil(aw / kSmoothRun)
};

// drawCorrect (images.nim:405-445): every pixel sampled at inv * (x+h, y+h).
struct CorrectArgs {
  uint32_t* dst;
  int32_t aw, ah;
  DrawSrc src;
  float inv[9];               // the inverse transform after minify / magnify
  int32_t mode;
  int32_t wrapped;            // drawTiled
};

int launch_minify(const MinifyArgs& a, hipStream_t s);
int launch_magnify(const MagnifyArgs& a, hipStream_t s);
int launch_draw_smooth(const SmoothArgs& a, hipStream_t s);
int launch_draw_correct(const CorrectArgs& a, hipStream_t s);

}  // namespace pxh
