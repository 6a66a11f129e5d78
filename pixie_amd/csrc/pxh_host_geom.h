// pxh_host_geom.h -- host-side geometry of pixie's fill/stroke path:
// parsePath, commandsToShapes, strokeShapes, shapesToSegments and
// partitionSegments (paths.nim:119-1262, 1914-2082), producing the device
// layout (FillInfo / PartInfo / Entry) the raster kernels consume.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "pxh_common.h"
#include "pxh_geom.h"

namespace pxh {

// Polygons as pixie's seq[Polygon]: points plus start offsets (n+1 entries).
struct Shapes {
  std::vector<V2> pts;
  std::vector<int32_t> ofs{0};
  int32_t count() const { return (int32_t)ofs.size() - 1; }
  void clear() {
    pts.clear();
    ofs.assign(1, 0);
  }
  void add(const V2* p, int32_t n) {
    pts.insert(pts.end(), p, p + n);
    ofs.push_back((int32_t)pts.size());
  }
};

struct StrokeParams {
  float width = 1.0f;
  int32_t cap = 0;   // LineCap: Butt, Round, Square
  int32_t join = 0;  // LineJoin: Miter, Round, Bevel
  float miterLimit = 4.0f;
  const float* dashes = nullptr;
  int32_t ndashes = 0;
};

// Device-ready geometry of one fill.
struct FillGeom {
  FillInfo info;
  std::vector<PartInfo> parts;
  std::vector<Entry> entries;
  std::vector<int8_t> windings;
  int32_t maxCount = 0;  // largest band
};

// Sets the message pxh_last_error() returns and returns `code` (pxh_api.cpp).
int set_last_error(int code, const std::string& msg);

int32_t parse_path(const char* s, std::vector<float>* out, std::string* err);
float pixel_scale(const float* mat9);
bool mat_is_identity(const float* mat9);
int32_t commands_to_shapes(const float* cmds, int32_t n, bool closeSubpaths, float pixelScale,
                           Shapes* out, std::string* err);
void transform_shapes(Shapes* s, const float* mat9);
int32_t stroke_shapes(const Shapes& in, const StrokeParams& sp, float pixelScale, Shapes* out,
                      std::string* err);
// shapesToSegments + bounds + partitionSegments.  W, H: canvas size.
int32_t build_fill(const Shapes& s, int32_t W, int32_t H, uint32_t rgbx, int32_t mode,
                   int32_t rule, FillGeom* out, std::string* err);

}  // namespace pxh
