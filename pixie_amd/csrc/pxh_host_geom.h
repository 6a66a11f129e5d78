// pxh_host_geom.h -- host-side geometry of pixie's fill/stroke path:
// parsePath, commandsToShapes, strokeShapes, shapesToSegments and
// partitionSegments (paths.nim:119-1262, 1914-2082), producing the device
// layout (FillInfo / PartInfo / Entry) the raster kernels consume.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pixie_hip.h"
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

// Text calls (pxh_api.cpp): span paths drawn through the library's scene,
// their glyphs' canonical commands kept once in a device glyph arena and
// placed per glyph on the device.  An instance is a glyph (ref from
// text_glyph_ref) transformed by m (columns 0 and 1 and the translation of
// a Mat3; identity = copy), or, with ref < 0, path.rect(m[0], m[1], m[2],
// m[3]).
struct TextInst {
  int32_t ref;
  int32_t identity;
  float m[6];
};
int text_glyph_ref(const void* typeface, uint32_t rune, const float* cmds, int32_t n, int32_t* ref);
void text_glyphs_forget();
pxh_scene* text_scene();  // cleared
int text_record(pxh_scene* sc, const std::vector<TextInst>& inst, const float* mat9, uint32_t rgbx, int32_t rule,
                int32_t mode, bool stroke, float width, int32_t cap, int32_t join, float miterLimit,
                const float* dashes, int32_t ndashes);
int text_scene_render(pxh_image* img);
bool paint_solid_rgbx(const pxh_paint* p, uint32_t* rgbx);

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
