/*
 * pixie_oracle.h -- CPU restatement of pixie v6.0.0's scalar raster path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: only the
 * test-suite, the smoke check and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product (libpixie_hip.so) never links or calls it.
 *
 * Semantics follow pixie's scalar build (-d:pixieNoSimd): every function cites
 * the Nim source it restates.  Float math is IEEE float32 in pixie's operation
 * order; build with -ffp-contract=off (see oracle/Makefile).
 *
 * Images are row-major premultiplied RGBA8 (r,g,b,a bytes), stride = width,
 * exactly pixie's Image.data layout (common.nim:34-37).
 */
#ifndef PIXIE_ORACLE_H
#define PIXIE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint8_t r, g, b, a; } orc_rgbx;

/* BlendMode ordinals (common.nim:6-29). */
enum {
  ORC_NORMAL = 0, ORC_DARKEN, ORC_MULTIPLY, ORC_COLOR_BURN, ORC_LIGHTEN,
  ORC_SCREEN, ORC_COLOR_DODGE, ORC_OVERLAY, ORC_SOFT_LIGHT, ORC_HARD_LIGHT,
  ORC_DIFFERENCE, ORC_EXCLUSION, ORC_HUE, ORC_SATURATION, ORC_COLOR,
  ORC_LUMINOSITY, ORC_MASK, ORC_OVERWRITE, ORC_SUBTRACT_MASK, ORC_EXCLUDE_MASK,
  ORC_NUM_BLEND_MODES
};

/* WindingRule (paths.nim:5-8), LineCap / LineJoin (paths.nim:10-16). */
enum { ORC_NONZERO = 0, ORC_EVENODD = 1 };
enum { ORC_BUTT = 0, ORC_ROUND_CAP = 1, ORC_SQUARE = 2 };
enum { ORC_MITER = 0, ORC_ROUND_JOIN = 1, ORC_BEVEL = 2 };

/* PathCommandKind ordinals (paths.nim:18-22). */
enum {
  ORC_CMD_CLOSE = 0, ORC_CMD_MOVE, ORC_CMD_LINE, ORC_CMD_HLINE, ORC_CMD_VLINE,
  ORC_CMD_CUBIC, ORC_CMD_SCUBIC, ORC_CMD_QUAD, ORC_CMD_TQUAD, ORC_CMD_ARC,
  ORC_CMD_RMOVE, ORC_CMD_RLINE, ORC_CMD_RHLINE, ORC_CMD_RVLINE, ORC_CMD_RCUBIC,
  ORC_CMD_RSCUBIC, ORC_CMD_RQUAD, ORC_CMD_RTQUAD, ORC_CMD_RARC
};

/* Error codes: 0 = ok, otherwise a PixieError was raised. */
enum {
  ORC_OK = 0, ORC_ERR_DISCRETIZE = 1, ORC_ERR_INT_OVERFLOW = 2,
  ORC_ERR_DASH = 3, ORC_ERR_NEG_BLUR = 4, ORC_ERR_PARSE = 5, ORC_ERR_ARG = 6,
  ORC_ERR_NOMEM = 7
};

/* Counters the oracle keeps while filling (for fixtures and the bench's
 * algorithmic-byte count). */
typedef struct {
  int64_t touched;      /* pixels blended or written (N_touch) */
  int64_t rows_trap;    /* rows taken by the trapezoid path */
  int64_t rows_aa;      /* rows taken by the 5-sample coverage path */
  int64_t rows_nonaa;   /* rows taken by the 1-sample (fillHits) path */
  int64_t rows_span2;   /* rows taken by the two-spanning non-AA fill */
  int64_t dropped;      /* coverage writes outside the path span (UB in pixie) */
  int64_t segments;     /* segments after shapesToSegments */
  int64_t partitions;   /* partitions built */
  int64_t max_part;     /* largest partition (entries) */
  int64_t max_sel;      /* most entries selected for one row */
} orc_stats;

const char* orc_last_error(void);

/* ---- pixel ops (blends.nim, common.nim) ---- */
orc_rgbx orc_blend(int mode, orc_rgbx backdrop, orc_rgbx source);
orc_rgbx orc_mul_cov(orc_rgbx c, uint8_t cov);       /* common.nim:79-90 */
orc_rgbx orc_mul_opacity(orc_rgbx c, float opacity);  /* common.nim:67-77 */
/* Blend n pixel pairs: out[i] = blend(mode, backdrop[i], source[i]). */
void orc_blend_pixels(int mode, const uint8_t* backdrop, const uint8_t* source,
                      uint8_t* out, int64_t n);
/* chroma conversions the blend modes go through (parity unpinned). */
orc_rgbx orc_rgbx_from_rgba(orc_rgbx straight);
orc_rgbx orc_rgba_from_rgbx(orc_rgbx premul);
/* Straight-alpha float colour -> premultiplied rgbx (chroma asRgbx). */
orc_rgbx orc_rgbx_from_color(float r, float g, float b, float a);
void orc_to_straight_alpha(uint8_t* data, int64_t n); /* internal.nim:68-84 */
void orc_to_premultiplied_alpha(uint8_t* data, int64_t n); /* internal.nim:86-96 */

/* ---- image ops (images.nim, internal.nim) ---- */
void orc_gaussian_kernel(int radius, uint16_t* out /* 2r+1 */);
void orc_fill(uint8_t* img, int w, int h, orc_rgbx c);
void orc_apply_opacity(uint8_t* img, int w, int h, float opacity);
int orc_blur(uint8_t* img, int w, int h, float radius, orc_rgbx oob);
int orc_spread(uint8_t* img, int w, int h, float spread);
int orc_shadow(uint8_t* out, const uint8_t* img, int w, int h,
               float offx, float offy, float spread, float blur,
               orc_rgbx color);
/* draw(a, b, translate(px, py), mode) for integer translations = blendRect. */
void orc_draw_rect(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh,
                   int px, int py, int mode);
/* images.nim:138 diff score (100 * sum|d| / (255*4*w*h)), equal sizes. */
float orc_diff_score(const uint8_t* master, const uint8_t* img, int w, int h);

/* ---- paths (paths.nim) ---- */
/* Parse an SVG path string into pixie's flat command array.  *out is
 * malloc'd; caller frees with orc_free. */
int orc_parse_path(const char* s, float** out, int* n);
void orc_free(void* p);

/* Flatten (commandsToShapes) and transform; returns polygons as a flat xy
 * array plus polygon start offsets (npolys+1 entries, in points). */
int orc_flatten(const float* cmds, int ncmds, const float* mat9, int close,
                float** xy, int** poly_ofs, int* npolys);
/* strokeShapes + transform. */
int orc_stroke_shapes(const float* cmds, int ncmds, const float* mat9,
                      float width, int cap, int join, float miter_limit,
                      const float* dashes, int ndashes,
                      float** xy, int** poly_ofs, int* npolys);
/* fillShapes on already flattened, transformed polygons. */
int orc_fill_shapes(uint8_t* img, int w, int h, const float* xy,
                    const int* poly_ofs, int npolys, orc_rgbx rgbx,
                    int winding, int blend, orc_stats* st);
/* fillPath / strokePath with a solid premultiplied colour (the colour is
 * already converted and opacity-applied, as fillPath does before
 * fillShapes, paths.nim:2101-2112). */
int orc_fill_path(uint8_t* img, int w, int h, const float* cmds, int ncmds,
                  const float* mat9, orc_rgbx rgbx, int winding, int blend,
                  orc_stats* st);
int orc_stroke_path(uint8_t* img, int w, int h, const float* cmds, int ncmds,
                    const float* mat9, orc_rgbx rgbx, float width, int cap,
                    int join, float miter_limit, const float* dashes,
                    int ndashes, int blend, orc_stats* st);

/* Debug views of the geometry pipeline, for cross-checking the device
 * pipeline.  Segments: 4 floats (at.x, at.y, to.x, to.y) + winding. */
int orc_segments(const float* xy, const int* poly_ofs, int npolys,
                 float** segs, int16_t** winding, int* nseg);
/* Partitions of one fillShapes call: per partition (top, bottom, nentries,
 * flags bit0=AA bit1=twoSpanning) and entries (aty, toy, m, b, atx, tox,
 * winding as float) in partition order. */
int orc_partitions(const float* xy, const int* poly_ofs, int npolys, int w,
                   int h, int** parts, int* nparts, float** entries,
                   int* nentries, int* bounds /* startX,startY,pathW,pathH */);

/* fillOverlaps / strokeOverlaps (paths.nim:2216-2277). */
int orc_fill_overlaps(const float* cmds, int ncmds, const float* mat9,
                      float x, float y, int winding, int* result);

/* ---- transformed draws (images.nim:168-259, 367-445, 531-683) ---- */
orc_rgbx orc_mix(orc_rgbx a, orc_rgbx b, float t);   /* common.nim:59-65 */
void orc_minify_size(int w, int h, int power, int* ow, int* oh);
int orc_minify_by2(const uint8_t* src, int w, int h, int power, uint8_t* dst);
int orc_magnify_by2(const uint8_t* src, int w, int h, int power, uint8_t* dst);
orc_rgbx orc_get_rgba_smooth(const uint8_t* img, int w, int h, float x, float y, int wrapped);
/* draw(a, b, transform, blendMode) and drawTiled (drawCorrect, tiled) */
int orc_draw(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh, const float* mat9,
             int mode);
int orc_draw_correct(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh,
                     const float* mat9, int mode, int tiled);
int orc_draw_tiled(uint8_t* a, int aw, int ah, const uint8_t* b, int bw, int bh, const float* mat9,
                   int mode);

/* ---- non-solid paints (paints.nim:12-248, paths.nim:2115-2142) ---- */
enum {
  ORC_IMAGE_PAINT = 1, ORC_TILED_IMAGE_PAINT = 2,
  ORC_LINEAR_GRADIENT = 3, ORC_RADIAL_GRADIENT = 4, ORC_ANGULAR_GRADIENT = 5
};
typedef struct { float r, g, b, a, pos; } orc_stop;  /* straight float colour, position */
typedef struct {
  int kind;              /* PaintKind ordinal (paints.nim:4-10) */
  int mode;              /* blendMode */
  float opacity;
  int nhandles;          /* gradientHandlePositions.len */
  float h[6];            /* gradientHandlePositions: 2 (linear) or 3 points */
  int nstops;
  const orc_stop* stops;
  const uint8_t* image;  /* ImagePaint / TiledImagePaint: premultiplied RGBA8 */
  int iw, ih;
  float image_mat[9];    /* imageMat */
} orc_paint;
int orc_fill_gradient(uint8_t* img, int w, int h, const orc_paint* p);
void orc_radial_matrix(const orc_paint* p, float* m9);
int orc_fill_path_paint(uint8_t* img, int w, int h, const float* cmds, int ncmds,
                        const float* mat9, const orc_paint* p, int winding, orc_stats* st);

#ifdef __cplusplus
}
#endif
#endif
