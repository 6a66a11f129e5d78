/*
 * pixie_hip.h -- C-ABI of libpixie_hip.so, the MI355X (gfx950) raster hot path
 * of pixie v6.0.0.
 *
 * Plain C types only (pointers, sizes, ints, floats).  Images are premultiplied
 * RGBA8 (r,g,b,a bytes per pixel, row-major, stride = width) kept resident in
 * HBM; 8-bit alpha images hold one byte per pixel and read as rgbx(v,v,v,v).
 * Colours cross the boundary already premultiplied (pixie converts its Color
 * with chroma's asRgbx before fillShapes, paths.nim:1603).
 *
 * Every int-returning function returns 0 on success or a PXH_ERR_* code; the
 * message is available from pxh_last_error() (thread-local), mirroring the
 * PixieError text the reference raises (genny's checkError/takeError model,
 * bindings/bindings.nim:3-10).  Calls are thread-compatible: one host thread
 * drives the library at a time.
 *
 * Enum ordinals are pixie's: BlendMode (common.nim:6-29), WindingRule
 * (paths.nim:5-8), LineCap / LineJoin (paths.nim:10-16), path command kinds
 * (paths.nim:18-22).
 */
#ifndef PIXIE_HIP_H
#define PIXIE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PXH_OK = 0,
  PXH_ERR_DISCRETIZE = 1,   /* "Unable to discretize cubic/quadratic/arc" */
  PXH_ERR_INT_OVERFLOW = 2, /* "Path int overflow detected" */
  PXH_ERR_DASH = 3,         /* "Invalid line dash value" */
  PXH_ERR_NEG_BLUR = 4,     /* "Cannot apply negative blur" */
  PXH_ERR_PARSE = 5,        /* "Invalid path ..." */
  PXH_ERR_ARG = 6,
  PXH_ERR_NOMEM = 7,
  PXH_ERR_DEVICE = 8,
  PXH_ERR_CAPACITY = 9
};

enum { PXH_FORMAT_RGBA8 = 0, PXH_FORMAT_A8 = 1 };

typedef struct pxh_image pxh_image;

/* Library ---------------------------------------------------------------- */
int pxh_init(void);                 /* selects HIP device 0; idempotent */
const char* pxh_last_error(void);
const char* pxh_version(void);
int pxh_synchronize(void);          /* waits for all queued device work */

/* Images ---------------------------------------------------------------- */
/* newImage (common.nim:39-47): zero-filled, width and height > 0. */
pxh_image* pxh_image_new(int32_t width, int32_t height);
pxh_image* pxh_image_new_a8(int32_t width, int32_t height);
void pxh_image_free(pxh_image* img);
int32_t pxh_image_width(const pxh_image* img);
int32_t pxh_image_height(const pxh_image* img);
int32_t pxh_image_format(const pxh_image* img);
void* pxh_image_device_ptr(pxh_image* img);
/* Host <-> device copies of the whole image (width*height*bpp bytes). */
int pxh_image_upload(pxh_image* img, const void* data);
int pxh_image_download(const pxh_image* img, void* data);
/* copy (common.nim:49-54) into an existing image of the same size/format. */
int pxh_image_copy(pxh_image* dst, const pxh_image* src);
/* fill (images.nim:55-57 / internal.nim:55-66). */
int pxh_image_fill(pxh_image* img, uint32_t rgbx);

/* Paths ----------------------------------------------------------------- */
/* parsePath (paths.nim:119-262): SVG path string -> pixie command array
 * (kind, params...).  *cmds is malloc'd; release with pxh_free. */
int pxh_parse_path(const char* path, float** cmds, int32_t* ncmds);
void pxh_free(void* p);

/* fillPath (paths.nim:2093-2113) with a solid paint: cmds is a pixie command
 * array, mat9 a vmath Mat3 (column-major, NULL = identity), rgbx the paint
 * colour with opacity applied, premultiplied (the fillShapes argument).  The
 * caller keeps fillPath's early returns, which test the float colour before
 * it is converted (opacity == 0; color.a == 0 unless OverwriteBlend,
 * paths.nim:2101-2107): this function always rasterises. */
int pxh_fill_path(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                  uint32_t rgbx, int32_t winding_rule, int32_t blend_mode);

/* strokePath (paths.nim:2144-2176) with a solid paint; the caller keeps the
 * early returns of paths.nim:2157-2162 as for pxh_fill_path. */
int pxh_stroke_path(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                    uint32_t rgbx, float stroke_width, int32_t line_cap, int32_t line_join,
                    float miter_limit, const float* dashes, int32_t ndashes,
                    int32_t blend_mode);

/* Transformed image draws.
 * draw(a, b, transform, blendMode) (images.nim:636-678): minifyBy2 /
 * magnifyBy2 until the sampling step is within (1/2, 2), then the integer
 * composite when the result is a whole-pixel translation, else drawSmooth's
 * bilinear samples (images.nim:531-634).  drawTiled (images.nim:680-683):
 * drawCorrect with wrapped samples over every pixel.  Both images RGBA8 and
 * distinct.  pixie loops forever on a transform whose sampling step is 0 or
 * infinite; here that returns PXH_ERR_ARG. */
int pxh_draw_transform(pxh_image* a, const pxh_image* b, const float* mat9, int32_t blend_mode);
int pxh_draw_tiled(pxh_image* a, const pxh_image* b, const float* mat9, int32_t blend_mode);
/* minifyBy2 / magnifyBy2 (images.nim:168-259): new images (NULL on error,
 * pxh_last_error() says why). */
pxh_image* pxh_minify_by2(const pxh_image* img, int32_t power);
pxh_image* pxh_magnify_by2(const pxh_image* img, int32_t power);

/* Paints (paints.nim:4-29): the PaintKind ordinal, blend mode and opacity;
 * a solid paint (0) carries its colour (straight float r, g, b, a), image
 * paints (1, 2) the image and imageMat, gradients (3-5) the handle positions
 * (image space: 2 points for linear, 3 for radial and angular) and colour
 * stops (straight float r, g, b, a and the stop position, in order). */
enum {
  PXH_SOLID_PAINT = 0,
  PXH_IMAGE_PAINT = 1,
  PXH_TILED_IMAGE_PAINT = 2,
  PXH_LINEAR_GRADIENT = 3,
  PXH_RADIAL_GRADIENT = 4,
  PXH_ANGULAR_GRADIENT = 5
};
typedef struct {
  int32_t kind;
  int32_t blend_mode;
  float opacity;
  int32_t nhandles;        /* gradientHandlePositions.len */
  float handles[6];
  int32_t nstops;
  const float* stops;      /* nstops x (r, g, b, a, position) */
  const pxh_image* image;  /* ImagePaint / TiledImagePaint, RGBA8 */
  float image_mat[9];      /* imageMat, m[3 * col + row] */
  float color[4];          /* SolidPaint: color (chroma Color, straight) */
} pxh_paint;
#define PXH_MAX_STOPS 32

/* fillPath / strokePath with any paint (paths.nim:2093-2214).  A solid
 * paint takes fillPath's own route: opacity clamped to [0, 1], nothing drawn
 * at opacity 0 or for a transparent colour unless it overwrites, else the
 * colour with the opacity multiplied into its alpha, converted as chroma's
 * asRgbx, filled as pxh_fill_path / pxh_stroke_path fill.  For the non-solid
 * kinds pixie builds a full-size coverage mask, a full-size fill
 * (the image drawn with imageMat, tiled or not, or the gradient), masks one
 * with the other and draws the result; here the mask is an 8-bit plane and
 * one pass masks the fill (a gradient evaluated per pixel, or the drawn
 * image) and blends it.  Opacity 0 draws nothing; gradient errors are
 * pixie's ("Gradient must have at least 1 color stop", "Linear gradient
 * requires 2 handles", ...), raised after the mask is built and before the
 * image changes. */
int pxh_fill_path_paint(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                        const pxh_paint* paint, int32_t winding_rule);
int pxh_stroke_path_paint(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                          const pxh_paint* paint, float stroke_width, int32_t line_cap,
                          int32_t line_join, float miter_limit, const float* dashes,
                          int32_t ndashes);

/* fillShapes (paths.nim:1593-1912): flattened, transformed polygons given as
 * xy pairs and polygon start offsets (npolys + 1 entries, in points). */
int pxh_fill_shapes(pxh_image* img, const float* xy, const int32_t* poly_ofs, int32_t npolys,
                    uint32_t rgbx, int32_t winding_rule, int32_t blend_mode);

/* Scenes (batched form) ------------------------------------------------- */
/* A scene records an ordered sequence of fillPath / strokePath calls
 * (paths.nim:2093-2113 / 2144-2176; solid paints, dashes and elliptical arcs
 * included) and renders them in one device pass.  The result is the image
 * the same calls leave when issued one by one in order.  Pixie has no batched
 * entry point; this one replaces a loop over fillPath / strokePath on one
 * image.  As for pxh_fill_path the caller keeps fillPath's paint checks;
 * strokes of width <= 0 (no stroke shapes, paths.nim:1931) are not
 * recorded.  The recording stays
 * resident on the device across renders; scene pools are sized on the first
 * render at a given image size (which waits for the device once) and reused
 * afterwards without host synchronisation.  Paths with elliptical arc
 * commands are flattened on the host when recorded (a flattening error is
 * returned by the record call); everything else is flattened on the device.
 * Dashes follow pxh_stroke_path (an invalid dash value is reported through
 * pxh_scene_status). */
typedef struct pxh_scene pxh_scene;
pxh_scene* pxh_scene_new(void);
void pxh_scene_free(pxh_scene* scene);
int pxh_scene_clear(pxh_scene* scene);
int32_t pxh_scene_count(const pxh_scene* scene); /* recorded calls */
int pxh_scene_fill_path(pxh_scene* scene, const float* cmds, int32_t ncmds, const float* mat9,
                        uint32_t rgbx, int32_t winding_rule, int32_t blend_mode);
int pxh_scene_stroke_path(pxh_scene* scene, const float* cmds, int32_t ncmds, const float* mat9,
                          uint32_t rgbx, float stroke_width, int32_t line_cap,
                          int32_t line_join, float miter_limit, const float* dashes,
                          int32_t ndashes, int32_t blend_mode);
/* Renders the recorded calls onto img (queued on the library's stream). */
int pxh_scene_render(pxh_scene* scene, pxh_image* img);
/* Autotuner (no reference counterpart): renders the scene onto a scratch
 * image of img's size with each phase-A launch shape (band entries kept in
 * LDS x row waves per CU, and whether tasks of large bands run in a launch
 * of their own), `reps` timed renders each, and keeps the fastest for later
 * renders.  *shape receives its index, *ms its time per render.  Results are
 * cached for the process, keyed by device arch, target size and format, and
 * the scene's call count and command count (to a power of two): a scene of
 * that key takes the cached shape without timing. */
int pxh_scene_autotune(pxh_scene* scene, pxh_image* img, int32_t reps, int32_t* shape,
                       float* ms);
/* Sets the launch shape directly (0..5; 0, one launch of every task, is the
 * default). */
int pxh_scene_set_shape(pxh_scene* scene, int32_t shape);
/* 1 when renders of this scene replay a captured HIP graph (after the first
 * render at a given image, lane and launch shape), 0 when they run eagerly. */
int32_t pxh_scene_graph_active(const pxh_scene* scene);
/* Per-call status of the last render: 0, or the PXH_ERR_* code the call
 * would have raised (e.g. PXH_ERR_INT_OVERFLOW); such calls draw nothing. */
int pxh_scene_status(pxh_scene* scene, int32_t* status, int32_t n);
/* Device geometry of recorded call i after a render, in the layout of
 * pxh_debug_fill_geometry (test access). */
int pxh_scene_debug_geometry(pxh_scene* scene, int32_t i, int32_t* info16, int32_t** parts4,
                             int32_t* nparts, float** entries4, int8_t** windings,
                             int32_t* nentries);
/* Test access: lowers the run-pool capacity the next renders see to `cap`
 * entries (never above the validated size), so that a test can force a
 * replayed render to run out of the pool.  Such a render leaves its canvas
 * untouched and its error surfaces at the lane's next synchronize
 * (pxh_synchronize / pxh_lane_synchronize / pxh_scene_status), after which
 * the scene re-validates on its next render. */
int pxh_scene_debug_limit_runs(pxh_scene* scene, int32_t cap);

/* Execution lanes (no reference counterpart: pixie is single-threaded).
 * The library owns pxh_lanes() independent device streams.  Lane 0 is the
 * one every other function uses.  Work issued on different lanes, each
 * driven by its own host thread, overlaps on the device.  Lanes have their
 * own temporaries, so one host thread per lane may call the *_on functions
 * concurrently, on distinct scenes and images.  Lane i's work is ordered
 * after its earlier work only. */
int32_t pxh_lanes(void);
int pxh_lane_synchronize(int32_t lane);
int pxh_scene_render_on(pxh_scene* scene, pxh_image* img, int32_t lane);
int pxh_blur_on(pxh_image* img, float radius, uint32_t oob_rgbx, int32_t lane);
int pxh_image_fill_on(pxh_image* img, uint32_t rgbx, int32_t lane);
/* A non-owning image over caller-owned device memory (width*height*bpp
 * bytes); pxh_image_free releases only the handle. */
pxh_image* pxh_image_view(void* device_ptr, int32_t width, int32_t height, int32_t format);

/* Serving (no reference counterpart: pixie is a library with no server).
 * A server renders scene requests with continuous batching: requests queue
 * in arrival order, and each of n_lanes worker threads (lanes 1..n_lanes,
 * so 1 <= n_lanes < pxh_lanes(); lane 0 stays the caller's) takes the oldest
 * queued request as soon as its canvas fits in the server's resident HBM
 * arena of hbm_budget bytes, fills the canvas transparent, renders the scene
 * on its lane and blurs it (images.nim:304-365) if blur_radius != 0.
 * Requests admitted while others run overlap on the device.  A request's
 * scene must stay alive and unchanged, and in no other unfinished request,
 * until the request is done.  A failing request (for example a negative
 * blur radius) is done with its PXH_ERR_* status and does not affect the
 * others.  A finished request's canvas (RGBA8, width x height) stays in the
 * arena until pxh_server_release: release requests to keep admitting. */
typedef struct pxh_server pxh_server;
pxh_server* pxh_server_new(int32_t n_lanes, int64_t hbm_budget);
/* Waits for running requests, drops queued ones and frees the arena. */
void pxh_server_free(pxh_server* server);
/* Queues a request; returns its ticket (>= 0), or -PXH_ERR_ARG /
 * -PXH_ERR_CAPACITY (canvas larger than the arena). */
int64_t pxh_server_enqueue(pxh_server* server, pxh_scene* scene, int32_t width, int32_t height,
                           float blur_radius);
/* *done = 1 once the request finished, with its status in *status. */
int pxh_server_poll(pxh_server* server, int64_t ticket, int32_t* done, int32_t* status);
int pxh_server_wait(pxh_server* server, int64_t ticket, int32_t* status);
/* The finished request's canvas (a view into the arena, valid until
 * release), its error text, and its enqueue / admission / completion times
 * in seconds on the server's clock (pxh_server_now). */
pxh_image* pxh_server_image(pxh_server* server, int64_t ticket);
const char* pxh_server_message(pxh_server* server, int64_t ticket);
int pxh_server_times(pxh_server* server, int64_t ticket, double* enqueued, double* admitted,
                     double* done);
double pxh_server_now(pxh_server* server);
/* Returns a finished request's canvas to the arena and forgets the ticket. */
int pxh_server_release(pxh_server* server, int64_t ticket);

/* Text (fonts.nim, fontformats/opentype.nim) ---------------------------- */
/* A typeface: an OpenType font with TrueType (glyf) or CFF (Type 2
 * charstring) outlines parsed from bytes (copied), as parseTtf / parseOtf
 * (fonts.nim:500-505, opentype.nim:2531-2588): cmap formats 4 and 12, head,
 * hhea, maxp, hmtx, name, OS/2, loca + glyf or CFF (opentype.nim:862-1678),
 * kern, GPOS pair adjustments, post.  Errors are
 * pixie's PixieError texts ("Unexpected error reading font data, EOF",
 * "Missing required font table: ...", "Unsupported font ...").  Glyph paths
 * are cached per rune on first use (opentype.nim:2456-2463). */
typedef struct pxh_typeface pxh_typeface;
pxh_typeface* pxh_typeface_parse(const uint8_t* data, int64_t len);
/* parseOpenTypeCollection (opentype.nim:2591-2616): face `index` of a .ttc;
 * *count (optional) receives the number of faces. */
pxh_typeface* pxh_typeface_parse_collection(const uint8_t* data, int64_t len, int32_t index,
                                            int32_t* count);
void pxh_typeface_free(pxh_typeface* tf);
/* Typeface.fallbacks (fonts.nim:12-16): replaces the list.  The typefaces
 * are borrowed and must outlive tf's use. */
int pxh_typeface_set_fallbacks(pxh_typeface* tf, pxh_typeface* const* fallbacks, int32_t n);
/* In font units (fonts.nim:59-115): out[0..8] = scale (unitsPerEm), ascent,
 * descent, lineGap, capHeight, underlinePosition, underlineThickness,
 * strikeoutPosition, strikeoutThickness. */
int pxh_typeface_metrics(const pxh_typeface* tf, float* out9);
/* fullName (opentype.nim:2523-2529), NUL-terminated into buf (cap bytes);
 * returns the name's length. */
int32_t pxh_typeface_name(const pxh_typeface* tf, char* buf, int32_t cap);
/* hasGlyph (fonts.nim:117-122) of tf itself; fallbacks are not searched. */
int32_t pxh_typeface_has_glyph(const pxh_typeface* tf, uint32_t rune);
/* getAdvance / getKerningAdjustment (fonts.nim:163-195), in font units,
 * through the fallbacks. */
float pxh_typeface_advance(const pxh_typeface* tf, uint32_t rune);
float pxh_typeface_kerning(const pxh_typeface* tf, uint32_t left, uint32_t right);
/* getGlyphPath (fonts.nim:145-161): the rune's outline in font units, y up
 * flipped to y down, through the fallbacks (empty when none has it).
 * *cmds is malloc'd; release with pxh_free. */
int pxh_typeface_glyph_path(pxh_typeface* tf, uint32_t rune, float** cmds, int32_t* ncmds);

/* A Font (fonts.nim:17-26).  line_height < 0 is autoLineHeight; text_case
 * is the TextCase ordinal (upper / lower / title case convert ASCII and
 * Latin-1 letters); paints are font.paints, drawn in order by
 * pxh_fill_text / pxh_stroke_text. */
typedef struct {
  pxh_typeface* typeface;
  float size;
  float line_height;
  int32_t text_case;
  int32_t underline;
  int32_t strikethrough;
  int32_t no_kerning_adjustments;
  int32_t npaints;
  const pxh_paint* paints;
} pxh_font;
/* A Span (fonts.nim:28-30): UTF-8 text (len bytes) and its font. */
typedef struct {
  const char* text;
  int64_t len;
  const pxh_font* font;
} pxh_span;

/* typeset (fonts.nim:256-466): the arrangement of the spans' runes (control
 * runes other than LF dropped) with per-rune positions and selection rects,
 * lines and spans.  The arrangement keeps the pxh_font pointers, as pixie's
 * keeps its Font refs: the fonts, their typefaces and paints must outlive
 * it, and their current values are read when it is filled or stroked.
 * h_align / v_align are the HorizontalAlignment / VerticalAlignment
 * ordinals.  NULL on error. */
typedef struct pxh_arrangement pxh_arrangement;
pxh_arrangement* pxh_typeset(const pxh_span* spans, int32_t nspans, float bounds_x, float bounds_y,
                             int32_t h_align, int32_t v_align, int32_t wrap);
void pxh_arrangement_free(pxh_arrangement* a);
/* Sizes: what = 0 runes, 1 lines, 2 spans. */
int32_t pxh_arrangement_count(const pxh_arrangement* a, int32_t what);
/* Copies out runes (n), positions (2n), selectionRects (4n: x, y, w, h),
 * lines and spans ((start, stop) pairs); any pointer may be NULL. */
int pxh_arrangement_read(const pxh_arrangement* a, uint32_t* runes, float* positions,
                         float* selection_rects, int32_t* lines, int32_t* spans);
/* layoutBounds (fonts.nim:478-492): out2 = width, height. */
int pxh_arrangement_layout_bounds(const pxh_arrangement* a, float* out2);
/* computePaths (fonts.nim:517-564): span `span`'s path (glyphs placed and
 * scaled, underline / strikethrough rects).  *cmds malloc'd (pxh_free). */
int pxh_arrangement_span_path(const pxh_arrangement* a, int32_t span, float** cmds,
                              int32_t* ncmds);
/* computeBounds (fonts.nim:590-597): out4 = x, y, w, h. */
int pxh_arrangement_compute_bounds(const pxh_arrangement* a, const float* mat9, float* out4);
/* fillText / strokeText of an arrangement (fonts.nim:566-588, 599-690):
 * every span's path filled (NonZero) or stroked with each of its font's
 * paints in order, through pxh_fill_path_paint / pxh_stroke_path_paint. */
int pxh_fill_text(pxh_image* img, const pxh_arrangement* a, const float* mat9);
int pxh_stroke_text(pxh_image* img, const pxh_arrangement* a, const float* mat9, float stroke_width,
                    int32_t line_cap, int32_t line_join, float miter_limit, const float* dashes,
                    int32_t ndashes);

/* SVG (fileformats/svg.nim) -------------------------------------------- */
/* parseSvg (svg.nim:502-555): the document's elements (path, line,
 * polyline, polygon, rect, circle, ellipse inside nested g) as pixie path
 * commands with their resolved properties, and its linearGradient
 * resources.  width = height = 0 keeps the view box size, else the content
 * is scaled to width x height.  Errors are pixie's texts ("Invalid SVG
 * data", "Unsupported SVG tag: ...", ...).  NULL on error. */
typedef struct pxh_svg pxh_svg;
pxh_svg* pxh_svg_parse(const char* data, int64_t len, int32_t width, int32_t height);
void pxh_svg_free(pxh_svg* svg);
int pxh_svg_size(const pxh_svg* svg, int32_t* width, int32_t* height);
int32_t pxh_svg_count(const pxh_svg* svg); /* elements */
/* An element's properties (SvgProperties, svg.nim:19-33) as resolved. */
typedef struct {
  int32_t display;
  int32_t fill_rule;
  char fill[64];          /* "none", a colour, or "url(#id)" (truncated) */
  uint32_t stroke_rgbx;   /* premultiplied; 0 = no stroke */
  float stroke_width;
  int32_t line_cap;
  int32_t line_join;
  float miter_limit;
  int32_t ndashes;
  const float* dashes;    /* valid while the svg lives */
  float transform[9];
  float opacity;
  float fill_opacity;
  float stroke_opacity;
} pxh_svg_props;
/* Element i: its commands (*cmds malloc'd, pxh_free; either may be NULL)
 * and properties (may be NULL). */
int pxh_svg_element(const pxh_svg* svg, int32_t i, float** cmds, int32_t* ncmds,
                    pxh_svg_props* props);
/* newImage(svg) (svg.nim:557-602) into img (RGBA8, the svg's size): cleared
 * to transparent, then every displayed element filled and stroked in
 * order, the first fill overwriting.  Solid fills and strokes are drawn in
 * one scene pass; gradient fills take the paint route in order between
 * scene passes. */
int pxh_svg_render(pxh_svg* svg, pxh_image* img);

/* Image operations ------------------------------------------------------ */
/* draw(a, b, translate(dx, dy), blend) for integer translations, i.e.
 * blendRect (images.nim:468-529, reached from draw images.nim:636-678). */
int pxh_draw(pxh_image* a, const pxh_image* b, int32_t dx, int32_t dy, int32_t blend_mode);
/* blur (images.nim:304-365); oob is the premultiplied out-of-bounds colour. */
int pxh_blur(pxh_image* img, float radius, uint32_t oob_rgbx);
/* Allocation-free blur.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  Same result as pxh_blur, on lane 0, but
 * the intermediate image and the Gaussian table live in a static device
 * array the library holds on each GPU until the process ends, so it calls
 * no hipMalloc / hipFree.  Limits: RGBA8 images of at most
 * PXH_BLUR_NOALLOC_MAX_PIXELS pixels (w * h), radius as pxh_blur's LDS tiles
 * allow; larger inputs fail with PXH_ERR_CAPACITY.  Every call on a GPU
 * shares that GPU's array (each call overwrites all of it that it reads),
 * so calls that overlap on one GPU (other threads or streams) race. */
#define PXH_BLUR_NOALLOC_MAX_PIXELS 4194304
int pxh_blur_noalloc(pxh_image* img, float radius, uint32_t oob_rgbx);
/* spread (images.nim:700-758). */
int pxh_spread(pxh_image* img, float spread);
/* shadow (images.nim:760-776) into `out` (same size as src).  Whole-pixel
 * offsets copy through the integer composite, fractional ones through
 * drawSmooth, as draw(image, translate(offset), OverwriteBlend) does. */
int pxh_shadow(pxh_image* out, const pxh_image* src, float offset_x, float offset_y,
               float spread, float blur, uint32_t rgbx);
/* applyOpacity (images.nim:261-277). */
int pxh_apply_opacity(pxh_image* img, float opacity);

/* Pixel ops on device buffers (test and benchmark access to blender(),
 * blends.nim:275-299): out[i] = blend(mode, backdrop[i], source[i]). */
int pxh_blend_pixels(int32_t blend_mode, const uint32_t* backdrop, const uint32_t* source,
                     uint32_t* out, int64_t n);

/* Measurement (no reference counterpart: pixie has no device timeline). */
/* Pixels written by fills since the last reset (N_touch); reset != 0 zeroes. */
int pxh_touched_pixels(int64_t* out, int32_t reset);
/* Records event slot 0..15 on the library's stream / elapsed ms a -> b. */
int pxh_event_record(int32_t slot);
int pxh_event_elapsed(int32_t a, int32_t b, float* ms);
/* Per-launch timing of raster kernels between begin and end (summed ms). */
int pxh_profile_begin(void);
int pxh_profile_end(double* kernel_ms, int64_t* launches);
/* The share of one kernel (or scene stage) in the last profile: names are
 * fill_rows_kernel, scene_geometry, scene_bands, scene_rows_kernel,
 * scene_compose_kernel. */
int pxh_profile_kernel(const char* name, double* kernel_ms, int64_t* launches);

/* Host geometry of one fillPath (stroke = 0) or strokePath (stroke = 1) call
 * without device work: info16 receives the 16-int fill header (startX,
 * startY, pathWidth, pathHeight, nParts, bandH, ...), parts4 4 ints per band
 * (entryBase, count, flags, 0), entries4 4 floats per entry (at.y, to.y, m,
 * b) and windings one int8 per entry.  Arrays are released with pxh_free.
 * Mirrors the locals of fillShapes (paths.nim:1602-1629) and the output of
 * partitionSegments (paths.nim:1168-1262); used by the test-suite. */
int pxh_debug_fill_geometry(const float* cmds, int32_t ncmds, const float* mat9, int32_t w,
                            int32_t h, int32_t stroke, float stroke_width, int32_t line_cap,
                            int32_t line_join, float miter_limit, int32_t* info16,
                            int32_t** parts4, int32_t* nparts, float** entries4,
                            int8_t** windings, int32_t* nentries);

#ifdef __cplusplus
}
#endif
#endif
