// pxh_api.cpp -- the C-ABI (include/pixie_hip.h): device context, images and
// the host side of each entry point (geometry, uploads, kernel launches).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pixie_hip.h"
#include "pxh_common.h"
#include "pxh_draw.h"
#include "pxh_host_geom.h"
#include "pxh_kernels.h"
#include "pxh_pixel.h"
#include "pxh_scene.h"

struct pxh_image {
  int32_t w, h, fmt;
  void* dev;
  size_t bytes;
  bool owned = true;  // false for views over caller memory
};

struct pxh_scene {
  std::vector<pxh::SceneItem> items;
  std::vector<int32_t> order;  // geometry workgroup -> call (built at upload)
  std::vector<pxh::SceneCmd> cmds;
  std::vector<pxh::ScenePoly> polys;
  std::vector<float> dashes;   // dash patterns of the recorded strokes
  // text calls: their commands and subpaths are written on the device by
  // glyph instances, in a region after the host-recorded ones (item index,
  // region-relative command and subpath bases)
  std::vector<pxh::GlyphInst> glyphInst;  // dstCmd / dstPoly region-relative
  std::vector<pxh::GlyphInst> glyphAbs;   // the uploaded copy, destinations absolute
  std::vector<std::array<int32_t, 3>> textItems;
  int32_t textCmds = 0, textPolys = 0;
  pxh::GlyphInst* dInst = nullptr;
  size_t dInstCap = 0;
  bool uploaded = false;
  int32_t validW = -1, validH = -1;  // target size the pools were validated for
  int32_t lane = 0;  // lane of the last render
  int32_t rowShape = pxh::kDefaultRowShape;  // phase-A launch shape (autotuned)
  // a validated render replays as a captured graph while its target, lane,
  // shape and buffers stay the same
  hipGraphExec_t exec = nullptr;
  bool graphsOff = false;
  uint64_t gen = 0;  // bumped whenever the recording or the pools change
  struct GraphKey {
    void* dev;
    int32_t w, h, fmt, lane, shape;
    const void* mem;
    const void* scratch;
    uint64_t gen;
    bool operator==(const GraphKey& o) const {
      return dev == o.dev && w == o.w && h == o.h && fmt == o.fmt && lane == o.lane &&
             shape == o.shape && mem == o.mem && scratch == o.scratch && gen == o.gen;
    }
  } key{};
  // device copies of the recording (grow-only)
  pxh::SceneItem* dItems = nullptr;
  pxh::SceneCmd* dCmds = nullptr;
  pxh::ScenePoly* dPolys = nullptr;
  int32_t* dOrder = nullptr;  // geometry workgroup -> call, longest call first
  int32_t nBig = 0;           // calls of more than 256 commands (first in order)
  int32_t bigCmds = 0;        // their commands
  float* dDashes = nullptr;
  size_t dItemsCap = 0, dCmdsCap = 0, dPolysCap = 0, dOrderCap = 0, dDashesCap = 0;
  // per-render buffers
  char* dMem = nullptr;
  size_t dMemBytes = 0;
  pxh::ScenePools pools{};
  int32_t ptCap = 0, segCap = 0, bandCap = 0, entryCap = 0, taskCap = 0, runCap = 0, binCap = 0;
  int32_t binSlots = 0;  // row bins the pools hold (height / kBinRows)
  int32_t gcap = 8192;   // band entries a large-band wave can hold in global row scratch
};

namespace pxh {
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define PXH_HIP(call)                                                            \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) return fail(kErrDevice, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// Growable device buffer.
constexpr int kLanes = 5;  // lane 0 + 4 lanes for concurrent requests

struct DevBuf {
  char* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return 0;
    if (p) {
      if (hipFree(p) != hipSuccess) return kErrDevice;
      p = nullptr;
      cap = 0;
    }
    size_t c = n < (1u << 20) ? (1u << 20) : n + n / 2;
    if (hipMalloc((void**)&p, c) != hipSuccess) return kErrNoMem;
    cap = c;
    return 0;
  }
};

// Static device scratch of pxh_blur_noalloc: the X pass's intermediate image
// (at most kNoAllocBlurPixels RGBA8 pixels, 16 MiB), then the Gaussian taps
// and the weight table built from them.
// Each GPU holds its own copy for the life of the process.  Host code never
// uses the name: it uses the address hipGetSymbolAddress gives for the device
// (Ctx::scratchPtr), taken once per device at init.
constexpr size_t kNoAllocBlurPixels = (size_t)1 << 22;
constexpr size_t kNoAllocLutWords = 4096;  // kMaxLutTaps taps + a weight table of up to 3072 dwords
static __device__ __align__(256) uint32_t scratch[kNoAllocBlurPixels + kNoAllocLutWords];
constexpr int kMaxDevices = 16;

struct Ctx {
  bool inited = false;
  int device = 0;                          // the device the library runs on
  uint32_t* scratchPtr[kMaxDevices] = {};  // device address of `scratch`, per device
  hipStream_t stream = nullptr;
  size_t maxLds = 65536;
  int numCUs = 256;
  // geometry upload: pinned host staging + device copy, reused per call once
  // the previous consumer kernel has finished (stageFree)
  char* hostStage = nullptr;
  size_t hostCap = 0;
  DevBuf devStage;
  hipEvent_t stageFree = nullptr;
  DevBuf gscratch;   // per-workgroup row scratch when it does not fit in LDS
  // Execution lanes: independent streams (lane 0 is `stream`), each with its
  // own blur temporary and scene row scratch, so that work issued on
  // different lanes (by different host threads) overlaps on the device.
  hipStream_t lane[kLanes] = {};
  // per lane: a second stream and fork / join events for phase A's two
  // concurrent launches
  hipStream_t laneAux[kLanes] = {};
  hipEvent_t laneFork[kLanes] = {}, laneJoin[kLanes] = {};
  DevBuf laneTmp[kLanes];
  DevBuf laneScratch[kLanes];
  std::mutex lutMu;
  DevBuf tmp;        // blur / shadow temporaries
  DevBuf tmp2;
  DevBuf plane;      // alpha plane
  DevBuf blendBuf;   // pxh_blend_pixels
  DevBuf minA, minB; // minifyBy2 steps of a transformed draw (ping-pong)
  DevBuf magBuf;     // a magnified source for the integer composite
  DevBuf paintFill;  // an image paint drawn onto a transparent canvas
  DevBuf drawRows;   // drawSmooth: per-row spans and run start positions
  pxh_scene* bigScene = nullptr;  // per-call fills / strokes of many commands
  // text: the canonical commands of every glyph drawn so far (grow-only),
  // keyed by (typeface, rune), and their device copy (re-uploaded as it grows)
  struct GlyphRec {
    int32_t cmd, nCmds, poly, nPolys;
  };
  std::map<std::pair<const void*, uint32_t>, int32_t> glyphKeys;
  std::vector<GlyphRec> glyphRecs;
  std::vector<SceneCmd> glyphCmds;
  std::vector<ScenePoly> glyphPolys;
  SceneCmd* dGlyphCmds = nullptr;
  ScenePoly* dGlyphPolys = nullptr;
  size_t dGlyphCmdsCap = 0, dGlyphPolysCap = 0, upGlyphCmds = 0, upGlyphPolys = 0;
  int32_t* errFlag = nullptr;
  // Per lane: error bits of scene renders that stopped before compose (an
  // overflowed pool); read and cleared when the lane is synchronized.  A
  // second row of kLanes words takes the reports of checked renders.
  int32_t* laneErr = nullptr;
  std::set<pxh_scene*> scenes;  // live scenes (their pools are re-validated after an error)
  std::mutex scenesMu;
  unsigned long long* touched = nullptr;  // N_touch counter of fills
  hipEvent_t ev[16] = {};
  // per-launch kernel timing (pxh_profile_begin / _end)
  bool profiling = false;
  std::vector<hipEvent_t> evPool;
  struct Timed {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Timed> evPairs;
  std::map<std::string, std::pair<double, int64_t>> kernelTimes;  // last profile: ms, launches
  std::map<int, uint32_t*> luts;  // per radius: device weight table of the blur kernels
  std::string arch;               // gcnArchName of the device
  // Autotuner results: (arch, target w, h, format, calls, commands bucket)
  // -> (phase-A shape, ms per render)
  std::map<std::string, std::pair<int, float>> tuned;
  std::mutex tunedMu;
};
Ctx g;

int init() {
  if (g.inited) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(kErrDevice, "no HIP device");
  PXH_HIP(hipSetDevice(0));
  g.device = 0;
  // the allocation-free blur's static scratch, addressed through the symbol's
  // device address on this device (the name itself is only a host stand-in)
  if (hipGetSymbolAddress((void**)&g.scratchPtr[g.device], HIP_SYMBOL(scratch)) != hipSuccess ||
      !g.scratchPtr[g.device])
    return fail(kErrDevice, "hipGetSymbolAddress(scratch) failed");
  int lds = 0, cus = 0;
  PXH_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0));
  PXH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  g.maxLds = (size_t)lds;
  g.numCUs = cus;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) g.arch = prop.gcnArchName;
  }
  PXH_HIP(hipStreamCreate(&g.stream));
  g.lane[0] = g.stream;
  for (int l = 1; l < kLanes; l++) PXH_HIP(hipStreamCreate(&g.lane[l]));
  for (int l = 0; l < kLanes; l++) {
    PXH_HIP(hipStreamCreate(&g.laneAux[l]));
    PXH_HIP(hipEventCreateWithFlags(&g.laneFork[l], hipEventDisableTiming));
    PXH_HIP(hipEventCreateWithFlags(&g.laneJoin[l], hipEventDisableTiming));
  }
  PXH_HIP(hipEventCreateWithFlags(&g.stageFree, hipEventDisableTiming));
  PXH_HIP(hipEventRecord(g.stageFree, g.stream));
  PXH_HIP(hipMalloc((void**)&g.errFlag, sizeof(int32_t)));
  PXH_HIP(hipMemsetAsync(g.errFlag, 0, sizeof(int32_t), g.stream));
  // kLanes words read at synchronize, then kLanes that checked renders write
  PXH_HIP(hipMalloc((void**)&g.laneErr, 2 * kLanes * sizeof(int32_t)));
  PXH_HIP(hipMemsetAsync(g.laneErr, 0, 2 * kLanes * sizeof(int32_t), g.stream));
  PXH_HIP(hipMalloc((void**)&g.touched, sizeof(unsigned long long)));
  PXH_HIP(hipMemsetAsync(g.touched, 0, sizeof(unsigned long long), g.stream));
  for (auto& e : g.ev) PXH_HIP(hipEventCreate(&e));
  if (set_fill_lds_limit(g.maxLds) != 0) return fail(kErrDevice, "hipFuncSetAttribute(fill)");
  if (set_blur_lds_limit(g.maxLds) != 0) return fail(kErrDevice, "hipFuncSetAttribute(blur)");
  if (scene_set_lds_limits(g.maxLds) != 0) return fail(kErrDevice, "hipFuncSetAttribute(scene)");
  g.inited = true;
  return 0;
}

#define PXH_INIT()          \
  do {                      \
    int r_ = init();        \
    if (r_ != 0) return r_; \
  } while (0)

// Stages `parts` contiguous host blobs into device memory; returns device
// pointers.  Waits for the previous consumer of the staging buffers.
int stage(const std::vector<std::pair<const void*, size_t>>& parts, std::vector<char*>* out) {
  size_t total = 0;
  std::vector<size_t> offs;
  for (auto& p : parts) {
    offs.push_back(total);
    total += (p.second + 255) & ~(size_t)255;
  }
  PXH_HIP(hipEventSynchronize(g.stageFree));
  if (total > g.hostCap) {
    if (g.hostStage) PXH_HIP(hipHostFree(g.hostStage));
    g.hostStage = nullptr;
    size_t c = total + total / 2 + (1 << 20);
    PXH_HIP(hipHostMalloc((void**)&g.hostStage, c, hipHostMallocDefault));
    g.hostCap = c;
  }
  if (int r = g.devStage.ensure(total)) return fail(r, "device staging allocation failed");
  for (size_t i = 0; i < parts.size(); i++)
    if (parts[i].second) memcpy(g.hostStage + offs[i], parts[i].first, parts[i].second);
  PXH_HIP(hipMemcpyAsync(g.devStage.p, g.hostStage, total, hipMemcpyHostToDevice, g.stream));
  out->clear();
  for (size_t i = 0; i < parts.size(); i++) out->push_back(g.devStage.p + offs[i]);
  return 0;
}

hipEvent_t pool_event() {
  if (!g.evPool.empty()) {
    hipEvent_t e = g.evPool.back();
    g.evPool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Brackets one kernel launch with events when profiling is on.
struct LaunchTimer {
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  explicit LaunchTimer(const char* kernel, hipStream_t s = nullptr) : name(kernel), st(s ? s : g.stream) {
    if (g.profiling) {
      a = pool_event();
      b = pool_event();
      (void)hipEventRecord(a, st);
    }
  }
  ~LaunchTimer() {
    if (a) {
      (void)hipEventRecord(b, st);
      g.evPairs.push_back({name, a, b});
    }
  }
};

int next_pow2(int v) {
  int p = 64;
  while (p < v) p <<= 1;
  return p;
}

// Largest number of band entries any row of the fill selects (entries with
// to.y > y and at.y < y + 1, the row program's selection), by a difference
// array over each band's rows.
int max_row_selection(const FillGeom& geo) {
  const FillInfo& f = geo.info;
  int best = 0;
  std::vector<int32_t> d;
  for (int32_t i = 0; i < f.nParts; i++) {
    const PartInfo& P = geo.parts[i];
    const int32_t top = f.startY + i * f.bandH;
    const int32_t bot = i == f.nParts - 1 ? f.pathHeight : top + f.bandH;
    if (bot <= top) continue;
    if (P.flags & 2) {
      best = std::max(best, 2);
      continue;
    }
    d.assign((size_t)(bot - top) + 1, 0);
    for (int32_t k = 0; k < P.count; k++) {
      const Entry& e = geo.entries[(size_t)P.entryBase + k];
      // rows y in [top, bot) with !(e.toy <= y || e.aty >= y + 1), the row
      // program's test (a NaN end selects every row, as there)
      int64_t y0 = top, y1 = bot;
      if (e.aty == e.aty && e.aty > (float)top - 2.0f) {
        const float a = std::min(e.aty, (float)bot + 2.0f);
        y0 = (int64_t)floor((double)a) - 1;
        while ((float)(y0 + 1) <= a) y0++;  // first y with a < y + 1
      }
      if (e.toy == e.toy && e.toy < (float)bot + 2.0f) {
        const float t = std::max(e.toy, (float)top - 2.0f);
        y1 = (int64_t)ceil((double)t) + 1;
        while (y1 > y0 && !(t > (float)(y1 - 1))) y1--;  // one past the last y with t > y
      }
      y0 = std::max<int64_t>(y0, top);
      y1 = std::min<int64_t>(y1, bot);
      if (y1 <= y0) continue;
      d[(size_t)(y0 - top)]++;
      d[(size_t)(y1 - top)]--;
    }
    int run = 0;
    for (int32_t y = 0; y < bot - top; y++) best = std::max(best, run += d[(size_t)y]);
  }
  return best;
}

// Launches the raster of one built fill onto img.
int raster_fill(pxh_image* img, FillGeom& geo) {
  FillInfo& f = geo.info;
  if (f.pathWidth == 0) return 0;
  const bool mask = f.mode == kMask;
  int32_t rowBase, nRows;
  if (mask) {
    rowBase = 0;
    nRows = img->h;
  } else {
    rowBase = f.startY;
    nRows = f.pathHeight - f.startY;
  }
  if (nRows <= 0) return 0;
  std::vector<char*> dp;
  int r = stage({{&f, sizeof(FillInfo)},
                 {geo.parts.data(), geo.parts.size() * sizeof(PartInfo)},
                 {geo.entries.data(), geo.entries.size() * sizeof(Entry)},
                 {geo.windings.data(), geo.windings.size()}},
                &dp);
  if (r) return r;
  // The row arrays hold one row's selection (the band entries that meet
  // [y, y + 1)), not the whole band: size them for the largest selection.
  const int maxSel = max_row_selection(geo);
  const int ww = 512;
  const int cap = next_pow2(maxSel > 2 ? maxSel : 2);
  const size_t fixedBytes = 8 * (size_t)ww + 16;
  const size_t capBytes = (size_t)cap * 24;
  FillArgs a;
  a.F = (const FillInfo*)dp[0];
  a.parts = (const PartInfo*)dp[1];
  a.E = (const Entry*)dp[2];
  a.EW = (const int8_t*)dp[3];
  a.canvas = img->dev;
  a.rowBase = rowBase;
  a.nRows = nRows;
  a.cap = cap;
  a.ww = ww;
  a.err = g.errFlag;
  a.touched = g.touched;
  const int32_t lo = mask ? 0 : f.xlo, hi = mask ? img->w : f.xhi;
  const int64_t nWin = hi > (lo & ~3) ? (hi - (lo & ~3) + ww - 1) / ww : 0;
  const int64_t nUnits = (int64_t)nRows * nWin;
  if (nUnits == 0) return 0;
  int grid = (int)std::min<int64_t>(nUnits, (int64_t)g.numCUs * 32);
  size_t lds = fixedBytes;
  if (fixedBytes + capBytes <= g.maxLds) {
    a.capInLds = 1;
    a.gscratch = nullptr;
    lds += capBytes;
  } else {
    a.capInLds = 0;
    size_t maxScratch = (size_t)1 << 30;
    if ((size_t)grid * capBytes > maxScratch) grid = (int)(maxScratch / capBytes);
    if (grid < 1) return fail(kErrCapacity, "a band holds too many segments");
    if (int e = g.gscratch.ensure((size_t)grid * capBytes)) return fail(e, "scratch allocation failed");
    a.gscratch = g.gscratch.p;
  }
  {
    LaunchTimer t("fill_rows_kernel");
    if (launch_fill_rows(a, grid, lds, img->fmt == PXH_FORMAT_A8, g.stream) != 0)
      return fail(kErrDevice, std::string("fill launch: ") + hipGetErrorString(hipGetLastError()));
  }
  PXH_HIP(hipEventRecord(g.stageFree, g.stream));
  return 0;
}

int check_image(const pxh_image* img) {
  if (!img || !img->dev) return fail(kErrArg, "invalid image handle");
  return 0;
}

int fill_shapes_impl(pxh_image* img, const Shapes& sh, uint32_t rgbx, int32_t rule, int32_t mode) {
  if (mode < 0 || mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  if (rule != kNonZero && rule != kEvenOdd) return fail(kErrArg, "invalid winding rule");
  FillGeom geo;
  std::string err;
  int32_t r = build_fill(sh, img->w, img->h, rgbx, mode, rule, &geo, &err);
  if (r != kOk) return fail(r, err);
  return raster_fill(img, geo);
}

const float kIdentity[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

// gaussianKernel (internal.nim:17-34): s and the exp() argument in float32,
// the 1/sqrt(2*PI*s^2) factor in float64 (2*PI is a float64 constant), taps
// summed in float64, stored to float32, normalised in float64.
// Writes the 2 * radius + 1 taps to out; f is caller-provided float storage
// of the same length.
template <class T>
void gaussian_lut_into(int radius, float* f, T* out) {
  int n = 2 * radius + 1;
  double total = 0.0;
  for (int step = -radius; step <= radius; step++) {
    float s = (float)radius / 2.2f;
    float s2 = s * s;
    float st = (float)step;
    double a = (1.0 / sqrt(6.283185307179586 * (double)s2)) *
               (double)expf((-1.0f * (st * st)) / (2.0f * s2));
    f[step + radius] = (float)a;
    total += a;
  }
  for (int i = 0; i < n; i++) {
    float v = (float)((double)f[i] / total);
    out[i] = (uint16_t)roundf(v * 255.0f * 256.0f);
  }
}

void gaussian_lut(int radius, std::vector<uint32_t>* out) {
  std::vector<float> f(2 * radius + 1);
  out->resize(2 * radius + 1);
  gaussian_lut_into(radius, f.data(), out->data());
}

int get_lut(int radius, uint32_t** dev) {
  std::lock_guard<std::mutex> lock(g.lutMu);
  auto it = g.luts.find(radius);
  if (it != g.luts.end()) {
    *dev = it->second;
    return 0;
  }
  // taps, then the weight table the blur kernels read (built on the device)
  std::vector<uint32_t> lut;
  gaussian_lut(radius, &lut);
  const size_t tabWords = 8 * (size_t)blur_table_dwords(radius);
  uint32_t* p = nullptr;
  PXH_HIP(hipMalloc((void**)&p, (lut.size() + tabWords) * 4));
  PXH_HIP(hipMemcpy(p, lut.data(), lut.size() * 4, hipMemcpyHostToDevice));
  if (launch_blur_table(p, radius, p + lut.size(), g.stream) != 0)
    return fail(kErrDevice, "blur table launch failed");
  PXH_HIP(hipStreamSynchronize(g.stream));
  g.luts[radius] = p + lut.size();
  *dev = p + lut.size();
  return 0;
}

int blur_impl(uint32_t* src_dst, uint32_t* out, int32_t w, int32_t h, int radius, uint32_t oob,
              bool shadowColor, uint32_t color, int lane = 0) {
  if (blur_x_lds(radius) > g.maxLds || blur_y_lds(radius) > g.maxLds)
    return fail(kErrCapacity, "blur radius too large for the LDS tiles");
  uint32_t* wtab = nullptr;
  if (int r = get_lut(radius, &wtab)) return r;
  DevBuf& tmp = lane == 0 ? g.tmp : g.laneTmp[lane];
  if (int e = tmp.ensure((size_t)w * h * 4)) return fail(e, "blur temporary allocation failed");
  const int nd = blur_table_dwords(radius);
  BlurArgs ax{src_dst, (uint32_t*)tmp.p, wtab, w, h, radius, oob, 0, 0, nd};
  BlurArgs ay{(uint32_t*)tmp.p, out, wtab, w, h, radius, oob, shadowColor ? 1 : 0, color, nd};
  if (launch_blur(ax, ay, g.lane[lane]) != 0) return fail(kErrDevice, "blur launch failed");
  return 0;
}


// ---- scene (batched form) ----

// A row whose selection exceeds the large-band launch's LDS capacity keeps
// its arrays in global row scratch, one slot of gcap entries per wave of that
// launch.  gcap starts at 8192 and grows with the largest band a validated
// render meets, up to kMaxSceneGcap and 8 GiB of slots in all.
const int kMaxSceneGcap = 1 << 18;

// Converts pixie commands to the device's canonical absolute form
// (commandsToShapes' command semantics, paths.nim:886-1057): relative,
// horizontal / vertical and smooth commands are resolved, every subpath
// starts with an explicit Move, and each command carries its subpath start.
int canonicalize(const float* c, int32_t n, std::vector<SceneCmd>* out,
                 std::vector<ScenePoly>* polys) {
  const int32_t polyBase0 = (int32_t)polys->size();  // first subpath of this call
  V2 start = v2(0, 0), at = v2(0, 0), prevCtrl = v2(0, 0), prevCtrl2 = v2(0, 0);
  int32_t prevKind = 1;  // Move
  bool open = false;     // a subpath (with its Move) has been started
  auto push = [&](int kind, std::initializer_list<float> p) {
    SceneCmd sc = {};
    sc.kind = (float)kind;
    int k = 0;
    for (float v : p) sc.p[k++] = v;
    sc.sx = start.x;
    sc.sy = start.y;
    sc.pad[0] = (int32_t)polys->size() - 1 - polyBase0;  // subpath index within the call
    out->push_back(sc);
    polys->back().numCmds++;
  };
  auto begin = [&](V2 s, bool prevNoClose) {
    if (open && prevNoClose) polys->back().noClose = 1;
    start = s;
    ScenePoly sp = {(int32_t)out->size(), 0, 0, 0};
    polys->push_back(sp);
    open = true;
    push(1, {s.x, s.y});
  };
  auto ensure_open = [&]() {
    if (!open) begin(at, false);
  };
  auto is_cubic = [](int k) { return k == 5 || k == 6 || k == 14 || k == 15; };
  auto is_quad = [](int k) { return k == 7 || k == 8 || k == 16 || k == 17; };
  int32_t i = 0;
  while (i < n) {
    int32_t kind = (int32_t)c[i];
    i++;
    static const int pc[19] = {0, 2, 2, 1, 1, 6, 4, 4, 2, 7, 2, 2, 1, 1, 6, 4, 4, 2, 7};
    if (kind < 0 || kind > 18 || i + pc[kind] > n) return fail(kErrArg, "Invalid path command");
    const float* a = c + i;
    switch (kind) {
      case 1:  // Move
        at = v2(a[0], a[1]);
        begin(at, false);
        break;
      case 10:  // RMove: ends the subpath without closing it (paths.nim:968-974)
        at = v2(at.x + a[0], at.y + a[1]);
        begin(at, true);
        break;
      case 2: case 3: case 4: case 11: case 12: case 13: {
        V2 to;
        if (kind == 2) to = v2(a[0], a[1]);
        else if (kind == 3) to = v2(a[0], at.y);
        else if (kind == 4) to = v2(at.x, a[0]);
        else if (kind == 11) to = v2(at.x + a[0], at.y + a[1]);
        else if (kind == 12) to = v2(at.x + a[0], at.y);
        else to = v2(at.x, at.y + a[0]);
        ensure_open();
        push(2, {to.x, to.y});
        at = to;
      } break;
      case 5: case 6: case 14: case 15: {
        V2 c1, c2, to;
        if (kind == 5) {
          c1 = v2(a[0], a[1]); c2 = v2(a[2], a[3]); to = v2(a[4], a[5]);
        } else if (kind == 14) {
          c1 = v2(at.x + a[0], at.y + a[1]); c2 = v2(at.x + a[2], at.y + a[3]);
          to = v2(at.x + a[4], at.y + a[5]);
        } else if (kind == 6) {
          c2 = v2(a[0], a[1]); to = v2(a[2], a[3]);
          c1 = is_cubic(prevKind) ? at * 2.0f - prevCtrl2 : at;
        } else {
          c2 = v2(at.x + a[0], at.y + a[1]); to = v2(at.x + a[2], at.y + a[3]);
          c1 = is_cubic(prevKind) ? at * 2.0f - prevCtrl2 : at;
        }
        ensure_open();
        push(5, {c1.x, c1.y, c2.x, c2.y, to.x, to.y});
        at = to;
        prevCtrl2 = c2;
      } break;
      case 7: case 8: case 16: case 17: {
        V2 cc, to;
        if (kind == 7) {
          cc = v2(a[0], a[1]); to = v2(a[2], a[3]);
        } else if (kind == 16) {
          cc = v2(at.x + a[0], at.y + a[1]); to = v2(at.x + a[2], at.y + a[3]);
        } else if (kind == 8) {
          to = v2(a[0], a[1]);
          cc = is_quad(prevKind) ? at * 2.0f - prevCtrl : at;
        } else {
          to = v2(at.x + a[0], at.y + a[1]);
          cc = is_quad(prevKind) ? at * 2.0f - prevCtrl : at;
        }
        ensure_open();
        push(7, {cc.x, cc.y, to.x, to.y});
        at = to;
        prevCtrl = cc;
      } break;
      case 0:  // Close
        ensure_open();
        push(0, {});
        at = start;
        open = false;
        break;
      default:
        return fail(kErrArg, "elliptical arc commands are only supported by the per-call form");
    }
    i += pc[kind];
    prevKind = kind;
  }
  return 0;
}

// Records flattened shapes as Move + Line subpaths that the device turns back
// into the same points: consecutive shape points differ, so every Line emits
// its end point, and noClose keeps the device from adding a closing point
// (a fill's shapes already hold it).
void canonicalize_shapes(const Shapes& sh, std::vector<SceneCmd>* out, std::vector<ScenePoly>* polys) {
  const int32_t polyBase0 = (int32_t)polys->size();
  for (int32_t k = 0; k < sh.count(); k++) {
    const int32_t a = sh.ofs[k], b = sh.ofs[k + 1];
    if (b <= a) continue;
    ScenePoly sp = {(int32_t)out->size(), 0, 1, 0};
    polys->push_back(sp);
    for (int32_t i = a; i < b; i++) {
      SceneCmd c = {};
      c.kind = i == a ? 1.0f : 2.0f;
      c.p[0] = sh.pts[i].x;
      c.p[1] = sh.pts[i].y;
      c.sx = sh.pts[a].x;
      c.sy = sh.pts[a].y;
      c.pad[0] = (int32_t)polys->size() - 1 - polyBase0;
      out->push_back(c);
      polys->back().numCmds++;
    }
  }
}

bool has_arc(const float* c, int32_t n) {
  static const int pc[19] = {0, 2, 2, 1, 1, 6, 4, 4, 2, 7, 2, 2, 1, 1, 6, 4, 4, 2, 7};
  for (int32_t i = 0; i < n;) {
    const int32_t k = (int32_t)c[i];
    if (k < 0 || k > 18) return false;  // canonicalize reports the bad command
    if (k == 9 || k == 18) return true;
    i += 1 + pc[k];
  }
  return false;
}

void item_params(pxh_scene* sc, SceneItem* itp, const float* mat9, uint32_t rgbx, int32_t rule, int32_t mode,
                 bool stroke, float width, int32_t cap, int32_t join, float miterLimit, const float* dashes,
                 int32_t ndashes);

int scene_record(pxh_scene* sc, const float* cmds, int32_t n, const float* mat9, uint32_t rgbx,
                 int32_t rule, int32_t mode, bool stroke, float width, int32_t cap, int32_t join,
                 float miterLimit, const float* dashes = nullptr, int32_t ndashes = 0) {
  if (!sc) return fail(kErrArg, "invalid scene handle");
  if (mode < 0 || mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  if (rule != kNonZero && rule != kEvenOdd) return fail(kErrArg, "invalid winding rule");
  if (ndashes < 0 || (ndashes > 0 && !dashes)) return fail(kErrArg, "invalid dash array");
  if (!mat9) mat9 = kIdentity;
  if (stroke && width <= 0.0f) return 0;  // strokeShapes yields nothing (paths.nim:1931-1932)
  SceneItem it = {};
  it.cmdBase = (int32_t)sc->cmds.size();
  it.polyBase = (int32_t)sc->polys.size();
  size_t polys0 = sc->polys.size(), cmds0 = sc->cmds.size();
  if (has_arc(cmds, n)) {
    // Elliptical arcs are flattened here, on the host, at the call's
    // tolerance: their flattening evaluates sinf / cosf / acosf, whose device
    // results can differ from the host libm's in the last bit, and the points
    // must be the ones pixie's addArc (paths.nim:764-884) emits.  The device
    // then meets them as line subpaths.
    Shapes sh;
    std::string err;
    int32_t r = commands_to_shapes(cmds, n, !stroke, pixel_scale(mat9), &sh, &err);
    if (r != kOk) return fail(r, err);
    canonicalize_shapes(sh, &sc->cmds, &sc->polys);
  } else if (int r = canonicalize(cmds, n, &sc->cmds, &sc->polys)) {
    sc->cmds.resize(cmds0);
    sc->polys.resize(polys0);
    return r;
  }
  it.cmdCount = (int32_t)sc->cmds.size() - it.cmdBase;
  it.polyCount = (int32_t)sc->polys.size() - it.polyBase;
  item_params(sc, &it, mat9, rgbx, rule, mode, stroke, width, cap, join, miterLimit, dashes, ndashes);
  sc->items.push_back(it);
  sc->uploaded = false;
  sc->gen++;
  sc->validW = sc->validH = -1;
  return 0;
}

// The paint, stroke and transform parameters of a recorded call, and its
// dash pattern (an odd one is repeated to even length, paths.nim:2025-2027).
void item_params(pxh_scene* sc, SceneItem* itp, const float* mat9, uint32_t rgbx, int32_t rule, int32_t mode,
                 bool stroke, float width, int32_t cap, int32_t join, float miterLimit, const float* dashes,
                 int32_t ndashes) {
  SceneItem& it = *itp;
  it.dashBase = (int32_t)sc->dashes.size();
  it.dashCount = 0;
  it.dashBad = 0;
  if (stroke && ndashes > 0) {
    for (int32_t k = 0; k < ndashes; k++) {
      sc->dashes.push_back(dashes[k]);
      if (dashes[k] <= 0.0f) it.dashBad = 1;
    }
    if (ndashes % 2 != 0)
      for (int32_t k = 0; k < ndashes; k++) sc->dashes.push_back(dashes[k]);
    it.dashCount = (int32_t)sc->dashes.size() - it.dashBase;
  }
  it.stroke = stroke ? 1 : 0;
  it.rgbx = rgbx;
  it.mode = mode;
  it.rule = rule;
  it.cap = cap;
  it.join = join;
  float ps = pixel_scale(mat9);
  it.halfStroke = width / 2.0f;
  it.miterAngleLimit = asinf(1.0f / miterLimit) * 2.0f;
  it.minArea = 0.2f / ps;
  it.errSq = powf(0.2f / ps, 2.0f);
  memcpy(it.mat, mat9, sizeof(it.mat));
  it.identity = mat_is_identity(mat9) ? 1 : 0;
  it.strokeNone = (stroke && width <= 0.0f) ? 1 : 0;
}

// A text call: its commands and subpaths are the glyph instances', placed
// on the device at upload (scene_launch_expand) in the text region.
int text_record_impl(pxh_scene* sc, const std::vector<TextInst>& inst, const float* mat9, uint32_t rgbx,
                     int32_t rule, int32_t mode, bool stroke, float width, int32_t cap, int32_t join,
                     float miterLimit, const float* dashes, int32_t ndashes) {
  if (!sc) return fail(kErrArg, "invalid scene handle");
  if (mode < 0 || mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  if (rule != kNonZero && rule != kEvenOdd) return fail(kErrArg, "invalid winding rule");
  if (ndashes < 0 || (ndashes > 0 && !dashes)) return fail(kErrArg, "invalid dash array");
  if (!mat9) mat9 = kIdentity;
  if (stroke && width <= 0.0f) return 0;  // strokeShapes yields nothing (paths.nim:1931-1932)
  const int32_t relC = sc->textCmds, relP = sc->textPolys;
  int32_t nC = 0, nP = 0;
  for (const TextInst& t : inst) {
    GlyphInst gi = {};
    if (t.ref < 0) {
      gi.srcCmd = gi.srcPoly = -1;
      gi.nCmds = 5;
      gi.nPolys = 1;
    } else {
      if (t.ref >= (int32_t)g.glyphRecs.size()) return fail(kErrArg, "invalid glyph reference");
      const Ctx::GlyphRec& r = g.glyphRecs[t.ref];
      gi.srcCmd = r.cmd;
      gi.nCmds = r.nCmds;
      gi.srcPoly = r.poly;
      gi.nPolys = r.nPolys;
    }
    if (gi.nCmds == 0) continue;  // an empty glyph adds nothing
    gi.dstCmd = relC + nC;
    gi.dstPoly = relP + nP;
    gi.polyIdx = nP;
    gi.identity = t.identity;
    memcpy(gi.m, t.m, sizeof(gi.m));
    sc->glyphInst.push_back(gi);
    nC += gi.nCmds;
    nP += gi.nPolys;
  }
  SceneItem it = {};
  it.cmdCount = nC;
  it.polyCount = nP;
  item_params(sc, &it, mat9, rgbx, rule, mode, stroke, width, cap, join, miterLimit, dashes, ndashes);
  sc->textItems.push_back({(int32_t)sc->items.size(), relC, relP});
  sc->textCmds += nC;
  sc->textPolys += nP;
  sc->items.push_back(it);
  sc->uploaded = false;
  sc->gen++;
  sc->validW = sc->validH = -1;
  return 0;
}

void scene_drop_graph(pxh_scene* sc) {
  if (sc->exec) (void)hipGraphExecDestroy(sc->exec);
  sc->exec = nullptr;
}

void scene_free_device(pxh_scene* sc) {
  scene_drop_graph(sc);
  if (sc->dItems) (void)hipFree(sc->dItems);
  if (sc->dCmds) (void)hipFree(sc->dCmds);
  if (sc->dPolys) (void)hipFree(sc->dPolys);
  if (sc->dOrder) (void)hipFree(sc->dOrder);
  if (sc->dDashes) (void)hipFree(sc->dDashes);
  sc->dDashes = nullptr;
  sc->dDashesCap = 0;
  if (sc->dInst) (void)hipFree(sc->dInst);
  sc->dInst = nullptr;
  sc->dInstCap = 0;
  if (sc->dMem) (void)hipFree(sc->dMem);
  sc->dItems = nullptr;
  sc->dCmds = nullptr;
  sc->dPolys = nullptr;
  sc->dOrder = nullptr;
  sc->dMem = nullptr;
  sc->dItemsCap = sc->dCmdsCap = sc->dPolysCap = sc->dOrderCap = 0;
  sc->dMemBytes = 0;
}

// (Re)allocates the per-render pools of a scene in one device block.
int scene_alloc(pxh_scene* sc) {
  sc->gen++;
  const size_t nI = sc->items.size(), nC = sc->cmds.size() + (size_t)sc->textCmds;
  const size_t nP = sc->polys.size() + (size_t)sc->textPolys;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t oGeo = carve(nI * sizeof(ItemGeom)), oFill = carve(nI * sizeof(FillInfo));
  size_t oCC = carve(nC * 4), oCO = carve(nC * 4), oPP = carve(nP * 4), oPN = carve(nP * 4);
  size_t oPts = carve((size_t)sc->ptCap * 8), oPtP = carve((size_t)sc->ptCap * 4);
  size_t oPtS = carve((size_t)sc->ptCap * 4);
  const size_t nChunk = (size_t)sc->ptCap / kBigChunk + (size_t)sc->nBig + 1;
  size_t oBCS = carve(((size_t)sc->nBig + 1) * 4), oBGS = carve(((size_t)sc->nBig + 1) * 4);
  size_t oBPS = carve(((size_t)sc->nBig + 1) * 4);
  size_t oBPC = carve((nP / kBigPolyChunk + (size_t)sc->nBig + 1) * 16);
  size_t oBCC = carve(nChunk * 8), oBCB = carve(nChunk * 16);
  size_t oSeg = carve((size_t)sc->segCap * 16), oSegW = carve((size_t)sc->segCap);
  size_t oSegB = carve((size_t)sc->segCap * 8), oSegC = carve(((size_t)sc->segCap / 64 + 1) * 8);
  size_t oParts = carve((size_t)sc->bandCap * sizeof(PartInfo)), oBI = carve((size_t)sc->bandCap * 4);
  size_t oEnt = carve((size_t)sc->entryCap * sizeof(Entry)), oEW = carve((size_t)sc->entryCap);
  size_t oTask = carve((size_t)sc->taskCap * sizeof(RowRuns)), oRun = carve((size_t)sc->runCap * 8);
  size_t oTaskItem = carve((size_t)sc->taskCap * 4), oTaskOrd = carve((size_t)sc->taskCap * 4);
  size_t oTaskCls = carve((size_t)sc->taskCap);
  size_t oRows = carve(nI * sizeof(int4)), oBinC = carve((size_t)sc->binSlots * 4);
  size_t oPaint = carve(nI * sizeof(int4)), oSpan = carve(nI * sizeof(int4));
  size_t oBinO = carve(((size_t)sc->binSlots + 1) * 4), oBinL = carve((size_t)sc->binCap * 4);
  size_t oQ = carve(kQueueSets * kQueues * kQueueStride * 4);
  size_t oCnt = carve(kCounterInts * 4);
  if (off > sc->dMemBytes) {
    PXH_HIP(hipStreamSynchronize(g.lane[sc->lane]));
    if (sc->dMem) PXH_HIP(hipFree(sc->dMem));
    sc->dMem = nullptr;
    if (hipMalloc((void**)&sc->dMem, off) != hipSuccess) return fail(kErrNoMem, "scene pool allocation failed");
    sc->dMemBytes = off;
  }
  char* b = sc->dMem;
  ScenePools& p = sc->pools;
  p.items = sc->dItems;
  p.cmds = sc->dCmds;
  p.polys = sc->dPolys;
  p.itemOrder = sc->dOrder;
  p.nBig = sc->nBig;
  p.bigCmdStart = sc->dOrder + sc->items.size();
  p.bigCmds = sc->bigCmds;
  p.dashes = sc->dDashes;
  p.nItems = (int32_t)nI;
  p.geo = (ItemGeom*)(b + oGeo);
  p.fills = (FillInfo*)(b + oFill);
  p.cmdCount = (int32_t*)(b + oCC);
  p.cmdOff = (int32_t*)(b + oCO);
  p.polyPt = (int32_t*)(b + oPP);
  p.polyN = (int32_t*)(b + oPN);
  p.pts = (float2*)(b + oPts);
  p.ptPoly = (int32_t*)(b + oPtP);
  p.ptSeg = (int32_t*)(b + oPtS);
  p.ptCap = sc->ptCap;
  p.bigChunkStart = (int32_t*)(b + oBCS);
  p.bigGroupStart = (int32_t*)(b + oBGS);
  p.bigPolyStart = (int32_t*)(b + oBPS);
  p.bigPolyChunk = (int4*)(b + oBPC);
  p.bigChunkCnt = (int2*)(b + oBCC);
  p.bigChunkBd = (float4*)(b + oBCB);
  p.segs = (float4*)(b + oSeg);
  p.segW = (int8_t*)(b + oSegW);
  p.segBand = (int2*)(b + oSegB);
  p.segChunk = (int2*)(b + oSegC);
  p.segCap = sc->segCap;
  p.parts = (PartInfo*)(b + oParts);
  p.bandItem = (int32_t*)(b + oBI);
  p.bandCap = sc->bandCap;
  p.entries = (Entry*)(b + oEnt);
  p.entW = (int8_t*)(b + oEW);
  p.entryCap = sc->entryCap;
  p.tasks = (RowRuns*)(b + oTask);
  p.taskItem = (int32_t*)(b + oTaskItem);
  p.taskOrder = (int32_t*)(b + oTaskOrd);
  p.taskClass = (int8_t*)(b + oTaskCls);
  p.taskCap = sc->taskCap;
  p.runs = (int2*)(b + oRun);
  p.runCap = sc->runCap;
  p.itemRows = (int4*)(b + oRows);
  p.itemPaint = (int4*)(b + oPaint);
  p.itemSpan = (int4*)(b + oSpan);
  p.binCount = (int32_t*)(b + oBinC);
  p.binOff = (int32_t*)(b + oBinO);
  p.binList = (int32_t*)(b + oBinL);
  p.binCap = sc->binCap;
  p.binSlots = sc->binSlots;
  p.queues = (int32_t*)(b + oQ);
  p.counters = (int32_t*)(b + oCnt);
  p.gcap = sc->gcap;
  return 0;
}

// Grow-only device buffer of a scene.
template <class T>
int scene_buf(T** p, size_t* cap, size_t n) {
  if (n <= *cap && *p) return 0;
  if (*p) PXH_HIP(hipFree(*p));
  *p = nullptr;
  size_t c = n + n / 2 + 64;
  if (hipMalloc((void**)p, c * sizeof(T)) != hipSuccess) return fail(kErrNoMem, "scene upload failed");
  *cap = c;
  return 0;
}

// The glyph arena's device copy, grown (whole) or extended by its new tail.
int glyph_arena_upload(hipStream_t st) {
  const size_t nc = g.glyphCmds.size(), np = g.glyphPolys.size();
  if (nc == g.upGlyphCmds && np == g.upGlyphPolys) return 0;
  if (nc > g.dGlyphCmdsCap || np > g.dGlyphPolysCap) {
    PXH_HIP(hipDeviceSynchronize());  // other lanes' expansions may read the old copy
    if (g.dGlyphCmds) PXH_HIP(hipFree(g.dGlyphCmds));
    if (g.dGlyphPolys) PXH_HIP(hipFree(g.dGlyphPolys));
    g.dGlyphCmds = nullptr;
    g.dGlyphPolys = nullptr;
    g.dGlyphCmdsCap = g.dGlyphPolysCap = 0;
    g.upGlyphCmds = g.upGlyphPolys = 0;
    if (int r = scene_buf(&g.dGlyphCmds, &g.dGlyphCmdsCap, nc * 2 + 4096)) return r;
    if (int r = scene_buf(&g.dGlyphPolys, &g.dGlyphPolysCap, np * 2 + 1024)) return r;
  }
  PXH_HIP(hipMemcpyAsync(g.dGlyphCmds + g.upGlyphCmds, g.glyphCmds.data() + g.upGlyphCmds,
                         (nc - g.upGlyphCmds) * sizeof(SceneCmd), hipMemcpyHostToDevice, st));
  if (np > g.upGlyphPolys)
    PXH_HIP(hipMemcpyAsync(g.dGlyphPolys + g.upGlyphPolys, g.glyphPolys.data() + g.upGlyphPolys,
                           (np - g.upGlyphPolys) * sizeof(ScenePoly), hipMemcpyHostToDevice, st));
  g.upGlyphCmds = nc;
  g.upGlyphPolys = np;
  PXH_HIP(hipStreamSynchronize(st));  // the host arena may grow (and move) after this
  return 0;
}

int scene_upload(pxh_scene* sc, int lane) {
  if (sc->uploaded) return 0;
  hipStream_t st = g.lane[lane];
  PXH_HIP(hipStreamSynchronize(g.lane[sc->lane]));  // a previous render may still read the buffers
  // text calls take the region after the host-recorded commands and subpaths
  const int32_t regC = (int32_t)sc->cmds.size(), regP = (int32_t)sc->polys.size();
  for (const auto& t : sc->textItems) {
    sc->items[t[0]].cmdBase = regC + t[1];
    sc->items[t[0]].polyBase = regP + t[2];
  }
  const size_t allC = (size_t)regC + sc->textCmds, allP = (size_t)regP + sc->textPolys;
  if (int r = scene_buf(&sc->dItems, &sc->dItemsCap, sc->items.size())) return r;
  if (int r = scene_buf(&sc->dCmds, &sc->dCmdsCap, allC)) return r;
  if (int r = scene_buf(&sc->dPolys, &sc->dPolysCap, allP)) return r;
  PXH_HIP(hipMemcpyAsync(sc->dItems, sc->items.data(), sc->items.size() * sizeof(SceneItem),
                         hipMemcpyHostToDevice, st));
  if (regC > 0)
    PXH_HIP(hipMemcpyAsync(sc->dCmds, sc->cmds.data(), (size_t)regC * sizeof(SceneCmd), hipMemcpyHostToDevice,
                           st));
  if (regP > 0)
    PXH_HIP(hipMemcpyAsync(sc->dPolys, sc->polys.data(), (size_t)regP * sizeof(ScenePoly),
                           hipMemcpyHostToDevice, st));
  if (!sc->glyphInst.empty()) {
    if (int r = glyph_arena_upload(st)) return r;
    // region-relative destinations -> absolute, in a copy that stays
    // untouched until the next upload (which first waits for the lane)
    const size_t n = sc->glyphInst.size();
    sc->glyphAbs = sc->glyphInst;
    for (GlyphInst& gi : sc->glyphAbs) {
      gi.dstCmd += regC;
      gi.dstPoly += regP;
    }
    if (int r = scene_buf(&sc->dInst, &sc->dInstCap, n)) return r;
    PXH_HIP(hipMemcpyAsync(sc->dInst, sc->glyphAbs.data(), n * sizeof(GlyphInst), hipMemcpyHostToDevice, st));
    if (scene_launch_expand(sc->dInst, (int32_t)n, g.dGlyphCmds, g.dGlyphPolys, sc->dCmds, sc->dPolys, st))
      return fail(kErrDevice, "glyph expansion launch failed");
  }
  if (int r = scene_buf(&sc->dDashes, &sc->dDashesCap, sc->dashes.size())) return r;
  if (!sc->dashes.empty())
    PXH_HIP(hipMemcpyAsync(sc->dDashes, sc->dashes.data(), sc->dashes.size() * sizeof(float),
                           hipMemcpyHostToDevice, st));
  // Geometry workgroups start longest call first, so that a long stroke does
  // not start late and hold the kernel's end: estimated cost = commands,
  // times the work of a stroke point (heavier with round joins or caps).
  {
    const size_t n = sc->items.size();
    std::vector<int64_t> cost(n);
    for (size_t i = 0; i < n; i++) {
      const SceneItem& it = sc->items[i];
      int64_t c = it.cmdCount;
      if (it.stroke && !it.strokeNone) c *= (it.join == 1 || it.cap == 1) ? 16 : 4;
      cost[i] = c;
    }
    // calls of more than 256 commands first: their geometry runs in
    // 1024-thread workgroups of a launch of its own
    auto big = [&](int32_t i) { return sc->items[i].cmdCount > 256; };
    sc->order.resize(n);
    for (size_t i = 0; i < n; i++) sc->order[i] = (int32_t)i;
    std::stable_sort(sc->order.begin(), sc->order.end(), [&](int32_t a, int32_t b) {
      if (big(a) != big(b)) return big(a);
      return cost[a] > cost[b];
    });
    sc->nBig = 0;
    for (size_t i = 0; i < n; i++) sc->nBig += big((int32_t)i) ? 1 : 0;
    // after the order: the big calls' command prefix
    sc->bigCmds = 0;
    for (int32_t k = 0; k < sc->nBig; k++) {
      sc->order.push_back(sc->bigCmds);
      sc->bigCmds += sc->items[sc->order[k]].cmdCount;
    }
    sc->order.push_back(sc->bigCmds);
    const size_t m = sc->order.size();
    if (int r = scene_buf(&sc->dOrder, &sc->dOrderCap, m)) return r;
    PXH_HIP(hipMemcpyAsync(sc->dOrder, sc->order.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, st));
  }
  sc->lane = lane;
  // first estimates of the pools; a render that overflows grows them.  The
  // sizes a scene has grown to are kept across clear() and re-recording, so
  // a scene reused for similar requests validates once, not per request.
  size_t nC = allC;
  auto keep = [](int32_t prev, int64_t est) { return (int32_t)std::max<int64_t>(prev, est); };
  sc->ptCap = keep(sc->ptCap, std::min<size_t>(64 * nC + 4096, 1u << 28));
  sc->segCap = keep(sc->segCap, (int64_t)sc->ptCap * 2);
  sc->bandCap = keep(sc->bandCap, std::min<size_t>(64 * sc->items.size() + 4096, 1u << 26));
  sc->entryCap = keep(sc->entryCap, (int64_t)sc->segCap * 2);
  sc->taskCap = keep(sc->taskCap, std::min<size_t>(256 * sc->items.size() + 4096, 1u << 28));
  sc->runCap = keep(sc->runCap, std::min<int64_t>(
      (int64_t)sc->taskCap * 8 + (int64_t)scene_row_waves(g.numCUs) * kRunChunk,
      (int64_t)1 << 30));
  sc->binCap = keep(sc->binCap, std::min<size_t>(8 * sc->items.size() + 4096, 1u << 28));
  sc->uploaded = true;
  return scene_alloc(sc);
}

// Queues every kernel of one render.  `checked` renders are those whose
// counters the host reads itself (pool validation, autotune timing): their
// compose kernel reports an early stop into a word nobody reads, so that
// only unchecked renders (replays) raise the lane's error.
int scene_launch(pxh_scene* sc, pxh_image* img, int lane, bool checked = false) {
  ScenePools& p = sc->pools;
  hipStream_t st = g.lane[lane];
  sc->lane = lane;
  p.W = img->w;
  p.H = img->h;
  p.items = sc->dItems;
  p.cmds = sc->dCmds;
  p.polys = sc->dPolys;
  p.itemOrder = sc->dOrder;
  p.nBig = sc->nBig;
  p.bigCmdStart = sc->dOrder + sc->items.size();
  p.bigCmds = sc->bigCmds;
  p.dashes = sc->dDashes;
  const RowShape shape = kRowShapes[sc->rowShape];
  p.ldsCap = shape.ldsCap;
  // tasks of bands of more than ldsCap entries go to the large-band launch;
  // a single-launch shape has none (class 32 is never reached)
  p.bigClass = shape.mode ? 32 - __builtin_clz((unsigned)shape.ldsCap) : kCostClasses;
  DevBuf& scr = g.laneScratch[lane];
  if (int e = scr.ensure(scene_gscratch_bytes(scene_scratch_slots(shape, g.numCUs), sc->gcap)))
    return fail(e, "scene row scratch allocation failed");
  p.gscratch = scr.p;
  int r = scene_launch_clear(p, st);
  if (!r) {
    LaunchTimer t("scene_geometry", st);
    r = scene_launch_geometry(p, st);
  }
  if (!r) {
    LaunchTimer t("scene_bands", st);
    r = scene_launch_bands(p, g.numCUs * 64, st);
  }
  if (!r) {
    LaunchTimer t("scene_bins", st);
    r = scene_launch_bins(p, g.numCUs * 4, st);
  }
  if (!r) {
    LaunchTimer t("scene_rows_kernel", st);
    r = scene_launch_rows(p, shape, g.numCUs, st, g.laneAux[lane], g.laneFork[lane], g.laneJoin[lane]);
  }
  if (!r) {
    LaunchTimer t("scene_compose_kernel", st);
    r = scene_launch_compose(p, img->dev, img->fmt == PXH_FORMAT_A8, g.touched,
                             g.laneErr + (checked ? kLanes : 0) + lane, scene_queue_grid(g.numCUs * 24),
                             st);
  }
  if (r) return fail(kErrDevice, std::string("scene launch: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

// A render of a validated scene: replays the captured graph of this render
// (capturing it first when anything it depends on changed).  Profiled
// renders run eagerly so that every launch can be timed.
int scene_replay(pxh_scene* sc, pxh_image* img, int lane) {
  if (g.profiling || sc->graphsOff) return scene_launch(sc, img, lane);
  hipStream_t st = g.lane[lane];
  const RowShape shape = kRowShapes[sc->rowShape];
  DevBuf& scr = g.laneScratch[lane];
  if (int e = scr.ensure(scene_gscratch_bytes(scene_scratch_slots(shape, g.numCUs), sc->gcap)))
    return fail(e, "scene row scratch allocation failed");
  pxh_scene::GraphKey key{img->dev, img->w, img->h, img->fmt, lane, sc->rowShape, sc->dMem, scr.p, sc->gen};
  if (!sc->exec || !(key == sc->key)) {
    scene_drop_graph(sc);
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      sc->graphsOff = true;
      return scene_launch(sc, img, lane);
    }
    int r = scene_launch(sc, img, lane);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (r != 0 || e != hipSuccess || !graph ||
        hipGraphInstantiate(&sc->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      sc->exec = nullptr;
      sc->graphsOff = true;
      (void)hipGetLastError();
      return scene_launch(sc, img, lane);
    }
    (void)hipGraphDestroy(graph);
    sc->key = key;
  }
  PXH_HIP(hipGraphLaunch(sc->exec, st));
  return 0;
}

// Reads and clears the scene error word of a synchronized lane.  A render
// replayed from a graph is not checked when it is issued; if one ran out of
// a pool, its compose kernel left the canvas untouched and set this word.
// The scenes of the lane are then re-validated (and their pools grown) on
// their next render.
int lane_scene_errors(int lane) {
  int32_t e = 0;
  PXH_HIP(hipMemcpy(&e, g.laneErr + lane, sizeof(e), hipMemcpyDeviceToHost));
  if (!e) return 0;
  PXH_HIP(hipMemset(g.laneErr + lane, 0, sizeof(int32_t)));
  {
    std::lock_guard<std::mutex> lk(g.scenesMu);
    for (pxh_scene* sc : g.scenes)
      if (sc->lane == lane) sc->validW = sc->validH = -1;
  }
  return fail(kErrCapacity, "a scene render on lane " + std::to_string(lane) +
                                " ran out of a pool (error bits " + std::to_string(e) +
                                ") and left its canvas untouched; render it again");
}

}  // namespace

int set_last_error(int code, const std::string& msg) { return fail(code, msg); }

}  // namespace pxh

using namespace pxh;

extern "C" {

int pxh_init(void) { return init(); }

const char* pxh_last_error(void) { return g_err.c_str(); }

const char* pxh_version(void) { return "pixie_hip 0.1 (gfx950)"; }

int pxh_synchronize(void) {
  PXH_INIT();
  PXH_HIP(hipStreamSynchronize(g.stream));
  if (int r = lane_scene_errors(0)) return r;
  int32_t e = 0;
  PXH_HIP(hipMemcpy(&e, g.errFlag, sizeof(e), hipMemcpyDeviceToHost));
  if (e) {
    PXH_HIP(hipMemset(g.errFlag, 0, sizeof(int32_t)));
    return fail(kErrCapacity, "a band held more segments than the row scratch capacity");
  }
  return 0;
}

static pxh_image* image_new(int32_t w, int32_t h, int32_t fmt) {
  if (init() != 0) return nullptr;
  if (w <= 0 || h <= 0) {
    fail(kErrArg, "Image width and height must be > 0");
    return nullptr;
  }
  pxh_image* img = new pxh_image;
  img->w = w;
  img->h = h;
  img->fmt = fmt;
  img->bytes = (size_t)w * h * (fmt == PXH_FORMAT_A8 ? 1 : 4);
  img->dev = nullptr;
  if (hipMalloc(&img->dev, img->bytes) != hipSuccess) {
    delete img;
    fail(kErrNoMem, "device allocation failed");
    return nullptr;
  }
  if (hipMemsetAsync(img->dev, 0, img->bytes, g.stream) != hipSuccess ||
      hipStreamSynchronize(g.stream) != hipSuccess) {
    (void)hipFree(img->dev);
    delete img;
    fail(kErrDevice, "device memset failed");
    return nullptr;
  }
  return img;
}

pxh_image* pxh_image_new(int32_t w, int32_t h) { return image_new(w, h, PXH_FORMAT_RGBA8); }
pxh_image* pxh_image_new_a8(int32_t w, int32_t h) { return image_new(w, h, PXH_FORMAT_A8); }

void pxh_image_free(pxh_image* img) {
  if (!img) return;
  if (img->owned) {
    (void)hipStreamSynchronize(g.stream);
    if (img->dev) (void)hipFree(img->dev);
  }
  delete img;
}

int32_t pxh_image_width(const pxh_image* img) { return img ? img->w : 0; }
int32_t pxh_image_height(const pxh_image* img) { return img ? img->h : 0; }
int32_t pxh_image_format(const pxh_image* img) { return img ? img->fmt : -1; }
void* pxh_image_device_ptr(pxh_image* img) { return img ? img->dev : nullptr; }

int pxh_image_upload(pxh_image* img, const void* data) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  PXH_HIP(hipMemcpyAsync(img->dev, data, img->bytes, hipMemcpyHostToDevice, g.stream));
  PXH_HIP(hipStreamSynchronize(g.stream));
  return 0;
}

int pxh_image_download(const pxh_image* img, void* data) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  PXH_HIP(hipMemcpyAsync(data, img->dev, img->bytes, hipMemcpyDeviceToHost, g.stream));
  return pxh_synchronize();
}

int pxh_image_copy(pxh_image* dst, const pxh_image* src) {
  PXH_INIT();
  if (int r = check_image(dst)) return r;
  if (int r = check_image(src)) return r;
  if (dst->bytes != src->bytes || dst->fmt != src->fmt) return fail(kErrArg, "image size mismatch");
  PXH_HIP(hipMemcpyAsync(dst->dev, src->dev, src->bytes, hipMemcpyDeviceToDevice, g.stream));
  return 0;
}

int pxh_image_fill(pxh_image* img, uint32_t rgbx) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (img->fmt == PXH_FORMAT_A8)
    PXH_HIP(hipMemsetAsync(img->dev, (int)(rgbx >> 24), img->bytes, g.stream));
  else
    PXH_HIP(hipMemsetD32Async((hipDeviceptr_t)img->dev, (int)rgbx, img->bytes / 4, g.stream));
  return 0;
}

int pxh_parse_path(const char* path, float** cmds, int32_t* ncmds) {
  std::vector<float> out;
  std::string err;
  *cmds = nullptr;
  *ncmds = 0;
  int32_t r = parse_path(path, &out, &err);
  if (r != kOk) return fail(r, err);
  if (!out.empty()) {
    *cmds = (float*)malloc(out.size() * sizeof(float));
    if (!*cmds) return fail(kErrNoMem, "allocation failed");
    memcpy(*cmds, out.data(), out.size() * sizeof(float));
  }
  *ncmds = (int32_t)out.size();
  return 0;
}

void pxh_free(void* p) { free(p); }

// A fillPath / strokePath of many commands is flattened on the device, where
// its commands spread over the CUs, instead of on one host thread: it is
// recorded as the only call of an internal scene and rendered (the batched
// form gives the per-call image bit for bit).  The scene keeps no graph (it
// is re-recorded per call).  A call the device reports an error for draws
// nothing there and goes through the host path, which raises the error with
// pixie's message.
constexpr int32_t kDeviceGeomFloats = 4096;  // about 500 cubics
static pxh_scene* big_call_scene() {
  if (!g.bigScene) {
    g.bigScene = pxh_scene_new();
    if (g.bigScene) g.bigScene->graphsOff = true;
  }
  return g.bigScene;
}
// 0 when drawn; -1 when the host path must run; else a status to return
static int big_call_render(pxh_image* img) {
  pxh_scene* sc = g.bigScene;
  if (pxh_scene_count(sc) == 0) return 0;  // nothing to draw (a stroke of width <= 0)
  if (int r = pxh_scene_render_on(sc, img, 0)) return r == kErrCapacity ? -1 : r;
  int32_t st = 0;
  if (int r = pxh_scene_status(sc, &st, 1)) return r;
  return st == kOk ? 0 : -1;
}

int pxh_fill_path(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9, uint32_t rgbx,
                  int32_t winding_rule, int32_t blend_mode) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (ncmds >= kDeviceGeomFloats && big_call_scene()) {
    pxh_scene* sc = g.bigScene;
    pxh_scene_clear(sc);
    int r = pxh_scene_fill_path(sc, cmds, ncmds, mat9, rgbx, winding_rule, blend_mode);
    if (r) return r;
    r = big_call_render(img);
    if (r >= 0) return r;
  }
  if (!mat9) mat9 = kIdentity;
  Shapes sh;
  std::string err;
  int32_t r = commands_to_shapes(cmds, ncmds, true, pixel_scale(mat9), &sh, &err);
  if (r != kOk) return fail(r, err);
  transform_shapes(&sh, mat9);
  return fill_shapes_impl(img, sh, rgbx, winding_rule, blend_mode);
}

// ---- vmath Mat3 in float32 (m3_apply / m3_mul in pxh_common.h) ----
// cofactor inverse expanded along the first column index
static void m3_inverse(const float* m, float* o) {
  auto A = [m](int i, int j) { return m[3 * i + j]; };
  const float det = A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
                    A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                    A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
  const float inv = 1.0f / det;
  float r[9];
  r[0] = (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) * inv;
  r[1] = -(A(0, 1) * A(2, 2) - A(0, 2) * A(2, 1)) * inv;
  r[2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * inv;
  r[3] = -(A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) * inv;
  r[4] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * inv;
  r[5] = -(A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2)) * inv;
  r[6] = (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1)) * inv;
  r[7] = -(A(0, 0) * A(2, 1) - A(2, 0) * A(0, 1)) * inv;
  r[8] = (A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1)) * inv;
  memcpy(o, r, sizeof(r));
}
static float fix_angle(float a) {  // vmath fixAngle
  const float PI = 3.14159265358979323846f;
  while (a > PI) a -= PI * 2.0f;
  while (a < -PI) a += PI * 2.0f;
  return a;
}

// The inverse gradient matrix of a radial paint, (translate(center) *
// rotate(angle) * scale(distanceX, distanceY)).inverse() (paints.nim:192-203).
static void radial_matrix(const float* hp, float* out) {
  auto len2 = [](float x, float y) { return sqrtf(x * x + y * y); };
  const float cx = hp[0], cy = hp[1];
  const float dX = len2(cx - hp[2], cy - hp[3]), dY = len2(cx - hp[4], cy - hp[5]);
  const float ln = len2(cx - hp[2], cy - hp[3]);
  const float ga = fix_angle(atan2f((cy - hp[3]) / ln, (cx - hp[2]) / ln));
  const float c = cosf(ga), s = sinf(ga);
  const float T[9] = {1, 0, 0, 0, 1, 0, cx, cy, 1};
  const float R[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
  const float S[9] = {dX, 0, 0, 0, dY, 0, 0, 0, 1};
  float TR[9], M[9];
  m3_mul(T, R, TR);
  m3_mul(TR, S, M);
  m3_inverse(M, out);
}

// ---- transformed draws (images.nim:405-445, 531-683) ----
static const float kH = 0.5f;  // images.nim:5

// p = inv * (h, h), dx / dy = inv * (1 + h, h) / (h, 1 + h) - p
static void movement(const float* inv, float* p, float* dx, float* dy) {
  float ax, ay, bx, by;
  m3_apply(inv, 0 + kH, 0 + kH, &p[0], &p[1]);
  m3_apply(inv, 1 + kH, 0 + kH, &ax, &ay);
  m3_apply(inv, 0 + kH, 1 + kH, &bx, &by);
  dx[0] = ax - p[0];
  dx[1] = ay - p[1];
  dy[0] = bx - p[0];
  dy[1] = by - p[1];
}

// The minifyBy2 (> 0) or magnifyBy2 (< 0) steps of a draw, with the vectors
// adjusted as the loops of images.nim:415-429 / 647-661 adjust them.  pixie
// loops forever on a transform whose filterBy2 is 0 or infinite; that raises.
static int filter_steps(float* p, float* dx, float* dy, int* steps) {
  const float lx = sqrtf(dx[0] * dx[0] + dx[1] * dx[1]), ly = sqrtf(dy[0] * dy[0] + dy[1] * dy[1]);
  float f = nim_maxf(lx, ly);
  if (!(f > 0.0f) || isinf(f)) return fail(kErrArg, "degenerate draw transform");
  int n = 0;
  while (f >= 2.0f) {
    for (int k = 0; k < 2; k++) {
      p[k] /= 2;
      dx[k] /= 2;
      dy[k] /= 2;
    }
    f /= 2;
    n++;
  }
  while (f <= 0.5f) {
    for (int k = 0; k < 2; k++) {
      p[k] *= 2;
      dx[k] *= 2;
      dy[k] *= 2;
    }
    f *= 2;
    n--;
  }
  *steps = n;
  return 0;
}

// The draw source after the filter steps: minified on the device into
// ping-pong temporaries, or magnified in place (DrawSrc::mag).
static int filter_source(const uint32_t* b, int32_t bw, int32_t bh, int steps, DrawSrc* out) {
  *out = DrawSrc{b, bw, bh, 0};
  if (steps < 0) {
    const int s = -steps;
    if (s > 20 || ((int64_t)bw << s) > (1 << 30) || ((int64_t)bh << s) > (1 << 30))
      return fail(kErrCapacity, "magnified draw source too large");
    out->mag = s;
    return 0;
  }
  const uint32_t* cur = b;
  int32_t w = bw, h = bh;
  for (int i = 0; i < steps; i++) {
    const int32_t nw = (w + 1) / 2, nh = (h + 1) / 2;
    DevBuf& dst = (i & 1) ? g.minB : g.minA;
    if (int e = dst.ensure((size_t)nw * nh * 4)) return fail(e, "minify allocation failed");
    MinifyArgs m{cur, w, h, (uint32_t*)dst.p};
    if (launch_minify(m, g.stream) != 0) return fail(kErrDevice, "minify launch failed");
    cur = (const uint32_t*)dst.p;
    w = nw;
    h = nh;
  }
  *out = DrawSrc{cur, w, h, 0};
  return 0;
}

// blendRect (images.nim:468-529) of raw RGBA8 / A8 buffers
static int composite_impl(uint32_t* dst, int32_t aw, int32_t ah, const void* src, bool srcA8,
                          int32_t bw, int32_t bh, int64_t dx, int64_t dy, int32_t mode) {
  if (dx >= aw || dx + bw <= 0 || dy >= ah || dy + bh <= 0) {
    if (mode == kMask) PXH_HIP(hipMemsetAsync(dst, 0, (size_t)aw * ah * 4, g.stream));
    return 0;
  }
  CompositeArgs c;
  c.dst = dst;
  c.src = src;
  c.aw = aw;
  c.ah = ah;
  c.bw = bw;
  c.bh = bh;
  c.px = (int32_t)dx;
  c.py = (int32_t)dy;
  c.xStart = -dx > 0 ? (int32_t)-dx : 0;
  c.yStart = -dy > 0 ? (int32_t)-dy : 0;
  c.xEnd = (int64_t)bw < aw - dx ? bw : (int32_t)(aw - dx);
  c.yEnd = (int64_t)bh < ah - dy ? bh : (int32_t)(ah - dy);
  c.mode = mode;
  if (launch_composite(c, srcA8, g.stream) != 0) return fail(kErrDevice, "composite launch failed");
  return 0;
}

// draw(a, b, transform, blendMode) (images.nim:636-678), or drawTiled
// (drawCorrect with wrapped samples, images.nim:405-445, 680-683)
static int draw_transform_impl(uint32_t* dst, int32_t aw, int32_t ah, const uint32_t* b, int32_t bw,
                               int32_t bh, const float* mat9, int32_t mode, bool tiled) {
  float t[9], inv[9], p[2], dx[2], dy[2];
  memcpy(t, mat9 ? mat9 : kIdentity, sizeof(t));
  m3_inverse(t, inv);
  movement(inv, p, dx, dy);
  int steps;
  if (int r = filter_steps(p, dx, dy, &steps)) return r;
  DrawSrc src;
  if (int r = filter_source(b, bw, bh, steps, &src)) return r;
  if (tiled) {
    // inverseTransform = scale(1/2) * inverse per minify step, scale(2) *
    // inverse per magnify step: the first two rows scale by 2^-steps
    const float k = ldexpf(1.0f, -steps);
    for (int c = 0; c < 3; c++) {
      inv[3 * c + 0] *= k;
      inv[3 * c + 1] *= k;
    }
    CorrectArgs ca;
    ca.dst = dst;
    ca.aw = aw;
    ca.ah = ah;
    ca.src = src;
    memcpy(ca.inv, inv, sizeof(inv));
    ca.mode = mode;
    ca.wrapped = 1;
    if (launch_draw_correct(ca, g.stream) != 0) return fail(kErrDevice, "draw launch failed");
    return 0;
  }
  // transform * scale(2) per minify step, * scale(1/2) per magnify step:
  // the first two columns scale by 2^steps
  const float k = ldexpf(1.0f, steps);
  for (int r = 0; r < 3; r++) {
    t[r] *= k;
    t[3 + r] *= k;
  }
  const bool rotScale = !(dx[0] == 1.0f && dx[1] == 0.0f && dy[0] == 0.0f && dy[1] == 1.0f);
  const float tx = t[6], ty = t[7];
  // vmath fractional(v) = |v| - floor(|v|); only its test against 0 is used
  const bool smooth = !(sqrtf(dx[0] * dx[0] + dx[1] * dx[1]) == 1.0f &&
                        sqrtf(dy[0] * dy[0] + dy[1] * dy[1]) == 1.0f &&
                        (fabsf(tx) - floorf(fabsf(tx))) == 0.0f &&
                        (fabsf(ty) - floorf(fabsf(ty))) == 0.0f);
  if (!rotScale && !smooth) {
    // blendRect at ivec2(transform[2, 0].int32, transform[2, 1].int32);
    // offsets past the int32 range put the image off the canvas
    const float lim = 2147483520.0f;
    int64_t ox = fabsf(tx) > lim ? (int64_t)1 << 40 : (int64_t)tx;
    int64_t oy = fabsf(ty) > lim ? (int64_t)1 << 40 : (int64_t)ty;
    const void* sp = src.px;
    int32_t sw = src.w, sh = src.h;
    if (src.mag > 0) {  // the integer composite reads the magnified image itself
      const int64_t mw = (int64_t)src.w << src.mag, mh = (int64_t)src.h << src.mag;
      if (mw * mh > ((int64_t)1 << 30)) return fail(kErrCapacity, "magnified draw source too large");
      if (int e = g.magBuf.ensure((size_t)(mw * mh * 4))) return fail(e, "magnify allocation failed");
      MagnifyArgs ma{src.px, src.w, src.h, src.mag, (uint32_t*)g.magBuf.p};
      if (launch_magnify(ma, g.stream) != 0) return fail(kErrDevice, "magnify launch failed");
      sp = g.magBuf.p;
      sw = (int32_t)mw;
      sh = (int32_t)mh;
    }
    return composite_impl(dst, aw, ah, sp, false, sw, sh, ox, oy, mode);
  }
  // drawSmooth (images.nim:531-634) with the adjusted transform
  SmoothArgs sa;
  float c[4][2];
  const float sw = (float)(src.w << src.mag), sh = (float)(src.h << src.mag);
  m3_apply(t, 0, 0, &c[0][0], &c[0][1]);
  m3_apply(t, sw, 0, &c[1][0], &c[1][1]);
  m3_apply(t, sw, sh, &c[2][0], &c[2][1]);
  m3_apply(t, 0, sh, &c[3][0], &c[3][1]);
  for (int i = 0; i < 4; i++) {
    sa.per[i][0] = c[i][0];
    sa.per[i][1] = c[i][1];
    sa.per[i][2] = c[(i + 1) & 3][0];
    sa.per[i][3] = c[(i + 1) & 3][1];
  }
  float inv2[9];
  m3_inverse(t, inv2);
  movement(inv2, sa.p, sa.dx, sa.dy);
  int64_t y0 = ah, y1 = 0;
  for (int i = 0; i < 4; i++) {
    const int64_t f = nim_f2i(floorf(sa.per[i][1])), ce = nim_f2i(ceilf(sa.per[i][1]));
    y0 = f < y0 ? f : y0;
    y1 = ce > y1 ? ce : y1;
  }
  sa.yStart = (int32_t)(y0 < 0 ? 0 : (y0 > ah ? ah : y0));
  sa.yEnd = (int32_t)(y1 < 0 ? 0 : (y1 > ah ? ah : y1));
  sa.dst = dst;
  sa.aw = aw;
  sa.ah = ah;
  sa.src = src;
  sa.mode = mode;
  sa.nRuns = (aw + kSmoothRun - 1) / kSmoothRun;
  const size_t spanBytes = ((size_t)ah * sizeof(int2) + 255) & ~(size_t)255;
  if (int e = g.drawRows.ensure(spanBytes + (size_t)ah * sa.nRuns * sizeof(float2)))
    return fail(e, "draw scratch allocation failed");
  sa.span = (int2*)g.drawRows.p;
  sa.starts = (float2*)(g.drawRows.p + spanBytes);
  if (launch_draw_smooth(sa, g.stream) != 0) return fail(kErrDevice, "draw launch failed");
  return 0;
}

static int check_draw_pair(pxh_image* a, const pxh_image* b, int32_t blend_mode) {
  if (int r = check_image(a)) return r;
  if (int r = check_image(b)) return r;
  if (a->fmt != PXH_FORMAT_RGBA8 || b->fmt != PXH_FORMAT_RGBA8)
    return fail(kErrArg, "transformed draws need RGBA8 images");
  if (blend_mode < 0 || blend_mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  if (a->dev == b->dev) return fail(kErrArg, "draw source and target must differ");
  return 0;
}

int pxh_draw_transform(pxh_image* a, const pxh_image* b, const float* mat9, int32_t blend_mode) {
  PXH_INIT();
  if (int r = check_draw_pair(a, b, blend_mode)) return r;
  return draw_transform_impl((uint32_t*)a->dev, a->w, a->h, (const uint32_t*)b->dev, b->w, b->h, mat9,
                             blend_mode, false);
}

int pxh_draw_tiled(pxh_image* a, const pxh_image* b, const float* mat9, int32_t blend_mode) {
  PXH_INIT();
  if (int r = check_draw_pair(a, b, blend_mode)) return r;
  return draw_transform_impl((uint32_t*)a->dev, a->w, a->h, (const uint32_t*)b->dev, b->w, b->h, mat9,
                             blend_mode, true);
}

pxh_image* pxh_minify_by2(const pxh_image* img, int32_t power) {
  if (pxh_init() != 0) return nullptr;
  if (check_image(img) != 0) return nullptr;
  if (img->fmt != PXH_FORMAT_RGBA8) {
    fail(kErrArg, "minifyBy2 needs an RGBA8 image");
    return nullptr;
  }
  if (power < 0) {
    fail(kErrArg, "Cannot minifyBy2 with negative power");
    return nullptr;
  }
  int32_t w = img->w, h = img->h;
  for (int i = 0; i < power; i++) {
    w = (w + 1) / 2;
    h = (h + 1) / 2;
  }
  pxh_image* out = pxh_image_new(w, h);
  if (!out) return nullptr;
  if (power == 0) {
    if (pxh_image_copy(out, img) != 0) {
      pxh_image_free(out);
      return nullptr;
    }
    return out;
  }
  DrawSrc src;
  if (filter_source((const uint32_t*)img->dev, img->w, img->h, power - 1, &src) != 0) {
    pxh_image_free(out);
    return nullptr;
  }
  MinifyArgs m{src.px, src.w, src.h, (uint32_t*)out->dev};
  if (launch_minify(m, g.stream) != 0) {
    fail(kErrDevice, "minify launch failed");
    pxh_image_free(out);
    return nullptr;
  }
  return out;
}

pxh_image* pxh_magnify_by2(const pxh_image* img, int32_t power) {
  if (pxh_init() != 0) return nullptr;
  if (check_image(img) != 0) return nullptr;
  if (img->fmt != PXH_FORMAT_RGBA8) {
    fail(kErrArg, "magnifyBy2 needs an RGBA8 image");
    return nullptr;
  }
  if (power < 0) {
    fail(kErrArg, "Cannot magnifyBy2 with negative power");
    return nullptr;
  }
  if (power > 20 || ((int64_t)img->w << power) > (1 << 30) || ((int64_t)img->h << power) > (1 << 30)) {
    fail(kErrCapacity, "magnified image too large");
    return nullptr;
  }
  pxh_image* out = pxh_image_new(img->w << power, img->h << power);
  if (!out) return nullptr;
  MagnifyArgs m{(const uint32_t*)img->dev, img->w, img->h, power, (uint32_t*)out->dev};
  if (launch_magnify(m, g.stream) != 0) {
    fail(kErrDevice, "magnify launch failed");
    pxh_image_free(out);
    return nullptr;
  }
  return out;
}

// ---- non-solid paints (paths.nim:2115-2142, 2176-2214) ----

// After the path's coverage is in g.plane: build the paint's colour
// (gradients per pixel in the paint kernel, images through a draw onto a
// transparent temporary), mask it with the coverage at the paint's opacity
// and blend it onto img, in one pass over the canvas.
static int paint_route(pxh_image* img, const pxh_paint* paint, float opacity) {
  PaintArgs a = {};
  a.dst = (uint32_t*)img->dev;
  a.mask = (const uint8_t*)g.plane.p;
  a.w = img->w;
  a.h = img->h;
  a.kind = paint->kind;
  a.mode = paint->blend_mode;
  a.useOpacity = opacity != 1.0f;
  a.o255 = (uint16_t)nim_f2u32(roundf(255.0f * opacity));  // images.nim:263
  if (paint->kind == PXH_IMAGE_PAINT || paint->kind == PXH_TILED_IMAGE_PAINT) {
    const pxh_image* im = paint->image;
    if (int r = check_image(im)) return r;
    if (im->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "paint image must be RGBA8");
    const size_t bytes = (size_t)img->w * img->h * 4;
    if (int e = g.paintFill.ensure(bytes)) return fail(e, "paint allocation failed");
    PXH_HIP(hipMemsetAsync(g.paintFill.p, 0, bytes, g.stream));
    if (int r = draw_transform_impl((uint32_t*)g.paintFill.p, img->w, img->h, (const uint32_t*)im->dev,
                                    im->w, im->h, paint->image_mat, kNormal,
                                    paint->kind == PXH_TILED_IMAGE_PAINT))
      return r;
    a.fill = (const uint32_t*)g.paintFill.p;
  } else {
    // fillGradient's checks (paints.nim:98-102, 182-186, 214-218, 247-248)
    const int k = paint->kind;
    if (k == PXH_LINEAR_GRADIENT && paint->nhandles != 2)
      return fail(kErrArg, "Linear gradient requires 2 handles");
    if (k == PXH_RADIAL_GRADIENT && paint->nhandles != 3)
      return fail(kErrArg, "Radial gradient requires 3 handles");
    if (k == PXH_ANGULAR_GRADIENT && paint->nhandles != 3)
      return fail(kErrArg, "Angular gradient requires 2 handles");  // sic: 3 are read
    if (k < PXH_LINEAR_GRADIENT || k > PXH_ANGULAR_GRADIENT) return fail(kErrArg, "Paint must be a gradient");
    if (paint->nstops <= 0 || !paint->stops) return fail(kErrArg, "Gradient must have at least 1 color stop");
    if (paint->nstops > PXH_MAX_STOPS) return fail(kErrCapacity, "too many gradient stops");
    memcpy(a.hx, paint->handles, sizeof(a.hx));
    if (k == PXH_RADIAL_GRADIENT) radial_matrix(paint->handles, a.m);
    if (k == PXH_ANGULAR_GRADIENT) {
      const float vx = paint->handles[2] - paint->handles[0], vy = paint->handles[3] - paint->handles[1];
      const float l = sqrtf(vx * vx + vy * vy);
      a.ga = fix_angle(atan2f(vy / l, vx / l));
    }
    a.nstops = paint->nstops;
    memcpy(a.stops, paint->stops, sizeof(float) * 5 * (size_t)paint->nstops);
  }
  if (launch_paint(a, g.stream) != 0) return fail(kErrDevice, "paint launch failed");
  return 0;
}

// The common head of fillPath / strokePath with a non-solid paint: clamps
// the opacity (0 draws nothing), checks the paint, and readies the 8-bit
// coverage plane (cleared) as an image view for the white mask fill.
static int paint_begin(pxh_image* img, const pxh_paint* paint, float* opacity, pxh_image** mask) {
  *mask = nullptr;
  if (int r = check_image(img)) return r;
  if (img->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "paints need an RGBA8 image");
  if (!paint) return fail(kErrArg, "invalid paint");
  if (paint->kind <= 0 || paint->kind > PXH_ANGULAR_GRADIENT) return fail(kErrArg, "invalid paint kind");
  if (paint->blend_mode < 0 || paint->blend_mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  float o = paint->opacity;  // paths.nim:2101 / 2156: clamp(opacity, 0, 1)
  if (!(o > 0.0f)) o = 0.0f;
  if (o > 1.0f) o = 1.0f;
  *opacity = o;
  if (o == 0.0f) return 0;
  const size_t npx = (size_t)img->w * img->h;
  if (int e = g.plane.ensure(npx)) return fail(e, "mask allocation failed");
  PXH_HIP(hipMemsetAsync(g.plane.p, 0, npx, g.stream));
  *mask = pxh_image_view(g.plane.p, img->w, img->h, PXH_FORMAT_A8);
  if (!*mask) return kErrArg;
  return 0;
}

// fillPath / strokePath's solid route (paths.nim:2101-2112, 2156-2174):
// false when nothing is drawn, else the colour with the clamped opacity in
// its alpha as asRgbx converts it (each channel clamped, times 255 rounded
// half away from zero, then premultiplied as (c * a + 127) div 255).
static bool solid_rgbx(const pxh_paint* p, uint32_t* rgbx) {
  float o = p->opacity;
  if (!(o > 0.0f)) o = 0.0f;
  if (o > 1.0f) o = 1.0f;
  if (o == 0.0f) return false;
  if (!(p->color[3] > 0.0f || p->blend_mode == kOverwrite)) return false;
  const float c[4] = {p->color[0], p->color[1], p->color[2], p->color[3] * o};
  uint32_t b[4];
  for (int k = 0; k < 4; k++) {
    const float v = c[k] < 0.0f ? 0.0f : (c[k] > 1.0f ? 1.0f : c[k]);
    b[k] = (uint32_t)floor((double)(v * 255.0f) + 0.5);
  }
  if (b[3] != 255)
    for (int k = 0; k < 3; k++) b[k] = (b[k] * b[3] + 127) / 255;
  *rgbx = pack(b[0], b[1], b[2], b[3]);
  return true;
}

int pxh_fill_path_paint(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                        const pxh_paint* paint, int32_t winding_rule) {
  PXH_INIT();
  if (paint && paint->kind == PXH_SOLID_PAINT) {
    uint32_t rgbx;
    if (!solid_rgbx(paint, &rgbx)) return check_image(img);
    return pxh_fill_path(img, cmds, ncmds, mat9, rgbx, winding_rule, paint->blend_mode);
  }
  float opacity;
  pxh_image* mask;
  if (int r = paint_begin(img, paint, &opacity, &mask)) return r;
  if (!mask) return 0;  // opacity 0
  // the path's coverage: a white fill onto a transparent 8-bit plane, which
  // holds the alpha pixie's RGBA mask holds (its rgb equals its alpha)
  int r = pxh_fill_path(mask, cmds, ncmds, mat9, 0xFFFFFFFFu, winding_rule, kNormal);
  pxh_image_free(mask);
  if (r) return r;
  return paint_route(img, paint, opacity);
}

int pxh_stroke_path_paint(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                          const pxh_paint* paint, float stroke_width, int32_t line_cap,
                          int32_t line_join, float miter_limit, const float* dashes,
                          int32_t ndashes) {
  PXH_INIT();
  if (paint && paint->kind == PXH_SOLID_PAINT) {
    uint32_t rgbx;
    if (!solid_rgbx(paint, &rgbx)) return check_image(img);
    return pxh_stroke_path(img, cmds, ncmds, mat9, rgbx, stroke_width, line_cap, line_join,
                           miter_limit, dashes, ndashes, paint->blend_mode);
  }
  float opacity;
  pxh_image* mask;
  if (int r = paint_begin(img, paint, &opacity, &mask)) return r;
  if (!mask) return 0;
  int r = pxh_stroke_path(mask, cmds, ncmds, mat9, 0xFFFFFFFFu, stroke_width, line_cap, line_join,
                          miter_limit, dashes, ndashes, kNormal);
  pxh_image_free(mask);
  if (r) return r;
  return paint_route(img, paint, opacity);
}

int pxh_stroke_path(pxh_image* img, const float* cmds, int32_t ncmds, const float* mat9,
                    uint32_t rgbx, float stroke_width, int32_t line_cap, int32_t line_join,
                    float miter_limit, const float* dashes, int32_t ndashes, int32_t blend_mode) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (ncmds >= kDeviceGeomFloats && big_call_scene()) {
    pxh_scene* sc = g.bigScene;
    pxh_scene_clear(sc);
    int r = pxh_scene_stroke_path(sc, cmds, ncmds, mat9, rgbx, stroke_width, line_cap, line_join,
                                  miter_limit, dashes, ndashes, blend_mode);
    if (r) return r;
    r = big_call_render(img);
    if (r >= 0) return r;
  }
  if (!mat9) mat9 = kIdentity;
  float ps = pixel_scale(mat9);
  Shapes sh, st;
  std::string err;
  int32_t r = commands_to_shapes(cmds, ncmds, false, ps, &sh, &err);
  if (r != kOk) return fail(r, err);
  StrokeParams sp;
  sp.width = stroke_width;
  sp.cap = line_cap;
  sp.join = line_join;
  sp.miterLimit = miter_limit;
  sp.dashes = dashes;
  sp.ndashes = ndashes;
  r = stroke_shapes(sh, sp, ps, &st, &err);
  if (r != kOk) return fail(r, err);
  transform_shapes(&st, mat9);
  return fill_shapes_impl(img, st, rgbx, kNonZero, blend_mode);
}

int pxh_fill_shapes(pxh_image* img, const float* xy, const int32_t* poly_ofs, int32_t npolys,
                    uint32_t rgbx, int32_t winding_rule, int32_t blend_mode) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  Shapes sh;
  for (int32_t k = 0; k < npolys; k++) {
    int32_t a = poly_ofs[k], b = poly_ofs[k + 1];
    std::vector<V2> p;
    for (int32_t i = a; i < b; i++) p.push_back(v2(xy[2 * i], xy[2 * i + 1]));
    sh.add(p.data(), (int32_t)p.size());
  }
  return fill_shapes_impl(img, sh, rgbx, winding_rule, blend_mode);
}

int pxh_draw(pxh_image* a, const pxh_image* b, int32_t dx, int32_t dy, int32_t blend_mode) {
  PXH_INIT();
  if (int r = check_image(a)) return r;
  if (int r = check_image(b)) return r;
  if (a->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "draw target must be RGBA8");
  if (blend_mode < 0 || blend_mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  // blendRect, images.nim:468-483
  if (dx >= a->w || dx + b->w <= 0 || dy >= a->h || dy + b->h <= 0) {
    if (blend_mode == kMask) return pxh_image_fill(a, 0);
    return 0;
  }
  CompositeArgs c;
  c.dst = (uint32_t*)a->dev;
  c.src = b->dev;
  c.aw = a->w;
  c.ah = a->h;
  c.bw = b->w;
  c.bh = b->h;
  c.px = dx;
  c.py = dy;
  c.xStart = -dx > 0 ? -dx : 0;
  c.yStart = -dy > 0 ? -dy : 0;
  c.xEnd = b->w < a->w - dx ? b->w : a->w - dx;
  c.yEnd = b->h < a->h - dy ? b->h : a->h - dy;
  c.mode = blend_mode;
  if (launch_composite(c, b->fmt == PXH_FORMAT_A8, g.stream) != 0)
    return fail(kErrDevice, "composite launch failed");
  return 0;
}

int pxh_blur(pxh_image* img, float radius, uint32_t oob_rgbx) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (img->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "blur needs an RGBA8 image");
  int64_t r = nim_f2i(roundf(radius));  // images.nim:308
  if (r == 0) return 0;
  if (r < 0) return fail(kErrNegBlur, "Cannot apply negative blur");
  if (r > 1 << 20) return fail(kErrCapacity, "blur radius too large");
  uint32_t* p = (uint32_t*)img->dev;
  return blur_impl(p, p, img->w, img->h, (int)r, oob_rgbx, false, 0);
}

// blur with no device allocation: the intermediate image and the table live
// in the static scratch; one kernel stores the table, two run the passes, all
// on lane 0.  Calls that overlap on one GPU (other host threads, other
// streams) share the scratch and race.
int pxh_blur_noalloc(pxh_image* img, float radius, uint32_t oob_rgbx) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (img->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "blur needs an RGBA8 image");
  int64_t r = nim_f2i(roundf(radius));  // images.nim:308
  if (r == 0) return 0;
  if (r < 0) return fail(kErrNegBlur, "Cannot apply negative blur");
  if ((size_t)img->w * (size_t)img->h > kNoAllocBlurPixels)
    return fail(kErrCapacity, "image larger than the allocation-free blur's scratch (4194304 pixels)");
  if (2 * r + 1 > (int64_t)kMaxLutTaps ||
      2 * r + 1 + 8 * (int64_t)blur_table_dwords((int)r) > (int64_t)kNoAllocLutWords ||
      blur_x_lds((int)r) > g.maxLds || blur_y_lds((int)r) > g.maxLds)
    return fail(kErrCapacity, "blur radius too large for the LDS tiles");
  uint32_t* base = g.scratchPtr[g.device];
  if (!base) return fail(kErrDevice, "allocation-free scratch is not set up on this device");
  LutTaps taps;
  float f[kMaxLutTaps];
  taps.n = (int32_t)(2 * r + 1);
  gaussian_lut_into((int)r, f, taps.v);
  uint32_t* tmp = base;
  uint32_t* lut = base + kNoAllocBlurPixels;
  uint32_t* wtab = lut + taps.n;
  if (launch_lut_store(taps, lut, g.stream) != 0 || launch_blur_table(lut, (int)r, wtab, g.stream) != 0)
    return fail(kErrDevice, "table store launch failed");
  uint32_t* p = (uint32_t*)img->dev;
  const int nd = blur_table_dwords((int)r);
  BlurArgs ax{p, tmp, wtab, img->w, img->h, (int32_t)r, oob_rgbx, 0, 0, nd};
  BlurArgs ay{tmp, p, wtab, img->w, img->h, (int32_t)r, oob_rgbx, 0, 0, nd};
  if (launch_blur(ax, ay, g.stream) != 0) return fail(kErrDevice, "blur launch failed");
  return 0;
}

int pxh_blur_on(pxh_image* img, float radius, uint32_t oob_rgbx, int32_t lane) {
  PXH_INIT();
  if (lane < 0 || lane >= kLanes) return fail(kErrArg, "invalid lane");
  if (int r = check_image(img)) return r;
  if (img->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "blur needs an RGBA8 image");
  int64_t r = nim_f2i(roundf(radius));  // images.nim:308
  if (r == 0) return 0;
  if (r < 0) return fail(kErrNegBlur, "Cannot apply negative blur");
  if (r > 1 << 20) return fail(kErrCapacity, "blur radius too large");
  uint32_t* p = (uint32_t*)img->dev;
  return blur_impl(p, p, img->w, img->h, (int)r, oob_rgbx, false, 0, lane);
}

int32_t pxh_lanes(void) { return kLanes; }

int pxh_lane_synchronize(int32_t lane) {
  PXH_INIT();
  if (lane < 0 || lane >= kLanes) return fail(kErrArg, "invalid lane");
  PXH_HIP(hipStreamSynchronize(g.lane[lane]));
  return lane_scene_errors(lane);
}

int pxh_image_fill_on(pxh_image* img, uint32_t rgbx, int32_t lane) {
  PXH_INIT();
  if (lane < 0 || lane >= kLanes) return fail(kErrArg, "invalid lane");
  if (int r = check_image(img)) return r;
  if (img->fmt == PXH_FORMAT_A8)
    PXH_HIP(hipMemsetAsync(img->dev, (int)(rgbx >> 24), img->bytes, g.lane[lane]));
  else
    PXH_HIP(hipMemsetD32Async((hipDeviceptr_t)img->dev, (int)rgbx, img->bytes / 4, g.lane[lane]));
  return 0;
}

pxh_image* pxh_image_view(void* device_ptr, int32_t width, int32_t height, int32_t format) {
  if (!device_ptr || width <= 0 || height <= 0 || (format != PXH_FORMAT_RGBA8 && format != PXH_FORMAT_A8)) {
    fail(kErrArg, "invalid image view");
    return nullptr;
  }
  pxh_image* img = new pxh_image();
  img->w = width;
  img->h = height;
  img->fmt = format;
  img->dev = device_ptr;
  img->bytes = (size_t)width * height * (format == PXH_FORMAT_A8 ? 1 : 4);
  img->owned = false;
  return img;
}

int pxh_spread(pxh_image* img, float spread) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (img->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "spread needs an RGBA8 image");
  int64_t s = nim_f2i(roundf(spread));  // images.nim:702
  if (s == 0) return 0;
  if (int e = g.plane.ensure((size_t)img->w * img->h)) return fail(e, "plane allocation failed");
  SpreadArgs a{(const uint32_t*)img->dev, (uint8_t*)g.plane.p, (uint32_t*)img->dev, img->w, img->h,
               (int32_t)(s > 0 ? s : -s), s > 0 ? 1 : 0, 0, 0};
  if (launch_spread(a, g.stream) != 0) return fail(kErrDevice, "spread launch failed");
  return 0;
}

int pxh_shadow(pxh_image* out, const pxh_image* src, float offset_x, float offset_y, float spread,
               float blur, uint32_t rgbx) {
  PXH_INIT();
  if (int r = check_image(out)) return r;
  if (int r = check_image(src)) return r;
  if (out->w != src->w || out->h != src->h || out->fmt != PXH_FORMAT_RGBA8 ||
      src->fmt != PXH_FORMAT_RGBA8)
    return fail(kErrArg, "shadow needs RGBA8 images of the same size");
  int64_t br = nim_f2i(roundf(blur));
  if (br < 0) return fail(kErrNegBlur, "Cannot apply negative blur");
  const int32_t w = src->w, h = src->h;
  if (int e = g.tmp2.ensure((size_t)w * h * 4)) return fail(e, "shadow temporary allocation failed");
  uint32_t* mask = (uint32_t*)g.tmp2.p;
  int64_t s = nim_f2i(roundf(spread));
  // mask = draw(image, translate(offset), OverwriteBlend) onto transparent
  // (images.nim:765-768): whole-pixel offsets fold into the spread pass's
  // reads; fractional ones go through drawSmooth first
  const bool whole = offset_x == floorf(offset_x) && offset_y == floorf(offset_y) &&
                     fabsf(offset_x) < 1073741824.0f && fabsf(offset_y) < 1073741824.0f;
  const uint32_t* from = (const uint32_t*)src->dev;
  int32_t dx = 0, dy = 0;
  if (whole) {
    dx = (int32_t)offset_x;
    dy = (int32_t)offset_y;
  } else {
    PXH_HIP(hipMemsetAsync(mask, 0, (size_t)w * h * 4, g.stream));
    const float t[9] = {1, 0, 0, 0, 1, 0, offset_x, offset_y, 1};
    if (int r = draw_transform_impl(mask, w, h, (const uint32_t*)src->dev, w, h, t, kOverwrite, false))
      return r;
    from = mask;  // spread in place: its passes read and write different buffers
  }
  SpreadArgs sa{from, nullptr, mask, w, h, (int32_t)(s > 0 ? s : -s), s > 0 ? 1 : 0, dx, dy};
  if (!whole && s == 0) {
    // the drawn mask is ready
  } else if (s != 0) {
    // mask = draw(image, translate(offset), OverwriteBlend); spread(mask)
    if (int e = g.plane.ensure((size_t)w * h)) return fail(e, "plane allocation failed");
    sa.plane = (uint8_t*)g.plane.p;
    if (launch_spread(sa, g.stream) != 0) return fail(kErrDevice, "spread launch failed");
  } else {
    if (launch_shift(sa, g.stream) != 0) return fail(kErrDevice, "shift launch failed");
  }
  if (br != 0) {
    // blur(mask); result = fill(color).draw(mask, MaskBlend), fused into the Y pass
    return blur_impl(mask, (uint32_t*)out->dev, w, h, (int)br, 0u, true, rgbx);
  }
  if (int r = pxh_image_fill(out, rgbx)) return r;
  CompositeArgs c{(uint32_t*)out->dev, mask, w, h, w, h, 0, 0, 0, w, 0, h, kMask};
  if (launch_composite(c, false, g.stream) != 0) return fail(kErrDevice, "composite launch failed");
  return 0;
}

int pxh_apply_opacity(pxh_image* img, float opacity) {
  PXH_INIT();
  if (int r = check_image(img)) return r;
  if (img->fmt != PXH_FORMAT_RGBA8) return fail(kErrArg, "applyOpacity needs an RGBA8 image");
  uint32_t o = (uint16_t)nim_f2u32(roundf(255.0f * opacity));  // images.nim:263
  if (o == 255) return 0;
  if (o == 0) return pxh_image_fill(img, 0);
  if (launch_opacity((uint32_t*)img->dev, (int64_t)img->w * img->h, o, g.stream) != 0)
    return fail(kErrDevice, "opacity launch failed");
  return 0;
}

int pxh_blend_pixels(int32_t blend_mode, const uint32_t* backdrop, const uint32_t* source,
                     uint32_t* out, int64_t n) {
  PXH_INIT();
  if (blend_mode < 0 || blend_mode >= kNumBlendModes) return fail(kErrArg, "invalid blend mode");
  if (n <= 0) return 0;
  size_t bytes = (size_t)n * 4;
  if (int e = g.blendBuf.ensure(bytes * 3)) return fail(e, "allocation failed");
  uint32_t* d = (uint32_t*)g.blendBuf.p;
  PXH_HIP(hipMemcpyAsync(d, backdrop, bytes, hipMemcpyHostToDevice, g.stream));
  PXH_HIP(hipMemcpyAsync(d + n, source, bytes, hipMemcpyHostToDevice, g.stream));
  if (launch_blend_pixels(d, d + n, d + 2 * n, n, blend_mode, g.stream) != 0)
    return fail(kErrDevice, "blend launch failed");
  PXH_HIP(hipMemcpyAsync(out, d + 2 * n, bytes, hipMemcpyDeviceToHost, g.stream));
  PXH_HIP(hipStreamSynchronize(g.stream));
  return 0;
}

pxh_scene* pxh_scene_new(void) {
  pxh_scene* sc = new pxh_scene();
  std::lock_guard<std::mutex> lk(g.scenesMu);
  g.scenes.insert(sc);
  return sc;
}

void pxh_scene_free(pxh_scene* sc) {
  if (!sc) return;
  {
    std::lock_guard<std::mutex> lk(g.scenesMu);
    g.scenes.erase(sc);
  }
  if (g.inited) (void)hipStreamSynchronize(g.lane[sc->lane]);
  scene_free_device(sc);
  delete sc;
}

int pxh_scene_clear(pxh_scene* sc) {
  if (!sc) return fail(kErrArg, "invalid scene handle");
  if (g.inited) PXH_HIP(hipStreamSynchronize(g.lane[sc->lane]));
  scene_drop_graph(sc);
  sc->items.clear();
  sc->cmds.clear();
  sc->polys.clear();
  sc->dashes.clear();
  sc->glyphInst.clear();
  sc->textItems.clear();
  sc->textCmds = sc->textPolys = 0;
  sc->uploaded = false;
  sc->validW = sc->validH = -1;
  return 0;
}

int32_t pxh_scene_count(const pxh_scene* sc) { return sc ? (int32_t)sc->items.size() : 0; }

int pxh_scene_fill_path(pxh_scene* sc, const float* cmds, int32_t ncmds, const float* mat9,
                        uint32_t rgbx, int32_t winding_rule, int32_t blend_mode) {
  return scene_record(sc, cmds, ncmds, mat9, rgbx, winding_rule, blend_mode, false, 0, 0, 0, 4.0f);
}

int pxh_scene_stroke_path(pxh_scene* sc, const float* cmds, int32_t ncmds, const float* mat9,
                          uint32_t rgbx, float stroke_width, int32_t line_cap, int32_t line_join,
                          float miter_limit, const float* dashes, int32_t ndashes,
                          int32_t blend_mode) {
  return scene_record(sc, cmds, ncmds, mat9, rgbx, kNonZero, blend_mode, true, stroke_width,
                      line_cap, line_join, miter_limit, dashes, ndashes);
}

int pxh_scene_render(pxh_scene* sc, pxh_image* img) { return pxh_scene_render_on(sc, img, 0); }

int pxh_scene_render_on(pxh_scene* sc, pxh_image* img, int32_t lane) {
  PXH_INIT();
  if (!sc) return fail(kErrArg, "invalid scene handle");
  if (lane < 0 || lane >= kLanes) return fail(kErrArg, "invalid lane");
  if (int r = check_image(img)) return r;
  if (sc->items.empty()) return 0;
  if (sc->lane != lane) {  // the scene's buffers move to another lane
    PXH_HIP(hipStreamSynchronize(g.lane[sc->lane]));
    sc->lane = lane;
  }
  if (int r = scene_upload(sc, lane)) return r;
  hipStream_t st = g.lane[lane];
  const int32_t nBins = (img->h + kBinRows - 1) / kBinRows;
  if (nBins > sc->binSlots) {  // row bins for a taller image
    sc->binSlots = nBins;
    sc->validW = sc->validH = -1;
    if (int r = scene_alloc(sc)) return r;
  }
  if (sc->validW == img->w && sc->validH == img->h) return scene_replay(sc, img, lane);
  // First render at this size: check the pools and grow them until the render
  // fits.  A render that overflows a pool stops before touching the canvas.
  for (int attempt = 0; attempt < 8; attempt++) {
    if (int r = scene_launch(sc, img, lane, true)) return r;
    int32_t cnt[kNumCounters];
    PXH_HIP(hipMemcpyAsync(cnt, sc->pools.counters, sizeof(cnt), hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    // the run pool must hold the chunks of any later render of the same
    // scene, whatever order its waves take tasks in (scene_run_bound)
    const int64_t runNeed =
        scene_run_bound(cnt[kCntRunsUsed], scene_row_waves(g.numCUs));
    if (!cnt[kCntError]) {  // this render is complete; later ones get the slack
      if (runNeed > sc->runCap) {
        sc->runCap = (int32_t)std::min<int64_t>(runNeed + runNeed / 4, (int64_t)1 << 30);
        if (int r = scene_alloc(sc)) return r;
      }
      sc->validW = img->w;
      sc->validH = img->h;
      return 0;
    }
    if (cnt[kCntError] & 8) {
      // a band larger than the global row scratch slots: grow them
      // (at most 8 GiB of slots in all)
      const int64_t slots = scene_scratch_slots(kRowShapes[sc->rowShape], g.numCUs);
      const int64_t maxCap = std::min<int64_t>(kMaxSceneGcap, ((int64_t)8 << 30) / (slots * 24));
      if (sc->gcap * 2 > maxCap)
        return fail(kErrCapacity, "a band holds more segments than the row scratch");
      int64_t want = std::max<int64_t>((int64_t)sc->gcap * 2, cnt[kCntMaxBand]);
      int32_t gc = sc->gcap;
      while (gc < want && gc * 2 <= maxCap) gc *= 2;
      sc->gcap = gc;
    }
    auto grow = [](int32_t cap, int64_t need) {
      int64_t c = std::max<int64_t>((int64_t)cap * 2, need + need / 4 + 1024);
      return (int32_t)std::min<int64_t>(c, (int64_t)1 << 30);
    };
    if (cnt[kCntError] == 8) {  // only the row scratch was short; every pool fit
      if (int r = scene_alloc(sc)) return r;
      continue;
    }
    sc->ptCap = grow(sc->ptCap, cnt[kCntPoints]);
    sc->segCap = grow(sc->segCap, cnt[kCntSegs]);
    sc->bandCap = grow(sc->bandCap, cnt[kCntBands]);
    sc->entryCap = grow(sc->entryCap, cnt[kCntEntries]);
    sc->taskCap = grow(sc->taskCap, cnt[kCntTasks]);
    if ((cnt[kCntError] & 16) || runNeed > sc->runCap) sc->runCap = grow(sc->runCap, runNeed);
    sc->binCap = grow(sc->binCap, cnt[kCntBinEntries]);
    if (int r = scene_alloc(sc)) return r;
  }
  return fail(kErrCapacity, "scene pools could not be sized");
}

// Sets the phase-A launch shape (as the autotuner would).
int pxh_scene_set_shape(pxh_scene* sc, int32_t shape) {
  if (!sc || shape < 0 || shape >= kNumRowShapes) return fail(kErrArg, "invalid scene or shape");
  sc->rowShape = shape;
  return 0;
}

// 1 when the scene's last render replayed a captured graph.
int32_t pxh_scene_graph_active(const pxh_scene* sc) { return sc && sc->exec && !sc->graphsOff ? 1 : 0; }

// Times every phase-A launch shape on a scratch image of img's size (graph
// replays, as renders are issued) and keeps the fastest for later renders
// of this scene; the result is cached per device arch, target and size of
// the scene.
int pxh_scene_autotune(pxh_scene* sc, pxh_image* img, int32_t reps, int32_t* shape_out,
                       float* ms_out) {
  PXH_INIT();
  if (!sc) return fail(kErrArg, "invalid scene handle");
  if (int r = check_image(img)) return r;
  if (reps < 1) reps = 1;
  // a scene of the same shape of work on the same device and target reuses
  // the shape found before
  size_t ncmd = sc->cmds.size() + (size_t)sc->textCmds, bucket = 0;
  while (ncmd >>= 1) bucket++;
  const std::string key = g.arch + "|" + std::to_string(img->w) + "x" + std::to_string(img->h) + "|" +
                          std::to_string(img->fmt) + "|" + std::to_string(sc->items.size()) + "|" +
                          std::to_string(bucket);
  {
    std::lock_guard<std::mutex> lk(g.tunedMu);
    auto it = g.tuned.find(key);
    if (it != g.tuned.end()) {
      if (sc->rowShape != it->second.first) sc->validW = sc->validH = -1;
      sc->rowShape = it->second.first;
      if (shape_out) *shape_out = it->second.first;
      if (ms_out) *ms_out = it->second.second;
      return 0;
    }
  }
  pxh_image* tmp = img->fmt == PXH_FORMAT_A8 ? pxh_image_new_a8(img->w, img->h) : pxh_image_new(img->w, img->h);
  if (!tmp) return kErrNoMem;
  int best = -1;
  float bestMs = 0;
  for (int k = 0; k < kNumRowShapes; k++) {
    sc->rowShape = k;
    // validate this shape's first render (it sizes the pools and checks
    // every counter) and capture its graph, then time renders as a caller
    // issues them (graph replays on lane 0)
    sc->validW = sc->validH = -1;
    int rc = pxh_scene_render_on(sc, tmp, 0);
    if (!rc) rc = pxh_scene_render_on(sc, tmp, 0);
    float total = 0;
    if (!rc && hipEventRecord(g.ev[14], g.stream) != hipSuccess) rc = kErrDevice;
    for (int r = 0; rc == 0 && r < reps; r++) rc = pxh_scene_render_on(sc, tmp, 0);
    if (!rc && hipEventRecord(g.ev[15], g.stream) != hipSuccess) rc = kErrDevice;
    if (!rc && (hipEventSynchronize(g.ev[15]) != hipSuccess ||
                hipEventElapsedTime(&total, g.ev[14], g.ev[15]) != hipSuccess))
      rc = kErrDevice;
    // a timed render that stopped early (an overflowed pool) did less work:
    // such a shape is not a candidate
    const bool stopped = !rc && lane_scene_errors(0) != 0;
    if (rc == kErrCapacity || stopped) continue;
    if (rc) {
      pxh_image_free(tmp);
      sc->rowShape = best < 0 ? kDefaultRowShape : best;
      return rc;
    }
    if (best < 0 || total / reps < bestMs) {
      bestMs = total / reps;
      best = k;
    }
  }
  if (best < 0) {
    pxh_image_free(tmp);
    sc->rowShape = kDefaultRowShape;
    sc->validW = sc->validH = -1;
    return fail(kErrCapacity, "no phase-A launch shape rendered this scene");
  }
  pxh_image_free(tmp);
  if (sc->rowShape != best) sc->validW = sc->validH = -1;
  sc->rowShape = best;
  {
    std::lock_guard<std::mutex> lk(g.tunedMu);
    g.tuned[key] = {best, bestMs};
  }
  if (shape_out) *shape_out = best;
  if (ms_out) *ms_out = bestMs;
  return 0;
}

// Per-call status of the last render (0 ok, else the PixieError code of that
// call, e.g. PXH_ERR_INT_OVERFLOW); waits for the render.
int pxh_scene_status(pxh_scene* sc, int32_t* status, int32_t n) {
  PXH_INIT();
  if (!sc || !sc->uploaded) return fail(kErrArg, "scene not rendered");
  int32_t m = std::min<int32_t>(n, (int32_t)sc->items.size());
  std::vector<ItemGeom> geo(m);
  hipStream_t st = g.lane[sc->lane];
  PXH_HIP(hipMemcpyAsync(geo.data(), sc->pools.geo, m * sizeof(ItemGeom), hipMemcpyDeviceToHost, st));
  PXH_HIP(hipStreamSynchronize(st));
  for (int32_t i = 0; i < m; i++) status[i] = geo[i].status;
  return lane_scene_errors(sc->lane);
}

int pxh_scene_debug_limit_runs(pxh_scene* sc, int32_t cap) {
  PXH_INIT();
  if (!sc || !sc->uploaded) return fail(kErrArg, "scene not rendered");
  if (cap < 0 || cap > sc->runCap) return fail(kErrArg, "run cap outside the validated pool");
  sc->pools.runCap = cap;
  sc->gen++;  // renders re-capture their graph with the lowered capacity
  return 0;
}

// Device geometry of call i of the last render, in the layout of
// pxh_debug_fill_geometry (for tests against the host geometry).
int pxh_scene_debug_geometry(pxh_scene* sc, int32_t i, int32_t* info16, int32_t** parts4,
                             int32_t* nparts, float** entries4, int8_t** windings,
                             int32_t* nentries) {
  PXH_INIT();
  if (!sc || !sc->uploaded || i < 0 || i >= (int32_t)sc->items.size())
    return fail(kErrArg, "invalid scene call");
  PXH_HIP(hipStreamSynchronize(g.lane[sc->lane]));
  FillInfo f;
  ItemGeom geo;
  PXH_HIP(hipMemcpy(&f, sc->pools.fills + i, sizeof(f), hipMemcpyDeviceToHost));
  PXH_HIP(hipMemcpy(&geo, sc->pools.geo + i, sizeof(geo), hipMemcpyDeviceToHost));
  memcpy(info16, &f, sizeof(FillInfo));
  int32_t np = (f.status == kOk && f.pathWidth > 0) ? f.nParts : 0;
  std::vector<PartInfo> parts(np);
  if (np) PXH_HIP(hipMemcpy(parts.data(), sc->pools.parts + geo.bandBase, np * sizeof(PartInfo),
                            hipMemcpyDeviceToHost));
  int32_t ne = 0;
  for (auto& p : parts) ne += p.count;
  *parts4 = (int32_t*)malloc(sizeof(PartInfo) * (np + 1));
  *entries4 = (float*)malloc(sizeof(Entry) * (ne + 1));
  *windings = (int8_t*)malloc(ne + 1);
  int32_t k = 0;
  for (int32_t b = 0; b < np; b++) {
    PartInfo q = parts[b];
    if (q.count) {
      PXH_HIP(hipMemcpy(*entries4 + 4 * k, sc->pools.entries + q.entryBase, q.count * sizeof(Entry),
                        hipMemcpyDeviceToHost));
      PXH_HIP(hipMemcpy(*windings + k, sc->pools.entW + q.entryBase, q.count, hipMemcpyDeviceToHost));
    }
    parts[b].entryBase = k;
    k += q.count;
  }
  if (np) memcpy(*parts4, parts.data(), np * sizeof(PartInfo));
  *nparts = np;
  *nentries = ne;
  return 0;
}

// Reads (and optionally resets) the count of pixels written by fills since
// the last reset (N_touch, the unit of the fill throughput metric).
int pxh_touched_pixels(int64_t* out, int32_t reset) {
  PXH_INIT();
  unsigned long long v = 0;
  PXH_HIP(hipMemcpyAsync(&v, g.touched, sizeof(v), hipMemcpyDeviceToHost, g.stream));
  PXH_HIP(hipStreamSynchronize(g.stream));
  *out = (int64_t)v;
  if (reset) PXH_HIP(hipMemsetAsync(g.touched, 0, sizeof(unsigned long long), g.stream));
  return 0;
}

// Per-launch timing of the raster kernels: between begin and end every fill
// kernel launch is bracketed by events; end returns the summed kernel time.
int pxh_profile_begin(void) {
  PXH_INIT();
  g.profiling = true;
  return 0;
}

int pxh_profile_end(double* kernel_ms, int64_t* launches) {
  PXH_INIT();
  g.profiling = false;
  double tot = 0;
  PXH_HIP(hipStreamSynchronize(g.stream));
  g.kernelTimes.clear();
  for (auto& pr : g.evPairs) {
    float ms = 0;
    PXH_HIP(hipEventElapsedTime(&ms, pr.a, pr.b));
    tot += ms;
    auto& kt = g.kernelTimes[pr.name];
    kt.first += ms;
    kt.second++;
    g.evPool.push_back(pr.a);
    g.evPool.push_back(pr.b);
  }
  *kernel_ms = tot;
  *launches = (int64_t)g.evPairs.size();
  g.evPairs.clear();
  return 0;
}

// Per-kernel share of the last profile (after pxh_profile_end).
int pxh_profile_kernel(const char* name, double* kernel_ms, int64_t* launches) {
  auto it = g.kernelTimes.find(name ? name : "");
  *kernel_ms = it == g.kernelTimes.end() ? 0.0 : it->second.first;
  *launches = it == g.kernelTimes.end() ? 0 : it->second.second;
  return 0;
}

// Records event slot (0..15) on the library stream.
int pxh_event_record(int32_t slot) {
  PXH_INIT();
  if (slot < 0 || slot >= 16) return fail(kErrArg, "event slot out of range");
  PXH_HIP(hipEventRecord(g.ev[slot], g.stream));
  return 0;
}

// Milliseconds between two recorded event slots (waits for the second).
int pxh_event_elapsed(int32_t a, int32_t b, float* ms) {
  PXH_INIT();
  if (a < 0 || a >= 16 || b < 0 || b >= 16) return fail(kErrArg, "event slot out of range");
  PXH_HIP(hipEventSynchronize(g.ev[b]));
  PXH_HIP(hipEventElapsedTime(ms, g.ev[a], g.ev[b]));
  return 0;
}

// Host geometry of one fillPath / strokePath call without touching the
// device: the FillInfo, bands, entries and windings the raster kernel would
// receive.  Arrays are malloc'd (release with pxh_free).  stroke_width <= 0
// means fill.  For tests of the geometry against the reference semantics.
int pxh_debug_fill_geometry(const float* cmds, int32_t ncmds, const float* mat9, int32_t w,
                            int32_t h, int32_t stroke, float stroke_width, int32_t line_cap,
                            int32_t line_join, float miter_limit, int32_t* info16,
                            int32_t** parts4, int32_t* nparts, float** entries4,
                            int8_t** windings, int32_t* nentries) {
  if (!mat9) mat9 = kIdentity;
  float ps = pixel_scale(mat9);
  Shapes sh, st;
  std::string err;
  int32_t r = commands_to_shapes(cmds, ncmds, !stroke, ps, &sh, &err);
  if (r != kOk) return fail(r, err);
  if (stroke) {
    StrokeParams sp;
    sp.width = stroke_width;
    sp.cap = line_cap;
    sp.join = line_join;
    sp.miterLimit = miter_limit;
    r = stroke_shapes(sh, sp, ps, &st, &err);
    if (r != kOk) return fail(r, err);
    sh = st;
  }
  transform_shapes(&sh, mat9);
  FillGeom geo;
  r = build_fill(sh, w, h, 0xFFFFFFFFu, 0, 0, &geo, &err);
  memcpy(info16, &geo.info, sizeof(FillInfo));
  *nparts = (int32_t)geo.parts.size();
  *nentries = (int32_t)geo.entries.size();
  *parts4 = (int32_t*)malloc(sizeof(PartInfo) * (geo.parts.size() + 1));
  *entries4 = (float*)malloc(sizeof(Entry) * (geo.entries.size() + 1));
  *windings = (int8_t*)malloc(geo.windings.size() + 1);
  if (!geo.parts.empty()) memcpy(*parts4, geo.parts.data(), sizeof(PartInfo) * geo.parts.size());
  if (!geo.entries.empty()) memcpy(*entries4, geo.entries.data(), sizeof(Entry) * geo.entries.size());
  if (!geo.windings.empty()) memcpy(*windings, geo.windings.data(), geo.windings.size());
  if (r != kOk) return fail(r, err);
  return 0;
}

}  // extern "C"

// ---- text calls through a scene (fonts.nim:566-690) ----
namespace pxh {

int text_glyph_ref(const void* typeface, uint32_t rune, const float* cmds, int32_t n, int32_t* ref) {
  const auto key = std::make_pair(typeface, rune);
  auto it = g.glyphKeys.find(key);
  if (it != g.glyphKeys.end()) {
    *ref = it->second;
    return 0;
  }
  Ctx::GlyphRec r;
  r.cmd = (int32_t)g.glyphCmds.size();
  r.poly = (int32_t)g.glyphPolys.size();
  if (int e = canonicalize(cmds, n, &g.glyphCmds, &g.glyphPolys)) {
    g.glyphCmds.resize(r.cmd);
    g.glyphPolys.resize(r.poly);
    return e;
  }
  r.nCmds = (int32_t)g.glyphCmds.size() - r.cmd;
  r.nPolys = (int32_t)g.glyphPolys.size() - r.poly;
  *ref = (int32_t)g.glyphRecs.size();
  g.glyphRecs.push_back(r);
  g.glyphKeys[key] = *ref;
  return 0;
}

// Typefaces changed (freed, or their fallbacks): later lookups re-record.
void text_glyphs_forget() { g.glyphKeys.clear(); }

pxh_scene* text_scene() {
  if (pxh_init() != 0) return nullptr;
  pxh_scene* sc = big_call_scene();
  if (sc) pxh_scene_clear(sc);
  return sc;
}

int text_record(pxh_scene* sc, const std::vector<TextInst>& inst, const float* mat9, uint32_t rgbx, int32_t rule,
                int32_t mode, bool stroke, float width, int32_t cap, int32_t join, float miterLimit,
                const float* dashes, int32_t ndashes) {
  return text_record_impl(sc, inst, mat9, rgbx, rule, mode, stroke, width, cap, join, miterLimit, dashes,
                          ndashes);
}

// 0 when drawn; -1 when the pools ran out (the canvas is untouched and the
// caller draws the calls one by one); else the first failing call's error
// (the other calls are drawn).
int text_scene_render(pxh_image* img) {
  pxh_scene* sc = g.bigScene;
  const int32_t n = pxh_scene_count(sc);
  if (n == 0) return 0;
  if (int r = pxh_scene_render_on(sc, img, 0)) return r == kErrCapacity ? -1 : r;
  std::vector<int32_t> st((size_t)n);
  if (int r = pxh_scene_status(sc, st.data(), n)) return r;
  for (int32_t k = 0; k < n; k++) {
    if (st[k] == kOk) continue;
    if (st[k] == kErrIntOverflow) return fail(st[k], "Path int overflow detected");
    if (st[k] == kErrDash) return fail(st[k], "Invalid line dash value");
    if (st[k] == kErrDiscretize) return fail(st[k], "Unable to discretize curve");
    return fail(st[k], "a text call failed");
  }
  return 0;
}

bool paint_solid_rgbx(const pxh_paint* p, uint32_t* rgbx) { return solid_rgbx(p, rgbx); }

}  // namespace pxh
