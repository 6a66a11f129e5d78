// pxh_scene.h -- device layout of a recorded scene (batched form) and the
// launchers of its kernels (pxh_scene.hip).
//
// A scene is an ordered list of fillPath / strokePath calls whose canonical
// commands stay resident in HBM.  Each render runs on the device:
//   geometry  one workgroup per call: adaptive flattening (count -> scan ->
//             emit), stroke expansion, shapesToSegments, bounds, band layout
//   bands     one wave per y-band: ordered entry lists (partitionSegments)
//   rows      phase A: every (call, row, window) row program in parallel,
//             each leaving its pixel codes as sparse (x, delta) runs
//   compose   phase B: each wave owns a canvas row window in registers and
//             applies the runs of every call covering it, in call order
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pxh_common.h"

namespace pxh {

// Canonical command: absolute Move / Line / Cubic / Quad / Close.  Every
// subpath starts with a Move (the host inserts the implicit ones).
struct SceneCmd {
  float kind;          // pixie command ordinal: Close 0, Move 1, Line 2, Cubic 5, Quad 7
  float p[6];          // absolute parameters
  float sx, sy;        // start point of the command's subpath
  int32_t pad[3];
};

struct ScenePoly {     // one subpath of a call
  int32_t firstCmd, numCmds;  // absolute SceneCmd indices; firstCmd is its Move
  int32_t noClose;            // ended by a relative move: no implicit closing point
  int32_t pad;
};

// One glyph (or underline / strikethrough rect) of a text call, placed into
// the call's command and subpath range on the device (scene_launch_expand):
// a glyph's canonical commands are copied from the glyph arena with every
// point transformed by m (a vmath Mat3 without its last row: m[0] m[1] are
// column 0, m[2] m[3] column 1, m[4] m[5] the translation); a rect
// (srcCmd < 0) is path.rect(m[0], m[1], m[2], m[3]) written out.
struct GlyphInst {
  int32_t srcCmd, nCmds;    // arena commands
  int32_t srcPoly, nPolys;  // arena subpaths
  int32_t dstCmd, dstPoly;  // absolute destination in the scene
  int32_t polyIdx;          // subpath index of the first subpath within its call
  int32_t identity;         // m is the identity (the points are copied)
  float m[6];
};

struct SceneItem {
  int32_t cmdBase, cmdCount;   // SceneCmd range
  int32_t polyBase, polyCount; // ScenePoly range
  int32_t stroke;              // 0 fill, 1 stroke
  uint32_t rgbx;
  int32_t mode, rule;
  int32_t cap, join;
  float halfStroke;            // strokeWidth / 2
  float miterAngleLimit;       // arcsin(1 / miterLimit) * 2
  float minArea;               // pixelErrorMargin / pixelScale
  float errSq;                 // (pixelErrorMargin / pixelScale)^2
  float mat[9];
  int32_t identity;
  int32_t strokeNone;          // strokeWidth <= 0: strokeShapes yields no shapes
  int32_t dashBase, dashCount; // dash pattern in ScenePools::dashes (even length)
  int32_t dashBad;             // a dash value <= 0: the stroke raises "Invalid line dash value"
  int32_t pad;
};

// Per-render derived data of one call.
struct ItemGeom {
  int32_t ptBase, ptCount;
  int32_t segBase, segCount;
  int32_t bandBase;
  int32_t rowStart, rowCount;  // rows that run a row program
  int32_t winBase, nWin;       // first column window, windows per row
  int32_t taskBase;            // first phase-A task
  int32_t status;
  int32_t pad;
};

// Phase-A output of one (call, row, window).
struct RowRuns {
  int32_t off, n;    // run entries in the run pool
  int32_t kind;      // RowKind
  int32_t pad;
};

// Run entry: x in the low 28 bits (tag in bit 30 = area factor entry),
// value = code delta at x, or the pixel's area factor.
constexpr int32_t kRunArea = 1 << 30;
constexpr int32_t kRunXMask = (1 << 28) - 1;

// Phase-A waves take run-pool space in chunks of kRunChunk entries.  One
// task of a 512-px window writes at most kMaxTaskRuns entries (513 deltas +
// 512 area factors), and a wave leaves a chunk only when the most the next
// task can write does not fit, so every chunk a wave leaves holds more than
// kRunChunk - kMaxTaskRuns written entries.  Whatever order the waves take tasks in, a
// render of U written entries takes at most U / (kRunChunk - kMaxTaskRuns)
// left chunks plus one open chunk per wave (scene_run_bound).
constexpr int kRowGroup = 8;       // rows of a grouped phase-A work item
constexpr int kRowGranule = 32;    // phase-A work items a queue takes in a row
// phase-B units a queue takes in a row: neighbouring units share a row bin's
// call list and call records (64 units are one 16-row bin of a 4096-px canvas)
constexpr int kComposeGranule = 64;
constexpr int kGroupMaxClass = 7;  // grouped rows' bands hold fewer than 128 entries
constexpr int kRunChunk = 2048;
constexpr int kMaxTaskRuns = 1025;
inline int64_t scene_run_bound(int64_t used, int64_t waves) {
  return (used / (kRunChunk - kMaxTaskRuns) + 1 + waves) * kRunChunk;
}

// segBand flag: the segment needs anti-aliasing (requiresAntiAliasing).
constexpr int32_t kSegAA = 1 << 30;
constexpr int32_t kBigChunk = 4096;  // points per chunk of a big fill's segment stage
constexpr int32_t kBigPolyChunk = 1024;  // subpaths per chunk of a big call's point layout

enum SceneCounter {
  kCntPoints = 0, kCntSegs, kCntBands, kCntEntries, kCntTasks,
  kCntRuns,      // run-pool entries handed out in chunks
  kCntRunsUsed,  // run entries written
  kCntError, kCntMaxBand, kCntBinEntries,
  kCntBigTasks,  // phase-A work items of the large-band launch (the first in their order)
  kCntGroups,    // phase-A work items (set by the class scan)
  kNumCounters
};

// A phase-A task is one (call, row, window).  A wave takes a work item:
// kRowGroup consecutive rows of one call and window when the bands of those
// rows hold at most kGroupMaxClass-bit entry counts (the call's records are
// then read once and the rows' band entries mostly come from the L1 cache),
// or else a single row, so that no item is long.  Items run longest first:
// an item's cost class is the bit length of the largest band entry count
// among its rows, and the items are ordered by class, descending.  Work
// items are named by the task of their first row; the task map marks
// grouped items (taskItem bit 31) and gives the other rows of a group no
// class (-1).
// The class histogram and the scatter cursors follow the counters.
constexpr int kCostClasses = 32;
constexpr int kCntClassHist = kNumCounters;                  // kCostClasses ints
constexpr int kCntClassCursor = kNumCounters + kCostClasses;  // kCostClasses ints
// The band-entry scan runs over kEntryTiles contiguous tiles of bands; the
// tile sums, then the tile offsets, follow the class cursors.
constexpr int kEntryTiles = 256;
constexpr int kCntTileSum = kNumCounters + 2 * kCostClasses;  // kEntryTiles ints
constexpr int kCounterInts = kNumCounters + 2 * kCostClasses + kEntryTiles;

// Phase B looks calls up by row bins of kBinRows canvas rows: each bin lists,
// in call order, the calls whose rows meet it.
constexpr int kBinRows = 16;

// Phase A and B deal out work through kQueues independent queues (queue q
// serves work items q, q + kQueues, ...), so that no single counter takes
// every fetch.
constexpr int kQueues = 16;
constexpr int kQueueStride = 64;  // ints between queue heads (256 B)
// Queue sets: phase A's large-band launch, its small-band launch, phase B.
constexpr int kQueueSets = 3;

// Grid of a queue-fed kernel: at least one wave per queue, a multiple of
// kQueues so that every queue is served by the same number of waves.
inline int scene_queue_grid(int waves) {
  if (waves < kQueues) waves = kQueues;
  return (waves + kQueues - 1) / kQueues * kQueues;
}

struct ScenePools {
  // recorded inputs
  const SceneItem* items;
  const SceneCmd* cmds;
  const ScenePoly* polys;
  const int32_t* itemOrder;  // geometry workgroup -> call
  const float* dashes;       // dash patterns of the recorded strokes
  int32_t nItems;
  int32_t nBig;              // calls of more than 256 commands, first in itemOrder
  const int32_t* bigCmdStart;  // their command counts' exclusive prefix (nBig + 1)
  int32_t bigCmds;           // bigCmdStart[nBig]
  int32_t W, H;              // target size
  // per-call outputs
  ItemGeom* geo;
  FillInfo* fills;
  // per-command / per-subpath scratch
  int32_t* cmdCount;
  int32_t* cmdOff;
  int32_t* polyPt;
  int32_t* polyN;
  // pools
  float2* pts;
  int32_t* ptPoly;
  int32_t* ptSeg;            // segments each point emits (geometry scratch)
  int32_t ptCap;
  // segments of the big fill calls, in chunks of kBigChunk points spread
  // over the device: chunk prefix per big call (nBig + 1), per chunk
  // (segment count, segment offset in the call) and bounds, and the prefix
  // of their 64-segment groups (nBig + 1)
  int32_t* bigChunkStart;
  int32_t* bigPolyStart;     // subpath-chunk prefix of the big calls (nBig + 1)
  int4* bigPolyChunk;        // per subpath chunk: points, error, point offset
  int2* bigChunkCnt;
  float4* bigChunkBd;
  int32_t* bigGroupStart;
  float4* segs;
  int8_t* segW;
  int2* segBand;             // per segment: first band, last band | kSegAA
  int2* segChunk;            // per 64-segment chunk (one call's): band range
  int32_t segCap;
  PartInfo* parts;
  int32_t* bandItem;
  int32_t bandCap;
  Entry* entries;
  int8_t* entW;
  int32_t entryCap;
  RowRuns* tasks;
  int32_t* taskItem;         // call of each work item
  int32_t* taskOrder;        // work items in the order phase A takes them (longest first)
  int8_t* taskClass;         // cost class of each work item
  int32_t taskCap;
  int2* runs;
  int32_t runCap;
  int4* itemRows;            // per call: first row, end row, x range (empty if it draws nothing)
  int4* itemPaint;           // per call: rgbx, mode, first task, windows per row
  int4* itemSpan;            // per call: rowStart, rowCount, xlo, xhi
  int32_t* binCount;         // nBins
  int32_t* binOff;           // nBins + 1
  int32_t* binList;          // call indices per bin
  int32_t binCap;
  int32_t binSlots;          // capacity of binCount / binOff (bins)
  int32_t* queues;           // work-queue heads: 2 phases x kQueues, one per 256 B
  int32_t* counters;         // SceneCounter
  // row scratch for bands larger than the LDS capacity
  char* gscratch;
  int32_t gcap;              // entries per wave slot
  int32_t ldsCap;            // entries that fit in LDS
  int32_t bigClass;          // first cost class of the large-band launch
};

// Phase-A launch shapes.  The row program is latency-bound, so each launch
// keeps about as many waves resident as LDS and registers allow for the
// band entries its waves keep in LDS.
//   kRowsSingle      one launch of every task: ldsCap entries in LDS, larger
//                    selections in the global row scratch
//   kRowsConcurrent  tasks whose band holds more than ldsCap entries (their
//                    cost class is above it; they come first in the task
//                    order) run in a second launch of bigWaves waves per CU
//                    with bigCap entries in LDS, beside the first launch
//   kRowsSequential  the same two launches, one after the other
// Only waves whose tasks may exceed their LDS capacity get global row
// scratch slots.  A scene's autotuner picks one shape; run-pool sizing
// assumes the most waves any shape has.
enum RowMode { kRowsSingle = 0, kRowsConcurrent = 1, kRowsSequential = 2 };
struct RowShape {
  int ldsCap, wavesPerCU;
  int mode;
  int bigCap, bigWaves;
};
constexpr RowShape kRowShapes[] = {
    {128, 20, kRowsSingle, 0, 0},
    {128, 20, kRowsConcurrent, 512, 6},
    {128, 20, kRowsSequential, 512, 9},
    {128, 16, kRowsConcurrent, 256, 10},
    {64, 24, kRowsSingle, 0, 0},
    {256, 14, kRowsSingle, 0, 0},
};
constexpr int kNumRowShapes = 6;
constexpr int kDefaultRowShape = 0;
// Waves of a shape's phase-A launches together, and the most of any shape.
inline int scene_shape_waves(const RowShape& s, int numCUs) {
  return scene_queue_grid(numCUs * s.wavesPerCU) + (s.mode ? scene_queue_grid(numCUs * s.bigWaves) : 0);
}
inline int scene_row_waves(int numCUs) {
  int m = 0;
  for (const RowShape& s : kRowShapes) m = m > scene_shape_waves(s, numCUs) ? m : scene_shape_waves(s, numCUs);
  return m;
}
// Wave slots of global row scratch a shape uses.
inline int scene_scratch_slots(const RowShape& s, int numCUs) {
  return s.mode ? scene_queue_grid(numCUs * s.bigWaves) : scene_queue_grid(numCUs * s.wavesPerCU);
}

int scene_set_lds_limits(size_t bytes);
size_t scene_gscratch_bytes(int rowGrid, int gcap);  // global row scratch of scene_launch_rows
int scene_launch_expand(const GlyphInst* inst, int32_t n, const SceneCmd* arenaCmds,
                        const ScenePoly* arenaPolys, SceneCmd* cmds, ScenePoly* polys, hipStream_t s);
int scene_launch_clear(const ScenePools& p, hipStream_t s);
int scene_launch_geometry(const ScenePools& p, hipStream_t s);
int scene_launch_bands(const ScenePools& p, int grid, hipStream_t s);
int scene_launch_bins(const ScenePools& p, int grid, hipStream_t s);
// Phase A in the given shape; a concurrent shape forks its small-band
// launch onto aux (fork / join through the two events).
int scene_launch_rows(const ScenePools& p, const RowShape& shape, int numCUs, hipStream_t s,
                      hipStream_t aux, hipEvent_t fork, hipEvent_t join);
int scene_launch_compose(const ScenePools& p, void* canvas, bool a8, unsigned long long* touched,
                         int32_t* errOut, int grid, hipStream_t s);

}  // namespace pxh
