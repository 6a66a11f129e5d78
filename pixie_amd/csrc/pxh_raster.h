// pxh_raster.h -- the per-row raster program of fillShapes (paths.nim:1631-1908)
// as wave-level (64-lane) device code, and its fused scan -> paint -> blend ->
// write-back epilogue.  Included by the HIP translation units only.
//
// One wavefront owns one (fill, row, window) item.  The row program decides
// pixie's per-row path (two-spanning fill, trapezoid, 5-sample coverage or
// 1-sample fill), computes the row's pixel operations as runs of codes and
// adds them to a difference array D in LDS.  The epilogue prefix-scans D
// across the window (one wave scan per 256 pixels) and applies each pixel's
// operation to the canvas in HBM with 16-byte accesses.
//
// Pixel codes (prefix sums of D):
//   coverage rows: 0..255 = accumulated coverage; kClear = cleared (mask mode)
//   other rows:    0 = untouched, kFill = blend with paint, kArea = blend with
//                  paint * AX[x] (trapezoid sliver), kClear = cleared
#pragma once

#include <hip/hip_runtime.h>

#include "pxh_common.h"
#include "pxh_geom.h"
#include "pxh_pixel.h"

namespace pxh {

constexpr int kWave = 64;
constexpr int32_t kFill = 256;
constexpr int32_t kArea = 512;
constexpr int32_t kClear = 1024;

enum RowKind : int32_t { kRowNone = 0, kRowCoverage = 1, kRowOps = 2 };

// Per-wave row scratch.  cap = capacity (power of two >= 64) for selected
// entries / hits of one row; ww = window width in pixels (multiple of 256).
// D and AX always live in LDS; the cap-sized arrays live in LDS when they fit
// and otherwise in a per-wave global scratch slot (same code either way).
struct RowLds {
  int32_t* D;        // ww + 4: difference array of pixel codes
  uint32_t* AX;      // ww: coverage factor of trapezoid sliver pixels
  uint64_t* keys;    // cap: sort keys (value << 32 | position)
  int32_t* A;        // cap: selected entries (band-local indices)
  int32_t* B;        // cap: reordered selection / hit windings by position
  float* tx0;        // cap: trapezoid x at scanTop (by selection position)
  float* tx1;        // cap: trapezoid x at scanBottom
  int cap, ww;
  uint64_t* wide;    // LDS staging of the wide path's hit keys (wideCap of
  int wideCap;       // them; 0 turns it off): AX, which coverage rows leave
                     // free, and the cap arrays after it when they are in LDS
};

PXH_HD size_t row_lds_fixed_bytes(int ww) { return (size_t)(ww + 4) * 4 + (size_t)ww * 4; }
PXH_HD size_t row_cap_bytes(int cap) { return (size_t)cap * 24; }

// D/AX from `fixed` (LDS), cap arrays from `arr` (LDS or global).
__device__ __forceinline__ RowLds row_lds_carve(char* fixed, char* arr, int cap, int ww) {
  RowLds L;
  L.cap = cap;
  L.ww = ww;
  L.D = (int32_t*)fixed;
  L.AX = (uint32_t*)(L.D + ww + 4);
  L.keys = (uint64_t*)arr;
  L.A = (int32_t*)(L.keys + cap);
  L.B = L.A + cap;
  L.tx0 = (float*)(L.B + cap);
  L.tx1 = L.tx0 + cap;
  L.wide = nullptr;
  L.wideCap = 0;
  return L;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Wave-wide scans with DPP lane moves (GFX9 row_shr / row_bcast): a
// Hillis-Steele scan inside each 16-lane row, then rows 1..3 pick up the
// totals of the rows before them.  Lanes a move has no source for keep the
// identity.  No LDS traffic, about two VALU operations per step.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143,
              kDppWaveShr1 = 0x138;

__device__ __forceinline__ int32_t wave_incl_sum(int32_t v) {
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr1, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr2, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr4, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr8, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowBcast15, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowBcast31, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

__device__ __forceinline__ int32_t wave_incl_max(int32_t v) {
  const int32_t lo = INT32_MIN;
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr1, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr2, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr4, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr8, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowBcast15, 0xa, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowBcast31, 0xc, 0xf, false));
  return v;
}

// Exclusive prefix sum over the wave; *total receives the wave's sum.
__device__ __forceinline__ int32_t wave_excl_scan(int32_t v, int32_t* total) {
  int32_t x = wave_incl_sum(v);
  *total = __builtin_amdgcn_readlane(x, kWave - 1);
  return x - v;
}

// Exclusive prefix max (INT32_MIN in lane 0).
__device__ __forceinline__ int32_t wave_excl_max(int32_t v) {
  int32_t x = wave_incl_max(v);
  return __builtin_amdgcn_update_dpp(INT32_MIN, x, kDppWaveShr1, 0xf, 0xf, false);
}

__device__ __forceinline__ int32_t wave_max(int32_t v) {
  return __builtin_amdgcn_readlane(wave_incl_max(v), kWave - 1);
}

// Scans inside independent SEG-lane segments of the wave (SEG = 16 or 32).
template <int SEG>
__device__ __forceinline__ int32_t seg_incl_sum(int32_t v) {
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr1, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr2, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr4, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr8, 0xf, 0xf, false);
  if (SEG == 32) v += __builtin_amdgcn_update_dpp(0, v, kDppRowBcast15, 0xa, 0xf, false);
  return v;
}

template <int SEG>
__device__ __forceinline__ int32_t seg_excl_max(int32_t v) {
  const int32_t lo = INT32_MIN;
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr1, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr2, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr4, 0xf, 0xf, false));
  v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr8, 0xf, 0xf, false));
  if (SEG == 32) v = imax(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowBcast15, 0xa, 0xf, false));
  v = __builtin_amdgcn_update_dpp(lo, v, kDppWaveShr1, 0xf, 0xf, false);
  return (lane_id() & (SEG - 1)) == 0 ? lo : v;
}

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// Bitonic sort of n keys in L.keys (n <= cap, cap a power of two).  Keys
// beyond n are padded with UINT64_MAX.  The whole wave participates.
__device__ void wave_sort_keys(uint64_t* keys, int n, int cap) {
  int lane = lane_id();
  int np = 64;
  while (np < n) np <<= 1;
  if (np > cap) np = cap;  // n <= cap guaranteed by the caller
  if (n <= 32) {
    // rank sort: keys are distinct (their low word is the position), so each
    // lane's rank is the number of smaller keys, found by broadcasting every
    // key in turn
    uint64_t k = lane < n ? keys[lane] : ~0ull;
    uint32_t klo = (uint32_t)k, khi = (uint32_t)(k >> 32);
    int rank = 0;
    for (int j = 0; j < n; j++) {
      uint32_t jlo = __builtin_amdgcn_readlane(klo, j), jhi = __builtin_amdgcn_readlane(khi, j);
      rank += (jhi < khi || (jhi == khi && jlo < klo)) ? 1 : 0;
    }
    wave_sync();
    if (lane < n) keys[rank] = k;
    wave_sync();
    return;
  }
  for (int i = n + lane; i < np; i += kWave) keys[i] = ~0ull;
  wave_sync();
  if (np == 64) {
    uint64_t k = keys[lane];
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        uint64_t o = __shfl_xor(k, stride, kWave);
        bool up = ((lane & size) == 0);
        bool lower = (lane & stride) == 0;
        uint64_t mn = k < o ? k : o, mx = k < o ? o : k;
        k = (lower == up) ? mn : mx;
      }
    }
    keys[lane] = k;
    wave_sync();
    return;
  }
  for (int size = 2; size <= np; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < np / 2; t += kWave) {
        int i = 2 * t - (t & (stride - 1));
        int j = i + stride;
        bool up = ((i & size) == 0);
        uint64_t a = keys[i], b = keys[j];
        if ((a > b) == up) {
          keys[i] = b;
          keys[j] = a;
        }
      }
      wave_sync();
    }
  }
}

// Order-preserving map of a float to an unsigned key (finite values).
__device__ __forceinline__ uint32_t float_key(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct RowCtx {
  const FillInfo* F;
  const PartInfo* parts;  // this fill's partitions
  const Entry* E;         // this fill's entries
  const int8_t* EW;       // windings
  int32_t y;
  int32_t wx0, wx1;       // window [wx0, wx1)
  int32_t cx0, cx1;       // clip of this row's writes
};

// Adds value v to pixels [a, b) of the window (clipped), via D.
__device__ __forceinline__ void add_run(const RowCtx& C, RowLds& L, int64_t a, int64_t b, int32_t v) {
  int64_t lo = a > C.cx0 ? a : C.cx0;
  int64_t hi = b < C.cx1 ? b : C.cx1;
  if (lo < C.wx0) lo = C.wx0;
  if (hi > C.wx1) hi = C.wx1;
  if (lo >= hi) return;
  atomicAdd(&L.D[lo - C.wx0], v);
  atomicAdd(&L.D[hi - C.wx0], -v);
}

// The walk iterator (paths.nim:1298-1330) over the n sorted hits in L.keys
// (at in the high word, hit position in the low word; windings in L.B by
// position), executed in parallel.  Calls span(prevAt, at) per yielded span.
__device__ __forceinline__ int32_t key_at(uint64_t k) {
  return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
}

// kPacked: the winding is the key's low byte (keys (x, index << 8 | winding))
// instead of L.B[low word].
__device__ __forceinline__ int32_t key_winding(uint64_t k) { return (int32_t)(int8_t)(uint8_t)k; }

// Walk state carried from one 64-hit chunk to the next.
struct WalkCarry {
  int32_t count = 0;      // winding count before the chunk
  int32_t prev = 0;       // prevAt so far (0 before any event)
  bool consumed = false;  // the chunk's first hit was consumed by a skip
};

// One chunk of the walk: lane i holds a sorted hit (at, w) when valid and,
// when hasNext, the hit after it (nextAt, nextW), which for the last lane is
// the next chunk's first hit.  Yields pixie's spans in parallel: winding
// counts by prefix sum, the zero-width skip of a cancelling pair by ballots
// (a skip consumes the next hit, so skips are taken left to right), the
// NonZero shortcut per lane, and prevAt as an exclusive prefix max of the
// event positions.
template <class SpanFn>
__device__ __forceinline__ void walk_chunk(int32_t at, int32_t w, bool valid, bool hasNext, int32_t nextAt,
                                           int32_t nextW, int32_t rule, WalkCarry& c, SpanFn& span) {
  const int lane = lane_id();
  int32_t tot;
  const int32_t countBefore = c.count + wave_excl_scan(w, &tot);
  const bool filling = should_fill(rule, countBefore);
  const bool canSkip = valid && at > 0 && filling && hasNext && nextAt == at && w + nextW == 0;
  uint64_t cm = ballot(canSkip);
  uint64_t skipMask = 0, consMask = c.consumed ? 1ull : 0ull;
  bool carryOut = false;
  while (cm) {
    const int b = __builtin_ctzll(cm);
    cm &= cm - 1;
    if (!((consMask >> b) & 1ull)) {
      skipMask |= 1ull << b;
      if (b + 1 < kWave) consMask |= 1ull << (b + 1);
      else carryOut = true;
    }
  }
  const bool consumed = (consMask >> lane) & 1ull;
  const bool skipStart = (skipMask >> lane) & 1ull;
  // 0 = none, 1 = start (prevAt = at), 2 = end (yield, prevAt = at)
  int ev = 0;
  if (valid && !consumed && !skipStart && at > 0) {
    if (filling) {
      const bool cont = (rule == kNonZero) && hasNext && (countBefore + w != 0);
      ev = cont ? 0 : 2;
    } else {
      ev = 1;
    }
  }
  const int32_t prevHere = wave_excl_max(ev ? at : INT32_MIN);
  const int32_t prevAt = prevHere > c.prev ? prevHere : c.prev;
  if (ev == 2) span(prevAt, at);
  const int32_t lastEv = wave_max(ev ? at : INT32_MIN);
  if (lastEv > c.prev) c.prev = lastEv;
  c.count += tot;
  c.consumed = carryOut;
}

template <bool kPacked = false, class SpanFn>
__device__ void wave_walk(RowLds& L, int n, int32_t rule, SpanFn& span) {
  const int lane = lane_id();
  WalkCarry c;
  for (int base = 0; base < n; base += kWave) {
    const int i = base + lane;
    const bool valid = i < n;
    const uint64_t k = valid ? L.keys[i] : 0ull;
    const int32_t at = valid ? key_at(k) : 0;
    const int32_t w = valid ? (kPacked ? key_winding(k) : L.B[(uint32_t)k]) : 0;
    const bool hasNext = valid && i < n - 1;
    const uint64_t kn = hasNext ? L.keys[i + 1] : 0ull;
    const int32_t nextAt = hasNext ? key_at(kn) : 0;
    const int32_t nextW = hasNext ? (kPacked ? key_winding(kn) : L.B[(uint32_t)kn]) : 0;
    walk_chunk(at, w, valid, hasNext, nextAt, nextW, rule, c, span);
  }
}

// wave_walk<true> for at most 64 sorted hits held in registers (hit i in
// lane i, padding keys ~0 beyond n), the next hit taken from the next lane.
template <class SpanFn>
__device__ __forceinline__ void wave_walk_reg(uint64_t k, int n, int32_t rule, SpanFn& span) {
  const int lane = lane_id();
  const bool valid = lane < n;
  const bool hasNext = valid && lane < n - 1;
  const uint64_t kn = __shfl(k, lane + 1 < kWave ? lane + 1 : lane, kWave);
  WalkCarry c;
  walk_chunk(valid ? key_at(k) : 0, valid ? key_winding(k) : 0, valid, hasNext,
             hasNext ? key_at(kn) : 0, hasNext ? key_winding(kn) : 0, rule, c, span);
}

// One odd-even transposition step over the wave's keys (pairs (2m, 2m+1), or
// with odd (2m+1, 2m+2)); the lower lane of a pair keeps the smaller key.
// Returns whether this lane took its partner's key.
__device__ __forceinline__ bool wave_transpose(uint64_t& k, bool odd) {
  const int lane = lane_id();
  bool lower, has;
  int src;
  if (!odd) {
    lower = (lane & 1) == 0;
    has = true;
    src = lane ^ 1;
  } else {
    lower = (lane & 1) == 1;
    has = lower ? lane + 1 < kWave : lane > 0;
    src = lower ? (lane + 1 < kWave ? lane + 1 : lane) : (lane > 0 ? lane - 1 : lane);
  }
  const uint64_t o = __shfl(k, src, kWave);
  const bool take = has && (lower ? o < k : o > k);
  if (take) k = o;
  return take;
}

// Coverage contributions of one span (computeCoverage 1390-1431).
struct CovSpan {
  const RowCtx* C;
  RowLds* L;
  int32_t sc;        // sampleCoverage
  int32_t covLo, covHi;  // coverage buffer span [startX, startX + pathWidth)
  __device__ void add_px(int64_t x, int32_t v) {
    if (x < covLo || x >= covHi) return;
    if (x < C->wx0 || x >= C->wx1) return;
    atomicAdd(&L->D[x - C->wx0], v);
    atomicAdd(&L->D[x - C->wx0 + 1], -v);
  }
  __device__ void operator()(int32_t prevAt, int32_t at) {
    int64_t fillStart = fx_int(prevAt);
    bool pixelCrossed = fx_int(at) != fx_int(prevAt);
    int32_t leftCover = pixelCrossed ? fx_trunc(prevAt) + 256 - prevAt : at - prevAt;
    if (leftCover != 0) {
      fillStart++;
      add_px(fx_int(prevAt), (int32_t)(uint8_t)fx_int(leftCover * sc));
    }
    if (pixelCrossed) {
      int32_t rightCover = at - fx_trunc(at);
      if (rightCover > 0) add_px(fx_int(at), (int32_t)(uint8_t)fx_int(rightCover * sc));
    }
    int64_t fillEnd = fx_int(at);
    if (fillEnd > fillStart) {
      int64_t lo = fillStart > covLo ? fillStart : covLo;
      int64_t hi = fillEnd < covHi ? fillEnd : covHi;
      if (lo < C->wx0) lo = C->wx0;
      if (hi > C->wx1) hi = C->wx1;
      if (lo < hi) {
        atomicAdd(&L->D[lo - C->wx0], sc);
        atomicAdd(&L->D[hi - C->wx0], -sc);
      }
    }
  }
};

// Fill runs of a walkInteger span (fillHits 1540-1591).
struct FillSpan {
  const RowCtx* C;
  RowLds* L;
  int32_t code;   // value added over the span
  int64_t minX;   // spans are cut below minX (mask mode: left of startX is cleared)
  __device__ void operator()(int32_t prevAt, int32_t at) {
    int64_t s = fx_int(prevAt), e = fx_int(at);
    if (e - s <= 0) return;
    if (s < minX) s = minX;
    add_run(*C, *L, s, e, code);
  }
};

// The row's selected entries.  When the selection fits in one wave (the
// common case), lane i keeps selection i in registers for all sample lines;
// otherwise entries are re-read through L.A.
struct SelCache {
  bool inReg;
  Entry e;
  int32_t w;
};

// Collects the hits of one sample line (paths.nim:1373-1388) from the
// selection L.A[0..ne) into L.keys sorted stably by x; windings go to L.B by
// hit position.  Returns the number of hits.
__device__ __forceinline__ int row_hits(const RowCtx& C, RowLds& L, int ne, float yLine, SelCache S) {
  int lane = lane_id();
  const float Wf = (float)C.F->canvasW;
  int nh = 0;
  for (int base = 0; base < ne; base += kWave) {
    int i = base + lane;
    bool hit = false;
    int32_t at = 0, w = 0;
    if (i < ne) {
      Entry e;
      int32_t ew;
      if (S.inReg) {
        e = S.e;
        ew = S.w;
      } else {
        int k = L.A[i];
        e = C.E[k];
        ew = C.EW[k];
      }
      if (e.aty <= yLine && e.toy >= yLine) {
        float x = e.m == 0.0f ? e.b : (yLine - e.b) / e.m;
        at = fixed32(nim_minf(x, Wf));
        w = ew;
        hit = true;
      }
    }
    uint64_t m = ballot(hit);
    int pos = nh + __popcll(m & ((1ull << lane) - 1ull));
    if (hit) {
      L.keys[pos] = ((uint64_t)((uint32_t)at ^ 0x80000000u) << 32) | (uint32_t)pos;
      L.B[pos] = w;
    }
    nh += __popcll(m);
  }
  wave_sync();
  if (nh > 1) wave_sort_keys(L.keys, nh, L.cap);
  return nh;
}

// The 5-sample coverage of a row whose selection holds at most SEG entries
// (computeCoverage 1350-1431 with walk 1298-1334), 64 / SEG sample lines at a
// time, one per SEG-lane segment of the wave: two passes for SEG = 16, three
// for SEG = 32, instead of five.  Lane SEG*g + i works on entry i for sample
// line g of the pass.  A segment holds one whole sample line, so the walk
// needs no carries between chunks.  S holds selection i in lane i.
template <int SEG>
__device__ __forceinline__ void row_coverage_small(const RowCtx& C, RowLds& L, int ne, SelCache S, float y0,
                                   float y1, float y2, float y3, float y4, int32_t rule,
                                   CovSpan& span) {
  constexpr int kPer = kWave / SEG;                 // sample lines per pass
  constexpr int kPasses = (5 + kPer - 1) / kPer;
  const int lane = lane_id();
  const int grp = lane / SEG, i = lane & (SEG - 1);
  const float Wf = (float)C.F->canvasW;
  // entry i of the selection in every segment
  Entry e;
  e.aty = __shfl(S.e.aty, i, kWave);
  e.toy = __shfl(S.e.toy, i, kWave);
  e.m = __shfl(S.e.m, i, kWave);
  e.b = __shfl(S.e.b, i, kWave);
  const int32_t ew = __shfl(S.w, i, kWave);
  uint64_t* sorted = (uint64_t*)L.tx0;  // 64 sorted keys (tx0/tx1 are free here)
  const int base = SEG * grp;
  for (int pass = 0; pass < kPasses; pass++) {
    const int sample = pass * kPer + grp;
    // the sample line of this segment, selected from registers
    const float yl = sample == 0 ? y0 : sample == 1 ? y1 : sample == 2 ? y2 : sample == 3 ? y3 : y4;
    // hits (1373-1388), compacted per segment in entry order
    bool hit = false;
    int32_t at = 0;
    if (sample < 5 && i < ne && e.aty <= yl && e.toy >= yl) {
      float x = e.m == 0.0f ? e.b : (yl - e.b) / e.m;
      at = fixed32(nim_minf(x, Wf));
      hit = true;
    }
    const uint64_t m = ballot(hit);
    const uint64_t segMask = (m >> base) & (SEG == 32 ? 0xffffffffull : 0xffffull);
    const int nh = __popcll(segMask);
    const int pos = __popcll(segMask & ((1ull << i) - 1ull));
    // (x, position << 8 | winding): the winding rides in the low byte
    const uint64_t key = ((uint64_t)((uint32_t)at ^ 0x80000000u) << 32) |
                         ((uint32_t)pos << 8) | (uint32_t)(uint8_t)ew;
    if (hit) L.keys[base + pos] = key;
    wave_sync();
    // stable sort by (x, position): rank among the segment's hits
    if (hit) {
      int rank = 0;
      for (int j = 0; j < nh; j++) rank += L.keys[base + j] < key ? 1 : 0;
      sorted[base + rank] = key;
    }
    wave_sync();
    // walk: lane SEG*g + p holds the p-th hit of segment g
    const bool valid = i < nh;
    const uint64_t k = valid ? sorted[base + i] : 0ull;
    const int32_t hat = valid ? key_at(k) : 0;
    const int32_t w = valid ? key_winding(k) : 0;
    const int32_t countBefore = seg_incl_sum<SEG>(w) - w;
    const bool filling = should_fill(rule, countBefore);
    const bool hasNext = valid && i < nh - 1;
    const uint64_t kn = hasNext ? sorted[base + i + 1] : 0ull;
    const int32_t nextAt = hasNext ? key_at(kn) : 0;
    const int32_t nextW = hasNext ? key_winding(kn) : 0;
    const bool canSkip = valid && hat > 0 && filling && hasNext && nextAt == hat && w + nextW == 0;
    uint64_t cm = ballot(canSkip);
    uint64_t skipMask = 0, consMask = 0;
    while (cm) {
      const int b = __builtin_ctzll(cm);
      cm &= cm - 1;
      if (!((consMask >> b) & 1ull)) {
        skipMask |= 1ull << b;
        if (((b + 1) & (SEG - 1)) != 0) consMask |= 1ull << (b + 1);  // stays in its segment
      }
    }
    const bool consumed = (consMask >> lane) & 1ull;
    const bool skipStart = (skipMask >> lane) & 1ull;
    int ev = 0;  // 0 none, 1 start, 2 end (yield)
    if (valid && !consumed && !skipStart && hat > 0) {
      if (filling) {
        const bool cont = (rule == kNonZero) && hasNext && (countBefore + w != 0);
        ev = cont ? 0 : 2;
      } else {
        ev = 1;
      }
    }
    int32_t prevAt = seg_excl_max<SEG>(ev ? hat : INT32_MIN);
    if (prevAt < 0) prevAt = 0;  // prevAt starts at 0 (walk 1300)
    if (ev == 2) span(prevAt, hat);
    wave_sync();
  }
}

// Finishes sorting n distinct keys that are nearly in order: odd-even
// transposition passes until one swaps nothing, and a full sort once
// kFixPasses passes have not been enough.
constexpr int kFixPasses = 8;     // keys in registers
constexpr int kFixPassesMem = 16;  // keys in LDS or global memory (more than 64 hits)
__device__ void wave_fix_keys(uint64_t* keys, int n, int cap) {
  const int lane = lane_id();
  for (int pass = 0;; pass++) {
    bool swapped = false;
    for (int par = 0; par < 2; par++) {
      for (int i = 2 * lane + par; i + 1 < n; i += 2 * kWave) {
        const uint64_t a = keys[i], b = keys[i + 1];
        if (a > b) {
          keys[i] = b;
          keys[i + 1] = a;
          swapped = true;
        }
      }
      wave_sync();
    }
    if (ballot(swapped) == 0) return;
    if (pass + 1 >= kFixPassesMem) {
      wave_sort_keys(keys, n, cap);
      return;
    }
  }
}

// The 5-sample coverage of a row with any number of selected entries
// (computeCoverage 1350-1431).  pixie sorts each sample line's hits from
// scratch (sort 1373-1388).  Here the selection is sorted once, by x where
// each entry is nearest the row's middle; every sample line then collects
// its hits in that order, which leaves them sorted except where entries cross
// inside the row, and wave_fix_keys finishes the sort.  The hit keys carry
// the selection index as tie-break, which orders equal x like pixie's
// insertion sort over hits collected in selection order, and the winding in
// their low byte, so the walk reads it from the key.
__device__ __forceinline__ void row_coverage_presorted(const RowCtx& C, RowLds& L, int ne, SelCache S,
                                       float yLine, float offset, int32_t rule, CovSpan& span) {
  const int lane = lane_id();
  const float Wf = (float)C.F->canvasW;
  const float yMid = (float)C.y + 0.5f;
  // By presorted position: the selection index with the winding in the low
  // byte (the hit key's low word), and the entry index.
  int32_t* order = (int32_t*)L.tx0;
  int32_t* entry = (int32_t*)L.tx1;
  for (int i = lane; i < ne; i += kWave) {
    Entry e;
    int32_t w;
    if (S.inReg) {
      e = S.e;
      w = S.w;
    } else {
      const int k = L.A[i];
      e = C.E[k];
      w = C.EW[k];
    }
    const float yc = yMid < e.aty ? e.aty : (yMid > e.toy ? e.toy : yMid);
    L.keys[i] = ((uint64_t)float_key(solve_x(e.m, e.b, yc)) << 32) | (uint32_t)i;
    L.B[i] = w;
  }
  wave_sync();
  if (ne > 1) wave_sort_keys(L.keys, ne, L.cap);
  for (int j = lane; j < ne; j += kWave) {
    const int s = (int)(uint32_t)L.keys[j];
    order[j] = (s << 8) | (int32_t)(uint8_t)L.B[s];
    if (!S.inReg) entry[j] = L.A[s];
  }
  wave_sync();
  // a selection held in registers moves to its presorted lanes
  Entry er = S.e;
  int32_t sr = 0;
  if (S.inReg) {
    sr = lane < ne ? order[lane] : 0;
    const int src = sr >> 8;
    er.aty = __shfl(S.e.aty, src, kWave);
    er.toy = __shfl(S.e.toy, src, kWave);
    er.m = __shfl(S.e.m, src, kWave);
    er.b = __shfl(S.e.b, src, kWave);
  }
  for (int m = 0; m < 5; m++) {
    yLine += offset;
    int nh = 0;
    for (int base = 0; base < ne; base += kWave) {
      const int j = base + lane;
      bool hit = false;
      int32_t at = 0, s = 0;
      if (j < ne) {
        Entry e;
        if (S.inReg) {
          e = er;
          s = sr;
        } else {
          e = C.E[entry[j]];
          s = order[j];
        }
        if (e.aty <= yLine && e.toy >= yLine) {
          const float x = e.m == 0.0f ? e.b : (yLine - e.b) / e.m;
          at = fixed32(nim_minf(x, Wf));
          hit = true;
        }
      }
      const uint64_t mk = ballot(hit);
      const int pos = nh + __popcll(mk & ((1ull << lane) - 1ull));
      if (hit) L.keys[pos] = ((uint64_t)((uint32_t)at ^ 0x80000000u) << 32) | (uint32_t)s;
      nh += __popcll(mk);
    }
    wave_sync();
    if (nh <= kWave) {
      // the hits fit the wave: finish the sort and walk in registers
      uint64_t k = lane < nh ? L.keys[lane] : ~0ull;  // padding sorts last
      bool done = false;
      for (int t = 0; t < kFixPasses; t++) {
        bool sw = wave_transpose(k, false);
        sw = wave_transpose(k, true) || sw;
        if (ballot(sw) == 0) {
          done = true;
          break;
        }
      }
      if (!done) {
        wave_sync();
        if (lane < nh) L.keys[lane] = k;
        wave_sync();
        wave_sort_keys(L.keys, nh, L.cap);
        k = lane < nh ? L.keys[lane] : ~0ull;
      }
      wave_walk_reg(k, nh, rule, span);
    } else {
      wave_fix_keys(L.keys, nh, L.cap);
      wave_walk<true>(L, nh, rule, span);
    }
    wave_sync();
  }
}

// Bitonic sort of the 64 * NS keys held in slots k[0..NS) (key s * 64 + lane
// in slot s of that lane).  Compare-exchanges at strides below 64 move keys
// between lanes; strides of 64 and more pair two slots of the same lane.
constexpr int kWideSlots = 4;
template <int SS>
__device__ __forceinline__ void slots_exchange(uint64_t (&k)[kWideSlots], int size) {
#pragma unroll
  for (int s = 0; s < kWideSlots; s++) {
    if (s & SS) continue;
    const bool up = ((s * kWave) & size) == 0;
    const uint64_t a = k[s], b = k[s | SS];
    const uint64_t mn = a < b ? a : b, mx = a < b ? b : a;
    k[s] = up ? mn : mx;
    k[s | SS] = up ? mx : mn;
  }
}
template <int NS>
__device__ __forceinline__ void slots_sort(uint64_t (&k)[kWideSlots]) {
  const int lane = lane_id();
#pragma unroll
  for (int size = 2; size <= kWave * NS; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 2 * kWave) {
        slots_exchange<2>(k, size);
      } else if (stride >= kWave) {
        slots_exchange<1>(k, size);
      } else {
        const bool lower = (lane & stride) == 0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
          const uint64_t o = __shfl_xor(k[s], stride, kWave);
          const bool up = (((s * kWave) | lane) & size) == 0;
          const uint64_t mn = k[s] < o ? k[s] : o, mx = k[s] < o ? o : k[s];
          k[s] = lower == up ? mn : mx;
        }
      }
    }
  }
}

// The 5-sample coverage of an anti-aliased row that selects 65 to
// kWideMax entries (computeCoverage 1350-1431 with walk 1298-1334), with no
// per-row arrays in memory.  A selection of at most 256 entries is held in
// registers (selection j in slot j / 64 of lane j % 64); a larger one is
// read from the band for every sample line (C.E / C.EW start at the band,
// count entries; a hit implies the selection test, which is repeated so
// that the two agree at any y).  Either way each line's hits are compacted
// in selection order, their keys (x, hit position << 8 | winding) staged in
// L.wide, and sorted, which gives pixie's stable insertion order: up to 256
// hits in four 64-key register slots, walked slot by slot with the walk's
// carries; more in the stage itself (LDS), whose kWideMax keys hold the
// sort's padding.  The selection indices are in L.A when ne <= L.cap.
constexpr int kWideMax = 8 * kWave;
__device__ __forceinline__ int wide_limit(const RowLds& L) {
  return L.wideCap >= kWideMax ? kWideMax : (L.wideCap < kWideSlots * kWave ? L.wideCap : kWideSlots * kWave);
}
__device__ __forceinline__ void row_coverage_wide(const RowCtx& C, RowLds& L, int count, int ne, float scanTop,
                                                  float scanBottom, float yLine, float offset, int32_t rule,
                                                  CovSpan& span) {
  const int lane = lane_id();
  const float Wf = (float)C.F->canvasW;
  uint64_t* stage = L.wide;
  const bool inReg = ne <= kWideSlots * kWave;
  Entry e[kWideSlots];
  int32_t w[kWideSlots];
#pragma unroll
  for (int s = 0; s < kWideSlots; s++) {
    e[s] = Entry{0.0f, 0.0f, 0.0f, 0.0f};
    w[s] = 0;
  }
  if (inReg) {
    const int32_t* selIdx = L.A;
    if (ne > L.cap) {  // select again, into the stage
      int32_t* idx = (int32_t*)stage;
      int n = 0;
      for (int base = 0; base < count; base += kWave) {
        const int i = base + lane;
        bool sel = false;
        if (i < count) {
          const Entry ei = C.E[i];
          sel = !(ei.toy <= scanTop || ei.aty >= scanBottom);
        }
        const uint64_t m = ballot(sel);
        if (sel) idx[n + __popcll(m & ((1ull << lane) - 1ull))] = i;
        n += __popcll(m);
      }
      wave_sync();
      selIdx = idx;
    }
#pragma unroll
    for (int s = 0; s < kWideSlots; s++) {
      const int j = s * kWave + lane;
      if (j < ne) {
        const int k = selIdx[j];
        e[s] = C.E[k];
        w[s] = C.EW[k];
      }
    }
    wave_sync();  // the stage now takes the hit keys
  }
  for (int m = 0; m < 5; m++) {
    yLine += offset;
    int nh = 0;
    if (inReg) {
#pragma unroll
      for (int s = 0; s < kWideSlots; s++) {
        bool hit = false;
        int32_t at = 0;
        if (s * kWave + lane < ne && e[s].aty <= yLine && e[s].toy >= yLine) {
          const float x = e[s].m == 0.0f ? e[s].b : (yLine - e[s].b) / e[s].m;
          at = fixed32(nim_minf(x, Wf));
          hit = true;
        }
        const uint64_t mk = ballot(hit);
        const int pos = nh + __popcll(mk & ((1ull << lane) - 1ull));
        if (hit)
          stage[pos] = ((uint64_t)((uint32_t)at ^ 0x80000000u) << 32) | ((uint32_t)pos << 8) |
                       (uint32_t)(uint8_t)w[s];
        nh += __popcll(mk);
      }
    } else {
#pragma unroll 2
      for (int base = 0; base < count; base += kWave) {
        const int i = base + lane;
        bool hit = false;
        int32_t at = 0, wi = 0;
        if (i < count) {
          const Entry ei = C.E[i];
          if (!(ei.toy <= scanTop || ei.aty >= scanBottom) && ei.aty <= yLine && ei.toy >= yLine) {
            const float x = ei.m == 0.0f ? ei.b : (yLine - ei.b) / ei.m;
            at = fixed32(nim_minf(x, Wf));
            wi = C.EW[i];
            hit = true;
          }
        }
        const uint64_t mk = ballot(hit);
        const int pos = nh + __popcll(mk & ((1ull << lane) - 1ull));
        if (hit)
          stage[pos] = ((uint64_t)((uint32_t)at ^ 0x80000000u) << 32) | ((uint32_t)pos << 8) |
                       (uint32_t)(uint8_t)wi;
        nh += __popcll(mk);
      }
    }
    wave_sync();
    if (nh > kWideSlots * kWave) {
      RowLds Ls = L;
      Ls.keys = stage;
      wave_sort_keys(stage, nh, kWideMax);
      wave_walk<true>(Ls, nh, rule, span);
      wave_sync();
      continue;
    }
    uint64_t k[kWideSlots];
#pragma unroll
    for (int s = 0; s < kWideSlots; s++) {
      const int j = s * kWave + lane;
      k[s] = j < nh ? stage[j] : ~0ull;  // padding sorts last
    }
    if (nh <= kWave)
      slots_sort<1>(k);
    else if (nh <= 2 * kWave)
      slots_sort<2>(k);
    else
      slots_sort<4>(k);
    WalkCarry c;
#pragma unroll
    for (int s = 0; s < kWideSlots; s++) {
      if (s * kWave >= nh) break;
      const int j = s * kWave + lane;
      const bool valid = j < nh, hasNext = j < nh - 1;
      uint64_t kn = __shfl(k[s], lane + 1 < kWave ? lane + 1 : lane, kWave);
      if (s + 1 < kWideSlots) {
        const uint64_t first = __shfl(k[s + 1], 0, kWave);
        if (lane == kWave - 1) kn = first;
      }
      walk_chunk(valid ? key_at(k[s]) : 0, valid ? key_winding(k[s]) : 0, valid, hasNext,
                 hasNext ? key_at(kn) : 0, hasNext ? key_winding(kn) : 0, rule, c, span);
    }
    wave_sync();
  }
}

// Entries row C.y selects (row_program's selection), so that a caller can
// place the row's arrays before running it: the arrays hold the selection,
// which is often much smaller than the band.
__device__ __forceinline__ int row_selection_count(const RowCtx& C) {
  const FillInfo& F = *C.F;
  const int32_t y = C.y;
  if (y < F.startY || y >= F.pathHeight) return 0;
  int32_t pi = (y - F.startY) / F.bandH;
  if (pi > F.nParts - 1) pi = F.nParts - 1;
  const int count = __builtin_amdgcn_readfirstlane(C.parts[pi].count);
  if (__builtin_amdgcn_readfirstlane(C.parts[pi].flags) & 2) return 2;
  const Entry* E = C.E + __builtin_amdgcn_readfirstlane(C.parts[pi].entryBase);
  const float scanTop = (float)y, scanBottom = (float)(y + 1);
  int ne = 0;
#pragma unroll 2
  for (int base = 0; base < count; base += kWave) {
    const int i = base + lane_id();
    bool s = false;
    if (i < count) {
      const Entry e = E[i];
      s = !(e.toy <= scanTop || e.aty >= scanBottom);
    }
    ne += __popcll(ballot(s));
  }
  return ne;
}

// Runs the row program for (fill, row y, window) and leaves the row's codes
// in L.D (and L.AX).  Returns the row kind; kRowNone means nothing to do and
// -1 means the selection holds more entries than L.cap (nothing was written).
__device__ int32_t row_program(const RowCtx& C, RowLds& L) {
  const FillInfo& F = *C.F;
  const int lane = lane_id();
  const int32_t W = F.canvasW;
  const int32_t y = C.y;
  const bool mask = F.mode == kMask;

  if (y < F.startY || y >= F.pathHeight) {
    if (mask) {  // rows outside the path are cleared (paths.nim:1910-1912)
      if (lane == 0) add_run(C, L, 0, W, kClear);
      wave_sync();
      return kRowOps;
    }
    return kRowNone;
  }
  int32_t pi = (y - F.startY) / F.bandH;
  if (pi > F.nParts - 1) pi = F.nParts - 1;
  PartInfo P = C.parts[pi];  // the same in every lane: scalar registers
  P.entryBase = __builtin_amdgcn_readfirstlane(P.entryBase);
  P.count = __builtin_amdgcn_readfirstlane(P.count);
  P.flags = __builtin_amdgcn_readfirstlane(P.flags);
  const bool aa = P.flags & 1, two = P.flags & 2;
  RowCtx Cp = C;
  Cp.E = C.E + P.entryBase;
  Cp.EW = C.EW + P.entryBase;
  const Entry* E = Cp.E;
  const int8_t* EW = Cp.EW;

  if (two && !aa) {  // paths.nim:1644-1668
    if (mask && lane == 0) add_run(C, L, 0, W, kClear);  // default for mask rows
    if (lane == 0) {
      int64_t left = nim_f2i(E[0].b), right = nim_f2i(E[1].b);
      int64_t minX = left < 0 ? 0 : (left > W ? W : left);
      int64_t maxX = right < 0 ? 0 : (right > W ? W : right);
      if (maxX > minX) add_run(C, L, minX, maxX, mask ? kFill - kClear : kFill);
    }
    wave_sync();
    return kRowOps;
  }

  const float scanTop = (float)y, scanBottom = (float)(y + 1);
  int ne = 0;
  bool allSpan = true;
  if (two) {
    if (lane == 0) {
      L.A[0] = 0;
      L.A[1] = 1;
    }
    ne = 2;
  } else {
#pragma unroll 2
    for (int base = 0; base < P.count; base += kWave) {
      int i = base + lane;
      bool s = false, notSpan = false;
      if (i < P.count) {
        Entry e = E[i];
        s = !(e.toy <= scanTop || e.aty >= scanBottom);
        notSpan = s && (e.aty > scanTop || e.toy < scanBottom);
      }
      uint64_t m = ballot(s);
      int pos = ne + __popcll(m & ((1ull << lane) - 1ull));
      if (s && pos < L.cap) L.A[pos] = i;
      ne += __popcll(m);
      if (ballot(notSpan)) allSpan = false;
    }
  }
  wave_sync();
  // An anti-aliased row of more than 64 selected entries that is not a
  // trapezoid candidate takes the wide path when its selection fits it.
  const bool wide = aa && !two && !(allSpan && (ne % 2) == 0) && ne > kWave && ne <= wide_limit(L);
  // the row's arrays hold the selection (not the whole band); D is untouched
  // so far, so a selection too large for them leaves nothing behind
  if (!wide && ne > L.cap) return -1;
  if (mask && lane == 0) add_run(C, L, 0, W, kClear);  // default for mask rows
  if (wide) {  // computeCoverage 1350-1431 at quality 5
    const float offset = 1.0f / 5.0f;
    const float yLine = (float)y + (offset / 2.0f + 0.00031415926535897936f) - offset;
    const int64_t covLo = F.startX, covHi = (int64_t)F.startX + F.pathWidth;
    if (mask && lane == 0) add_run(C, L, covLo, covHi, -kClear);  // coverage range
    CovSpan cs{&C, &L, 255 / 5, (int32_t)covLo, (int32_t)covHi};
    row_coverage_wide(Cp, L, P.count, ne, scanTop, scanBottom, yLine, offset, F.rule, cs);
    return kRowCoverage;
  }
  const bool inReg = ne <= kWave;
  float sAty = 0.0f, sToy = 0.0f, sM = 0.0f, sB = 0.0f;
  int32_t sW = 0;
  if (inReg && lane < ne) {
    const int k = L.A[lane];
    sAty = E[k].aty;
    sToy = E[k].toy;
    sM = E[k].m;
    sB = E[k].b;
    sW = EW[k];
  }
  const SelCache S{inReg, Entry{sAty, sToy, sM, sB}, sW};

  if (allSpan && (ne % 2) == 0) {  // trapezoid fast path, 1691-1872
    for (int i = lane; i < ne; i += kWave) {
      Entry e = S.e;  // (a conditional copy, not a ?: on the two lvalues: that would keep S in memory)
      if (!S.inReg) e = E[L.A[i]];
      float a = solve_x(e.m, e.b, scanTop), b = solve_x(e.m, e.b, scanBottom);
      L.tx0[i] = a;
      L.tx1[i] = b;
      L.keys[i] = ((uint64_t)float_key(midpoint_x(v2(a, scanTop), v2(b, scanBottom))) << 32) |
                  (uint32_t)i;
    }
    wave_sync();
    if (ne > 1) wave_sort_keys(L.keys, ne, L.cap);
    // the partial coverage areas must not overlap (1720-1728) and the pairs
    // must be simple fills (1731-1745)
    bool bad = false;
    int32_t carry = 0;
    for (int base = 0; base < ne; base += kWave) {
      int i = base + lane;
      bool fail = false;
      int32_t w = 0;
      if (i < ne) {
        int pi0 = (int)(uint32_t)L.keys[i];
        w = EW[L.A[pi0]];
        if (i < ne - 1) {
          int pi1 = (int)(uint32_t)L.keys[i + 1];
          float lmax = nim_maxf(L.tx0[pi0], L.tx1[pi0]);
          float rmin = nim_minf(L.tx0[pi1], L.tx1[pi1]);
          if (nim_f2i(ceilf(lmax)) > nim_f2i(rmin)) fail = true;
        }
      }
      int32_t tot;
      int32_t incl = carry + wave_excl_scan(w, &tot) + w;
      if (i < ne) {
        bool f = should_fill(F.rule, incl);
        if ((i % 2 == 0) != f) fail = true;
      }
      carry += tot;
      if (ballot(fail)) bad = true;
    }
    if (!bad) {
      const int32_t areaCode = mask ? kArea - kClear : kArea;
      for (int p = lane; p < ne / 2; p += kWave) {  // one lane per pair
        int pl = (int)(uint32_t)L.keys[2 * p], pr = (int)(uint32_t)L.keys[2 * p + 1];
        Entry le = E[L.A[pl]], re = E[L.A[pr]];
        float la = L.tx0[pl], lt = L.tx1[pl];
        float ra = L.tx0[pr], rt = L.tx1[pr];
        float leftMaxX = nim_maxf(la, lt);
        float rightMinX = nim_minf(ra, rt);
        int64_t leftCoverEnd = nim_f2i(ceilf(leftMaxX));
        int64_t rightCoverBegin = nim_f2i(truncf(rightMinX));
        {  // left sliver 1772-1809
          bool inverted = la < lt;
          float sliverStart = nim_minf(la, lt);
          float rectStart = leftMaxX;
          float pen = sliverStart, prevPen = pen;
          float penY = inverted ? (float)y : (float)(y + 1);
          float prevPenY = penY;
          int64_t x0 = nim_f2i(sliverStart), x1 = nim_f2i(ceilf(rectStart));
          for (int64_t x = x0; x < x1; x++) {
            prevPen = pen;
            pen = (float)(x + 1);
            float rightRectArea = 0.0f;
            if (pen > rectStart) {
              rightRectArea = pen - rectStart;
              pen = rectStart;
            }
            prevPenY = penY;
            penY = solve_y(le.m, le.b, pen);
            if (x < 0 || x >= W) continue;
            float run = pen - prevPen;
            float triangleArea = 0.5f * run * fabsf(penY - prevPenY);
            float rectArea = inverted ? (prevPenY - (float)y) * run : ((float)(y + 1) - prevPenY) * run;
            float area = triangleArea + rectArea + rightRectArea;
            if (x >= C.cx0 && x < C.cx1 && x >= C.wx0 && x < C.wx1) {
              L.AX[x - C.wx0] = area_factor(area);
              atomicAdd(&L.D[x - C.wx0], areaCode);
              atomicAdd(&L.D[x - C.wx0 + 1], -areaCode);
            }
          }
        }
        {  // right sliver 1811-1847
          bool inverted = ra > rt;
          float rectEnd = rightMinX;
          float sliverEnd = nim_maxf(ra, rt);
          float pen = rectEnd, prevPen = pen;
          float penY = inverted ? (float)(y + 1) : (float)y;
          float prevPenY = penY;
          int64_t x0 = nim_f2i(rectEnd), x1 = nim_f2i(ceilf(sliverEnd));
          for (int64_t x = x0; x < x1; x++) {
            prevPen = pen;
            pen = (float)(x + 1);
            float ap = fabsf(prevPen);
            float leftRectArea = ap - floorf(ap);
            if (pen > sliverEnd) pen = sliverEnd;
            prevPenY = penY;
            penY = solve_y(re.m, re.b, pen);
            if (x < 0 || x >= W) continue;
            float run = pen - prevPen;
            float triangleArea = 0.5f * run * fabsf(penY - prevPenY);
            float rectArea = inverted ? (penY - (float)y) * run : ((float)(y + 1) - penY) * run;
            float area = leftRectArea + triangleArea + rectArea;
            if (x >= C.cx0 && x < C.cx1 && x >= C.wx0 && x < C.wx1) {
              L.AX[x - C.wx0] = area_factor(area);
              atomicAdd(&L.D[x - C.wx0], areaCode);
              atomicAdd(&L.D[x - C.wx0 + 1], -areaCode);
            }
          }
        }
        int64_t fb = leftCoverEnd < 0 ? 0 : (leftCoverEnd > W ? W : leftCoverEnd);
        int64_t fe = rightCoverBegin < 0 ? 0 : (rightCoverBegin > W ? W : rightCoverBegin);
        // interior: fillHits on (fixed32(fb), +1), (fixed32(fe), -1)
        if (fe > fb) add_run(C, L, fb, fe, mask ? kFill - kClear : kFill);
      }
      wave_sync();
      return kRowOps;
    }
    // Not a simple trapezoid row: pixie keeps the midpoint order of the
    // selection for the coverage pass below.
    for (int i = lane; i < ne; i += kWave) L.B[i] = L.A[(int)(uint32_t)L.keys[i]];
    wave_sync();
    for (int i = lane; i < ne; i += kWave) L.A[i] = L.B[i];
    wave_sync();
  }

  // computeCoverage 1350-1431 / fillHits 1898
  const int quality = aa ? 5 : 1;
  const int32_t sampleCoverage = 255 / quality;
  const float offset = 1.0f / (float)quality;
  const float initialOffset = offset / 2.0f + 0.00031415926535897936f;
  float yLine = (float)y + initialOffset - offset;
  const int64_t covLo = F.startX, covHi = (int64_t)F.startX + F.pathWidth;
  if (aa) {
    if (mask && lane == 0) add_run(C, L, covLo, covHi, -kClear);  // coverage range
    CovSpan cs{&C, &L, sampleCoverage, (int32_t)covLo, (int32_t)covHi};
    // (two sample lines per pass for selections of 17-32 entries measured
    // slower than the five single passes, so only 16-entry rows take it)
    if (S.inReg && ne <= 16 && L.cap >= kWave) {
      const float y0 = yLine + offset, y1 = y0 + offset, y2 = y1 + offset, y3 = y2 + offset,
                  y4 = y3 + offset;
      row_coverage_small<16>(Cp, L, ne, S, y0, y1, y2, y3, y4, F.rule, cs);
      return kRowCoverage;
    }
    row_coverage_presorted(Cp, L, ne, S, yLine, offset, F.rule, cs);
    return kRowCoverage;
  }
  yLine += offset;
  int nh = row_hits(Cp, L, ne, yLine, S);
  FillSpan fs{&C, &L, mask ? kFill - kClear : kFill, mask ? (int64_t)F.startX : (int64_t)INT64_MIN};
  wave_walk(L, nh, F.rule, fs);
  wave_sync();
  return kRowOps;
}

// Applies one pixel's code.  Returns the new pixel value.
__device__ __forceinline__ uint32_t apply_code(int32_t kind, int32_t code, uint32_t ax,
                                               uint32_t dst, uint32_t rgbx, int32_t mode,
                                               bool* touched) {
  if (kind == kRowCoverage) {
    if (code >= kClear) {
      *touched = true;
      return 0;
    }
    uint32_t c = (uint32_t)code & 255u;
    bool doit = (mode == kMask) ? (c != 255u) : (c != 0u);
    if (!doit) return dst;
    *touched = true;
    uint32_t src = mul_cov(rgbx, c);
    return mode == kOverwrite ? src : blend(mode, dst, src);
  }
  if (code == 0) return dst;
  *touched = true;
  if (code >= kClear) return 0;
  uint32_t src = code >= kArea ? mul_cov(rgbx, ax) : rgbx;
  return blend(mode, dst, src);
}

// Scans L.D over the window and applies the codes to canvas row `row`.
// Leaves L.D zeroed for the next row.  Canvas rows are W pixels wide.
__device__ uint32_t row_epilogue(const RowCtx& C, RowLds& L, int32_t kind, uint32_t* canvas,
                                 int32_t W, uint32_t rgbx, int32_t mode) {
  const int lane = lane_id();
  const int32_t wlen = C.wx1 - C.wx0;
  uint32_t* rowp = canvas + (size_t)C.y * (size_t)W;
  const bool vec_ok = ((W & 3) == 0) && ((C.wx0 & 3) == 0);
  int32_t carry = 0;
  uint32_t ntouch = 0;
  for (int base = 0; base < wlen; base += 4 * kWave) {
    int i0 = base + 4 * lane;
    int32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    if (i0 < wlen + 1) {  // D has ww + 4 entries; indices up to ww + 3 valid
      d0 = L.D[i0];
      d1 = L.D[i0 + 1];
      d2 = L.D[i0 + 2];
      d3 = L.D[i0 + 3];
      L.D[i0] = 0;
      L.D[i0 + 1] = 0;
      L.D[i0 + 2] = 0;
      L.D[i0 + 3] = 0;
    }
    int32_t c0 = d0, c1 = c0 + d1, c2 = c1 + d2, c3 = c2 + d3;
    int32_t tot;
    int32_t off = carry + wave_excl_scan(c3, &tot);
    carry += tot;
    int32_t code[4] = {off + c0, off + c1, off + c2, off + c3};
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (i0 + k < wlen && code[k] != 0) any = true;
    if (kind == kRowCoverage && mode == kMask) any = true;  // mask coverage rows write every pixel
    if (!any) continue;
    int32_t x0 = C.wx0 + i0;
    uint32_t px[4];
    bool full = vec_ok && (i0 + 3 < wlen);
    if (full) {
      uint4 v = *(const uint4*)(rowp + x0);
      px[0] = v.x; px[1] = v.y; px[2] = v.z; px[3] = v.w;
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) px[k] = (i0 + k < wlen) ? rowp[x0 + k] : 0u;
    }
    bool touched = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (i0 + k < wlen) {
        uint32_t ax = (code[k] >= kArea && code[k] < kClear && kind == kRowOps) ? L.AX[i0 + k] : 0u;
        bool t = false;
        px[k] = apply_code(kind, code[k], ax, px[k], rgbx, mode, &t);
        ntouch += t ? 1u : 0u;
        touched |= t;
      }
    }
    if (!touched) continue;
    if (full) {
      *(uint4*)(rowp + x0) = make_uint4(px[0], px[1], px[2], px[3]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (i0 + k < wlen) rowp[x0 + k] = px[k];
    }
  }
  // clear the tail entries the loop did not visit
  for (int i = lane; i < 4; i += kWave) {
    int idx = ((wlen + 4 * kWave - 1) / (4 * kWave)) * 4 * kWave + i;
    if (idx < L.ww + 4) L.D[idx] = 0;
  }
  wave_sync();
  return ntouch;
}

// Epilogue for an A8 target (masks, glyph atlas): pixel v reads as
// rgbx(v,v,v,v) and the alpha of the result is stored.
__device__ uint32_t row_epilogue_a8(const RowCtx& C, RowLds& L, int32_t kind, uint8_t* canvas,
                                    int32_t W, uint32_t rgbx, int32_t mode) {
  const int lane = lane_id();
  const int32_t wlen = C.wx1 - C.wx0;
  uint8_t* rowp = canvas + (size_t)C.y * (size_t)W;
  const bool vec_ok = ((W & 3) == 0) && ((C.wx0 & 3) == 0);
  int32_t carry = 0;
  uint32_t ntouch = 0;
  for (int base = 0; base < wlen; base += 4 * kWave) {
    int i0 = base + 4 * lane;
    int32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    if (i0 < wlen + 1) {
      d0 = L.D[i0];
      d1 = L.D[i0 + 1];
      d2 = L.D[i0 + 2];
      d3 = L.D[i0 + 3];
      L.D[i0] = 0;
      L.D[i0 + 1] = 0;
      L.D[i0 + 2] = 0;
      L.D[i0 + 3] = 0;
    }
    int32_t c0 = d0, c1 = c0 + d1, c2 = c1 + d2, c3 = c2 + d3;
    int32_t tot;
    int32_t off = carry + wave_excl_scan(c3, &tot);
    carry += tot;
    int32_t code[4] = {off + c0, off + c1, off + c2, off + c3};
    bool any = kind == kRowCoverage && mode == kMask;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (i0 + k < wlen && code[k] != 0) any = true;
    if (!any) continue;
    int32_t x0 = C.wx0 + i0;
    uint32_t px[4];
    bool full = vec_ok && (i0 + 3 < wlen);
    if (full) {
      uint32_t v = *(const uint32_t*)(rowp + x0);
#pragma unroll
      for (int k = 0; k < 4; k++) px[k] = ((v >> (8 * k)) & 255u) * 0x01010101u;
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) px[k] = (i0 + k < wlen) ? rowp[x0 + k] * 0x01010101u : 0u;
    }
    bool touched = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (i0 + k < wlen) {
        uint32_t ax = (code[k] >= kArea && code[k] < kClear && kind == kRowOps) ? L.AX[i0 + k] : 0u;
        bool t = false;
        px[k] = apply_code(kind, code[k], ax, px[k], rgbx, mode, &t);
        ntouch += t ? 1u : 0u;
        touched |= t;
      }
    }
    if (!touched) continue;
    if (full) {
      *(uint32_t*)(rowp + x0) = ch_a(px[0]) | (ch_a(px[1]) << 8) | (ch_a(px[2]) << 16) | (ch_a(px[3]) << 24);
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (i0 + k < wlen) rowp[x0 + k] = (uint8_t)ch_a(px[k]);
    }
  }
  for (int i = lane; i < 4; i += kWave) {
    int idx = ((wlen + 4 * kWave - 1) / (4 * kWave)) * 4 * kWave + i;
    if (idx < L.ww + 4) L.D[idx] = 0;
  }
  wave_sync();
  return ntouch;
}

}  // namespace pxh
