/* oracle_internal.h -- helpers shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY. */
#ifndef ORACLE_INTERNAL_H
#define ORACLE_INTERNAL_H

#include <math.h>
#include <stdint.h>

void orc_set_error(const char* msg);

/* Nim's float -> int conversion in the C backend is a plain C cast; on
 * x86-64 gcc emits cvttss2si with a 64-bit destination, so NaN and values
 * outside the int64 range become INT64_MIN ("integer indefinite"). */
static inline int64_t nim_f2i(float x) {
  if (!(x >= -9.2233720368547758e18f && x < 9.2233720368547758e18f))
    return INT64_MIN;
  return (int64_t)x;
}

/* float -> uint32: the same 64-bit conversion, truncated to 32 bits. */
static inline uint32_t nim_f2u32(float x) { return (uint32_t)(uint64_t)nim_f2i(x); }

/* float -> int32 (Fixed32(f * 256), paths.nim:1268-1269): cvttss2si with a
 * 32-bit destination, NaN / out of range -> INT32_MIN. */
static inline int32_t nim_f2i32(float x) {
  if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
  return (int32_t)x;
}

/* Nim's NaN-aware min/max for floats (system.nim):
 *   min(x, y) = if x <= y or y != y: x else: y
 *   max(x, y) = if y <= x or y != y: x else: y */
static inline float nim_minf(float x, float y) { return (x <= y || y != y) ? x : y; }
static inline float nim_maxf(float x, float y) { return (y <= x || y != y) ? x : y; }

#endif
