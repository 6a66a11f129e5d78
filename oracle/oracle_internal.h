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

/* vmath Mat3 in float32, m[3 * col + row] (a[i, j] = m[3 * i + j]).  These
 * restate vmath >= 3.0.0 (not vendored; parity at the last bit of an entry
 * is pinned only through the masters that use them). */
static inline void orc_m3_apply(const float* m, float x, float y, float* ox, float* oy) {
  *ox = m[0] * x + m[3] * y + m[6];
  *oy = m[1] * x + m[4] * y + m[7];
}
static inline void orc_m3_mul(const float* a, const float* b, float* o) {
  float r[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r[3 * i + j] = b[3 * i + 0] * a[0 * 3 + j] + b[3 * i + 1] * a[1 * 3 + j] + b[3 * i + 2] * a[2 * 3 + j];
  for (int k = 0; k < 9; k++) o[k] = r[k];
}
static inline void orc_m3_inverse(const float* m, float* o) {
#define ORC_A(i, j) m[3 * (i) + (j)]
  const float det = ORC_A(0, 0) * (ORC_A(1, 1) * ORC_A(2, 2) - ORC_A(2, 1) * ORC_A(1, 2)) -
                    ORC_A(0, 1) * (ORC_A(1, 0) * ORC_A(2, 2) - ORC_A(1, 2) * ORC_A(2, 0)) +
                    ORC_A(0, 2) * (ORC_A(1, 0) * ORC_A(2, 1) - ORC_A(1, 1) * ORC_A(2, 0));
  const float inv = 1.0f / det;
  float r[9];
  r[0] = (ORC_A(1, 1) * ORC_A(2, 2) - ORC_A(2, 1) * ORC_A(1, 2)) * inv;
  r[1] = -(ORC_A(0, 1) * ORC_A(2, 2) - ORC_A(0, 2) * ORC_A(2, 1)) * inv;
  r[2] = (ORC_A(0, 1) * ORC_A(1, 2) - ORC_A(0, 2) * ORC_A(1, 1)) * inv;
  r[3] = -(ORC_A(1, 0) * ORC_A(2, 2) - ORC_A(1, 2) * ORC_A(2, 0)) * inv;
  r[4] = (ORC_A(0, 0) * ORC_A(2, 2) - ORC_A(0, 2) * ORC_A(2, 0)) * inv;
  r[5] = -(ORC_A(0, 0) * ORC_A(1, 2) - ORC_A(1, 0) * ORC_A(0, 2)) * inv;
  r[6] = (ORC_A(1, 0) * ORC_A(2, 1) - ORC_A(2, 0) * ORC_A(1, 1)) * inv;
  r[7] = -(ORC_A(0, 0) * ORC_A(2, 1) - ORC_A(2, 0) * ORC_A(0, 1)) * inv;
  r[8] = (ORC_A(0, 0) * ORC_A(1, 1) - ORC_A(1, 0) * ORC_A(0, 1)) * inv;
#undef ORC_A
  for (int k = 0; k < 9; k++) o[k] = r[k];
}

#endif
