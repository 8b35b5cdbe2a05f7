/*
 * half.h — IEEE binary16 conversions for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * f2h rounds to nearest, ties to even, keeps subnormals, and overflows to
 * infinity from 65520 up: the conversion the GPU's FP16 stores perform in
 * the default rounding mode.  h2f is exact.  Pinned against numpy's float16
 * conversion (every half value, random and boundary floats) by
 * tests/test_oracle_invariants.py.
 */
#ifndef CF_ORACLE_HALF_H
#define CF_ORACLE_HALF_H
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint16_t cfo_f2h1(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
  const uint32_t ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return sign | 0x7e00u;  /* NaN */
  if (ax >= 0x477ff000u) return sign | 0x7c00u; /* >= 65520 (and inf): infinity */
  if (ax < 0x38800000u) {                       /* below 2^-14: subnormal or zero */
    float af;
    memcpy(&af, &ax, 4);
    /* af * 2^24 < 1024 is exact; rintf rounds to nearest even (1024 = the smallest normal) */
    return sign | (uint16_t)rintf(af * 16777216.0f);
  }
  const uint32_t e = (ax >> 23) - 112u; /* rebias 127 -> 15 */
  const uint32_t m = ax & 0x7fffffu;
  uint32_t h = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++; /* a carry moves into the exponent, as it should */
  return sign | (uint16_t)h;
}

static inline float cfo_h2f1(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    const float v = (float)m * (1.0f / 16777216.0f); /* m * 2^-24, exact */
    memcpy(&x, &v, 4);
    x |= sign;
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

#endif
