/*
 * half.c -- IEEE-754 binary16 conversions and correctly rounded ops.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the arithmetic of the cuda_fp16 intrinsics the reference's FP16
 * build uses (src/include/common.hpp:32-45: __floats2half2_rn, __hsub2,
 * __hfma2, __hmin2): each is IEEE binary16 with round-to-nearest-even.
 *
 * +, -, * of two binary16 values are evaluated in binary32 and rounded once
 * to binary16; that double rounding is innocuous because 24 >= 2*11 + 2
 * (Figueroa, "When is double rounding innocuous?", 1995).  The fused form
 * (a*b + c, one rounding) is evaluated exactly in 128-bit fixed point.
 *
 * Parity pinned by: tests/test_oracle_half.py compares orc_f32_to_f16 with
 * numpy's float16 cast (round-half-even) over every binary16 boundary and a
 * large random sample of float32 bit patterns, orc_f16_to_f32 over all 65536
 * encodings, and the ops against numpy float16 arithmetic.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

uint16_t orc_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) { /* inf or NaN */
    if (ax == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x1ffu));
  }
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 */
  if (ax < 0x38800000u) {                                     /* < 2^-14 */
    if (ax <= 0x33000000u) return (uint16_t)sign;             /* <= 2^-25 */
    uint32_t e = ax >> 23;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    int shift = 126 - (int)e; /* value in units of 2^-24 is m >> shift */
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (ax >> 23) - 127u + 15u;
  uint32_t m = ax & 0x7fffffu;
  uint32_t q = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
  uint32_t m = (uint32_t)h & 0x3ffu;
  uint32_t x;
  float f;
  if (e == 0) {
    f = (float)m * 5.9604644775390625e-08f; /* m * 2^-24, exact */
    memcpy(&x, &f, 4);
    x |= sign;
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  memcpy(&f, &x, 4);
  return f;
}

void orc_f32_to_f16_array(const float* in, uint16_t* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = orc_f32_to_f16(in[i]);
}
void orc_f16_to_f32_array(const uint16_t* in, float* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = orc_f16_to_f32(in[i]);
}

uint16_t orc_h_add(uint16_t a, uint16_t b) {
  return orc_f32_to_f16(orc_f16_to_f32(a) + orc_f16_to_f32(b));
}
uint16_t orc_h_sub(uint16_t a, uint16_t b) {
  return orc_f32_to_f16(orc_f16_to_f32(a) - orc_f16_to_f32(b));
}
uint16_t orc_h_mul(uint16_t a, uint16_t b) {
  return orc_f32_to_f16(orc_f16_to_f32(a) * orc_f16_to_f32(b));
}

/* Finite binary16 value as a signed integer multiple of 2^-24. */
static int64_t h_fixed24(uint16_t h) {
  int64_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
  int64_t v = (e == 0) ? m : ((m | 0x400) << (e - 1));
  return (h & 0x8000) ? -v : v;
}

/* Round a signed multiple of 2^-48 to binary16 (nearest even). */
static uint16_t round_fixed48_to_f16(__int128 v) {
  uint16_t sign = 0;
  if (v < 0) { sign = 0x8000; v = -v; }
  if (v == 0) return sign;
  /* binary16 quantum for exponent field e (>=1): 2^(e-25) = 2^(e+23) units;
   * subnormals share the e=1 quantum 2^-24 = 2^24 units. */
  int msb = 127;
  while (!((v >> msb) & 1)) msb--;
  /* value = v * 2^-48; normal exponent field e has leading bit at 48-15+e-1+... */
  int e = msb - 48 + 15; /* unbiased exponent msb-48, field = that + 15 */
  if (e < 1) e = 1;      /* subnormal range uses the smallest quantum */
  int qshift = e + 23;   /* quantum is 2^qshift units of 2^-48 */
  __int128 q = v >> qshift;
  __int128 rem = v - (q << qshift);
  __int128 halfway = ((__int128)1) << (qshift - 1);
  if (rem > halfway || (rem == halfway && (q & 1))) q++;
  /* q counts quanta; rebuild encoding: for e>=1 normal numbers q in [1024,2048) */
  if (q >= 2048) { q >>= 1; e++; }
  if (e >= 31) return (uint16_t)(sign | 0x7c00);
  if (q < 1024) return (uint16_t)(sign | (uint16_t)q); /* subnormal (e==1 range) */
  return (uint16_t)(sign | (uint16_t)(e << 10) | (uint16_t)(q - 1024));
}

uint16_t orc_h_fma(uint16_t a, uint16_t b, uint16_t c) {
  if (((a & 0x7c00) == 0x7c00) || ((b & 0x7c00) == 0x7c00) ||
      ((c & 0x7c00) == 0x7c00)) /* inf/NaN operand: binary32 gives the answer */
    return orc_f32_to_f16(orc_f16_to_f32(a) * orc_f16_to_f32(b) + orc_f16_to_f32(c));
  __int128 p = (__int128)h_fixed24(a) * (__int128)h_fixed24(b); /* 2^-48 units */
  __int128 s = p + (((__int128)h_fixed24(c)) << 24);
  return round_fixed48_to_f16(s);
}
