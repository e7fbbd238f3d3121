// Morton quantization of one coordinate, mortonTree.js:53-58:
//   q = Math.min(((p - min) * s) | 0, 1023)
// where p and min are f32 values widened to JS Numbers and s = 1024 / size is
// a binary64 value: a binary64 subtraction and product, each rounded to
// nearest even, then ToInt32.  Computed here without binary64 hardware
// arithmetic, and bit-exact with it:
//
//   fast path, f32 only: p >= min, q32 = f32(f32(p - min) * f32(s)).  Each of
//     the three roundings has relative error <= 2^-24, so with x = (p - min) s
//     exactly, |q32 - x| <= 3.0002 * 2^-24 * q32 < 2^-22.4 * q32, and the
//     binary64 result X differs from x by <= 2^-51 * x.  If q32 is at least
//     e = 2^-21 * q32 away from both integers around it (or below 1 - e, since
//     x >= 0), X lies strictly between the same integers as q32 and
//     trunc(X) = floor(q32).  q32 < 2^20 keeps floor exact and far from the
//     ToInt32 wrap.
//   exact path, integer soft-double (softdouble.h): every other finite p --
//     within the margin of an integer (about 1 coordinate in 1000 in a
//     galaxy), below min (a particle left out of the bounding box for a
//     non-finite coordinate) or too far above it -- where ToInt32 may wrap.
// Compiles for the host too, so tests/native/quant_check.cpp can compare it
// with native binary64 arithmetic.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "softdouble.h"

namespace gsq {

SD_HD uint32_t f32_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

// p: coordinate; mn, mn_bits: the bounding-box minimum as f32 and as binary64
// bits (exact widening); s32, s_bits: f32(s) and the binary64 bits of s.
// Returns the low 10 bits that part1by2 uses, min(ToInt32(X), 1023).
SD_HD uint32_t quantize(float p, float mn, float s32, uint64_t mn_bits, uint64_t s_bits,
                        bool* exact_path = nullptr) {
  if (!(p - p == 0.0f)) return 0;  // NaN or +-Inf: (NaN | 0) = 0
  if (p >= mn) {
    const float q = (p - mn) * s32;  // two f32 roundings (no contraction)
    if (q < 1048576.0f) {
      const float k = floorf(q);
      const float e = q * 0x1p-21f;
      if ((k == 0.0f || q - k >= e) && (k + 1.0f) - q >= e) {
        if (exact_path) *exact_path = false;
        const uint32_t v = (uint32_t)k;
        return v < 1023u ? v : 1023u;
      }
    }
  }
  if (exact_path) *exact_path = true;
  const int32_t v = sd::to_int32(sd::mul(sd::sub(sd::from_f32(f32_bits(p)), mn_bits), s_bits));
  return (uint32_t)(v < 1023 ? v : 1023);
}

}  // namespace gsq
