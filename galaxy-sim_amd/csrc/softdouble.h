// Integer-only IEEE-754 binary64 arithmetic (round to nearest, ties to even).
//
// The Morton quantization of mortonTree.js:53-58 runs in JS Numbers (IEEE
// double): size = max(dx,dy,dz)*1.0001 + 1e-9, s = 1024/size and
// q = ToInt32((p - min) * s).  Reproducing those codes bit-for-bit needs the
// exact double roundings; this header performs them with 64-bit integer ops.
// The kernels compute with binary64 hardware operations (equally correctly
// rounded) and use this header for ToInt32 of out-of-range values only.  The
// same code compiles for the host (g++) so tests can check it against native
// double arithmetic.
//
// Values are passed as raw uint64 bit patterns.  Inputs may be zero, normal or
// subnormal; results round correctly, including subnormal results and
// overflow to infinity.  Inf/NaN inputs are not handled here: callers test
// finiteness first (as the encode path must anyway, for JS semantics).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SD_HD __host__ __device__ inline
#else
#define SD_HD static inline
#endif

namespace sd {

SD_HD int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __clzll((long long)x);
#else
  return x ? __builtin_clzll(x) : 64;
#endif
}

// high 64 bits of a 64x64 product
SD_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

struct Unpacked {
  uint64_t sign;  // 0 or 1ull << 63
  int exp;        // biased exponent of the leading mantissa bit
  uint64_t mant;  // mantissa with the leading bit at position 52; 0 for zero
};

SD_HD Unpacked unpack(uint64_t x) {
  Unpacked u;
  u.sign = x & 0x8000000000000000ull;
  int e = (int)((x >> 52) & 0x7ff);
  uint64_t m = x & 0x000fffffffffffffull;
  if (e == 0) {
    if (m == 0) {
      u.exp = 0;
      u.mant = 0;
      return u;
    }
    // subnormal: normalise so the leading bit sits at 52
    int sh = clz64(m) - 11;
    u.mant = m << sh;
    u.exp = 1 - sh;
    return u;
  }
  u.mant = m | 0x0010000000000000ull;
  u.exp = e;
  return u;
}

// Round and pack.  value = r * 2^(e - 1085) with r in [2^62, 2^63) (or 0),
// bit 0 of r is sticky.  Handles subnormal results and overflow.
SD_HD uint64_t pack(uint64_t sign, int e, uint64_t r) {
  if (r == 0) return sign;
  if (e >= 2047) return sign | 0x7ff0000000000000ull;
  uint64_t ebase;
  if (e <= 0) {
    int sh = 1 - e;
    r = sh >= 64 ? (r != 0) : ((r >> sh) | ((r & ((1ull << sh) - 1)) != 0));
    ebase = 0;
  } else {
    ebase = (uint64_t)(e - 1);
  }
  uint64_t mant = r >> 10;
  uint64_t rb = r & 0x3ff;
  if (rb > 0x200 || (rb == 0x200 && (mant & 1))) mant += 1;
  return sign | ((ebase << 52) + mant);
}

SD_HD uint64_t neg(uint64_t a) { return a ^ 0x8000000000000000ull; }

// a + b
SD_HD uint64_t add(uint64_t a, uint64_t b) {
  Unpacked x = unpack(a), y = unpack(b);
  if (x.mant == 0 && y.mant == 0) return x.sign & y.sign;  // +0 unless both -0
  if (x.mant == 0) return b;
  if (y.mant == 0) return a;
  // |x| >= |y|
  if (y.exp > x.exp || (y.exp == x.exp && y.mant > x.mant)) {
    Unpacked t = x;
    x = y;
    y = t;
  }
  uint64_t mx = x.mant << 10;  // leading bit at 62
  uint64_t my = y.mant << 10;
  int d = x.exp - y.exp;
  if (d > 0) {
    if (d >= 64) {
      my = 1;  // pure sticky
    } else {
      uint64_t lost = my & ((1ull << d) - 1);
      my = (my >> d) | (lost != 0);
    }
  }
  int e = x.exp;
  uint64_t r;
  if (x.sign == y.sign) {
    r = mx + my;  // < 2^64
    if (r & 0x8000000000000000ull) {
      r = (r >> 1) | (r & 1);
      e += 1;
    }
  } else {
    r = mx - my;
    if (r == 0) return 0;  // exact cancellation gives +0 in RNE
    int sh = clz64(r) - 1;  // bring the leading bit to 62
    r <<= sh;               // exact: a shift > 1 implies d <= 1 (no sticky)
    e -= sh;
  }
  return pack(x.sign, e, r);
}

SD_HD uint64_t sub(uint64_t a, uint64_t b) { return add(a, neg(b)); }

// a * b
SD_HD uint64_t mul(uint64_t a, uint64_t b) {
  Unpacked x = unpack(a), y = unpack(b);
  uint64_t sign = x.sign ^ y.sign;
  if (x.mant == 0 || y.mant == 0) return sign;
  // product of two 53-bit mantissas: < 2^106.  Shift operands so the
  // 128-bit product's top word carries the leading bits.
  uint64_t hi = mulhi64(x.mant << 11, y.mant << 11);  // P * 2^22 >> 64
  uint64_t lo = (x.mant << 11) * (y.mant << 11);
  // P*2^22 < 2^128, leading bit at 126 or 127
  int e;
  uint64_t r;
  if (hi & 0x8000000000000000ull) {
    r = (hi >> 1) | (hi & 1) | (lo != 0);
    e = x.exp + y.exp - 1022;
  } else {
    r = hi | (lo != 0);
    e = x.exp + y.exp - 1023;
  }
  return pack(sign, e, r);
}

// a / b  (b != 0)
SD_HD uint64_t div(uint64_t a, uint64_t b) {
  Unpacked x = unpack(a), y = unpack(b);
  uint64_t sign = x.sign ^ y.sign;
  if (x.mant == 0) return sign;
  if (y.mant == 0) return sign | 0x7ff0000000000000ull;
  // q = floor(mx / my * 2^63) by restoring division, 64 quotient bits
  uint64_t rem = x.mant, q = 0;
  for (int i = 0; i < 64; i++) {
    uint64_t bit = rem >= y.mant;
    if (bit) rem -= y.mant;
    q = (q << 1) | bit;
    rem <<= 1;
  }
  uint64_t sticky = rem != 0;
  int e;
  uint64_t r;
  if (q & 0x8000000000000000ull) {
    r = (q >> 1) | (q & 1) | sticky;
    e = x.exp - y.exp + 1023;
  } else {
    r = q | sticky;
    e = x.exp - y.exp + 1022;
  }
  return pack(sign, e, r);
}

// exact widening of an f32 bit pattern (finite) to binary64
SD_HD uint64_t from_f32(uint32_t f) {
  uint64_t sign = (uint64_t)(f >> 31) << 63;
  uint32_t e = (f >> 23) & 0xff;
  uint32_t m = f & 0x7fffff;
  if (e == 0xff) return sign | 0x7ff0000000000000ull | ((uint64_t)m << 29);
  if (e == 0) {
    if (m == 0) return sign;
    int sh = clz64((uint64_t)m) - 40;  // leading bit of m to position 23
    m <<= sh;
    return sign | ((uint64_t)(1023 - 126 - sh) << 52) | ((uint64_t)(m & 0x7fffff) << 29);
  }
  return sign | ((uint64_t)(e - 127 + 1023) << 52) | ((uint64_t)m << 29);
}

// binary64 -> binary32, round to nearest even (finite input)
SD_HD uint32_t to_f32(uint64_t a) {
  uint32_t sign = (uint32_t)(a >> 63) << 31;
  uint32_t ef = (uint32_t)((a >> 52) & 0x7ff);
  if (ef == 0x7ff) return sign | 0x7f800000u | ((a & 0x000fffffffffffffull) ? 0x400000u : 0);
  Unpacked x = unpack(a);
  if (x.mant == 0) return sign;
  int e = x.exp - 1023 + 127;  // f32 biased exponent of the leading bit
  uint64_t r = x.mant << 10;   // leading bit at 62
  if (e >= 255) return sign | 0x7f800000u;
  uint32_t ebase;
  if (e <= 0) {
    int sh = 1 - e;
    r = sh >= 64 ? (r != 0) : ((r >> sh) | ((r & ((1ull << sh) - 1)) != 0));
    ebase = 0;
  } else {
    ebase = (uint32_t)(e - 1);
  }
  uint64_t mant = r >> 39;  // 24 bits
  uint64_t rb = r & ((1ull << 39) - 1);
  uint64_t half = 1ull << 38;
  if (rb > half || (rb == half && (mant & 1))) mant += 1;
  return sign | ((ebase << 23) + (uint32_t)mant);
}

// ECMAScript ToInt32 (the `| 0` of mortonTree.js:56-58); NaN/Inf -> 0
SD_HD int32_t to_int32(uint64_t a) {
  if (((a >> 52) & 0x7ff) == 0x7ff) return 0;
  Unpacked x = unpack(a);
  if (x.mant == 0) return 0;
  int k = x.exp - 1075;  // value = mant * 2^k
  uint32_t v;
  if (k >= 32) {
    v = 0;
  } else if (k >= 0) {
    v = (uint32_t)(x.mant << k);
  } else if (k > -64) {
    v = (uint32_t)(x.mant >> (-k));
  } else {
    v = 0;
  }
  if (x.sign) v = 0u - v;
  return (int32_t)v;
}

SD_HD bool is_nan(uint64_t a) {
  return ((a >> 52) & 0x7ff) == 0x7ff && (a & 0x000fffffffffffffull);
}

// JS `a < b` on finite-or-infinite non-NaN values
SD_HD bool lt(uint64_t a, uint64_t b) {
  bool sa = a >> 63, sb = b >> 63;
  uint64_t ma = a & 0x7fffffffffffffffull, mb = b & 0x7fffffffffffffffull;
  if (ma == 0 && mb == 0) return false;
  if (sa != sb) return sa;
  return sa ? ma > mb : ma < mb;
}

}  // namespace sd
