// Host-side rounding of doubles to the 16-bit MFMA operand formats (bit
// patterns), used to lay out the reference gradients of the BF16 / FP16
// matrix-free kernels.  Plain C++: no device code, testable on the host.
#pragma once
#include <cstdint>
#include <cstring>

namespace ngsamd {

inline uint32_t FloatBits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// bfloat16, round to nearest even (via FP32; NaN stays NaN)
inline uint16_t ToBf16(double v) {
  const uint32_t u = FloatBits((float)v);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

// IEEE binary16, round to nearest even (via FP32), with subnormals, overflow
// to infinity and NaN kept quiet
inline uint16_t ToF16(double v) {
  const uint32_t u = FloatBits((float)v);
  const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
  const uint32_t a = u & 0x7FFFFFFFu;
  if (a >= 0x7F800000u) return (uint16_t)(sign | (a > 0x7F800000u ? 0x7E00u : 0x7C00u));
  const int e = (int)(a >> 23) - 127 + 15;  // biased binary16 exponent
  uint32_t m = a & 0x007FFFFFu;
  if (e >= 31) return (uint16_t)(sign | 0x7C00u);
  if (e <= 0) {
    if (e < -10) return sign;  // below half the smallest subnormal: +-0
    m |= 0x00800000u;
    const int shift = 14 - e;  // 14..24
    uint32_t h = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return (uint16_t)(sign | h);  // a carry into bit 10 gives the smallest normal
  }
  uint32_t h = ((uint32_t)e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // a carry may reach infinity
  return (uint16_t)(sign | h);
}

// value of an IEEE binary16 bit pattern (exact)
inline double F16ToDouble(uint16_t h) {
  const int e = (h >> 10) & 0x1F, m = h & 0x3FF;
  const double sign = (h & 0x8000u) ? -1.0 : 1.0;
  if (e == 31) return m ? sign * (0.0 / 0.0) : sign * (1.0 / 0.0);
  if (e == 0) return sign * m * 0x1p-24;  // subnormal: m 2^-24
  double v = (1024 + m) * 0x1p-10;
  for (int i = 15; i < e; i++) v *= 2.0;
  for (int i = e; i < 15; i++) v *= 0.5;
  return sign * v;
}

// Two-term FP16 split v ~ hi + lo: hi = fp16(v), lo = fp16(v - hi) (the
// operands of the split-FP16 matrix-free kernels)
inline void SplitF16(double v, uint16_t* hi, uint16_t* lo) {
  *hi = ToF16(v);
  *lo = ToF16(v - F16ToDouble(*hi));
}

}  // namespace ngsamd
