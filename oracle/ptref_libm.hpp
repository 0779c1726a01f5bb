// ptref_libm.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Elementary functions of the path tracer, specified here so that the CPU
// oracle and the gfx950 kernels evaluate them identically (same operations,
// same order, no fused multiply-add):
//   sin, cos, tan: Cody-Waite reduction by pi/2 (three-part constant) and
//     Taylor polynomials of degree 9 / 10 on [-pi/4, pi/4];
//   atan, atan2, asin, acos: reduction to |t| <= tan(pi/12) via
//     atan(a) = pi/6 + atan((a*sqrt3 - 1) / (a + sqrt3)) and the odd Taylor
//     series to t^11 (Cody and Waite, "Software Manual for the Elementary
//     Functions", 1980);
//   pow: exp(y * ln x) in double precision, ln by the atanh series of
//     (m - 1) / (m + 1) on m in [sqrt(1/2), sqrt(2)), exp by reduction
//     modulo ln 2 and its Taylor series; pow5(x) = x^5 by products.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace ptref {

inline uint32_t lm_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float lm_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// returns false for arguments outside |x| < 2^21 (results are NaN there)
inline void psincos(float x, float& s, float& c) {
  if (!(std::fabs(x) < 2097152.0f)) {
    s = c = std::numeric_limits<float>::quiet_NaN();
    return;
  }
  float kf = std::floor(x * 0x1.45f306p-1f + 0.5f);
  float r = ((x - kf * 0x1.92p+0f) - kf * 0x1.fb4p-12f) - kf * 0x1.4442d2p-24f;
  int q = ((int)kf) & 3;
  float r2 = r * r;
  float sp = r + r * (r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f))));
  float cp = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f))));
  switch (q) {
    case 0: s = sp; c = cp; break;
    case 1: s = cp; c = -sp; break;
    case 2: s = -sp; c = -cp; break;
    default: s = -cp; c = sp; break;
  }
}
inline float psin(float x) { float s, c; psincos(x, s, c); return s; }
inline float pcos(float x) { float s, c; psincos(x, s, c); return c; }
inline float ptan(float x) { float s, c; psincos(x, s, c); return s / c; }

// atan of a in [0, 1]
inline float patan_unit(float a) {
  float off = 0.0f;
  if (a > 0x1.126146p-2f) {
    a = (a * 0x1.bb67aep+0f - 1.0f) / (a + 0x1.bb67aep+0f);
    off = 0x1.0c1524p-1f;
  }
  float a2 = a * a;
  float p = a + a * (a2 * (-0x1.555556p-2f + a2 * (0x1.99999ap-3f + a2 * (-0x1.24924ap-3f + a2 * (0x1.c71c72p-4f + a2 * -0x1.745d18p-4f)))));
  return off + p;
}
inline float patan(float x) {
  float ax = std::fabs(x);
  if (ax != ax) return x;
  float r = ax > 1.0f ? 0x1.921fb6p+0f - patan_unit(1.0f / ax) : patan_unit(ax);
  return std::copysign(r, x);
}
inline float patan2(float y, float x) {
  if (x != x || y != y) return x + y;
  float ax = std::fabs(x), ay = std::fabs(y);
  float r;
  if (ay == 0.0f) {
    r = std::signbit(x) ? 0x1.921fb6p+1f : 0.0f;
  } else if (ax == 0.0f) {
    r = 0x1.921fb6p+0f;
  } else if (std::isinf(ax) && std::isinf(ay)) {
    r = std::signbit(x) ? 0x1.2d97c8p+1f : 0x1.921fb6p-1f;
  } else {
    r = ay > ax ? 0x1.921fb6p+0f - patan_unit(ax / ay) : patan_unit(ay / ax);
    if (x < 0.0f) r = 0x1.921fb6p+1f - r;
  }
  return std::copysign(r, y);
}
inline float pasin(float x) { return patan2(x, std::sqrt((1.0f - x) * (1.0f + x))); }
inline float pacos(float x) { return patan2(std::sqrt((1.0f - x) * (1.0f + x)), x); }

// natural log of a positive finite float, in double
inline double pln(float x) {
  uint32_t b = lm_bits(x);
  int e = 0;
  if (b < 0x00800000u) {  // subnormal: scale by 2^24
    x = x * 16777216.0f;
    b = lm_bits(x);
    e = -24;
  }
  e += (int)((b >> 23) & 0xFFu) - 127;
  double m = (double)lm_float((b & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = s2 * (1.0 / 3.0 + s2 * (1.0 / 5.0 + s2 * (1.0 / 7.0 + s2 * (1.0 / 9.0 + s2 * (1.0 / 11.0 + s2 * (1.0 / 13.0))))));
  return (double)e * 0x1.62e42fefa39efp-1 + 2.0 * (s + s * p);
}
// exp of a double, rounded to float
inline float pexp_d(double z) {
  if (z > 89.0) return std::numeric_limits<float>::infinity();
  if (z < -104.0) return 0.0f;
  double k = std::floor(z * 0x1.71547652b82fep+0 + 0.5);
  double r = z - k * 0x1.62e42fefa39efp-1;
  double p = 1.0 + r * (1.0 + r * (1.0 / 2.0 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0 + r * (1.0 / 720.0 + r * (1.0 / 5040.0 + r * (1.0 / 40320.0 + r * (1.0 / 362880.0 + r * (1.0 / 3628800.0))))))))));
  return (float)std::ldexp(p, (int)k);
}
inline float ppow(float x, float y) {
  if (x != x || y != y) return x + y;
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (x < 0.0f) return std::numeric_limits<float>::quiet_NaN();
  if (x == 0.0f) return y > 0.0f ? 0.0f : std::numeric_limits<float>::infinity();
  if (std::isinf(x)) return y > 0.0f ? x : 0.0f;
  return pexp_d((double)y * pln(x));
}
inline float ppow5(float x) { return (x * x) * (x * x) * x; }
// natural log in float through pln; -inf at 0
inline float plogf(float x) { return x > 0.0f ? (float)pln(x) : -std::numeric_limits<float>::infinity(); }

}  // namespace ptref
