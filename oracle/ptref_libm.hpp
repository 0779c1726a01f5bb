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
//   pow: 2^(y log2 x) in float, ln by the atanh series of (m - 1) / (m + 1)
//     on m in [sqrt(1/2), sqrt(2)), 2^r by reduction to |r| <= 1/2, its
//     Taylor series and exact scaling; pow5(x) = x^5 by products.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace ptref {

#ifndef PTREF_DOUBLE
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

// ln(m) of x = m 2^e, m in [sqrt(1/2), sqrt(2)): 2 atanh(s), s = (m - 1) /
// (m + 1), by the odd series to s^11 (|s| <= 0.1716: the next term is below
// 2^-30 relative); e returned
inline float pln_m(float x, int& e) {
  uint32_t b = lm_bits(x);
  e = 0;
  if (b < 0x00800000u) {  // subnormal: scale by 2^24
    x = x * 16777216.0f;
    b = lm_bits(x);
    e = -24;
  }
  e += (int)((b >> 23) & 0xFFu) - 127;
  float m = lm_float((b & 0x007FFFFFu) | 0x3F800000u);
  if (m > 0x1.6a09e6p+0f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = s2 * (0x1.555556p-2f + s2 * (0x1.99999ap-3f + s2 * (0x1.24924ap-3f + s2 * (0x1.c71c72p-4f + s2 * 0x1.745d18p-4f))));
  return 2.0f * (s + s * p);
}
// natural log of a positive finite float: e ln2 with ln2 split in a 16-bit
// head (e * head exact) and its tail
inline float pln(float x) {
  int e;
  float lm = pln_m(x, e);
  return (float)e * 0x1.62e4p-1f + ((float)e * 0x1.7f7d1cp-20f + lm);
}
// 2^z: z = k + r, |r| <= 1/2, e^(r ln 2) by Taylor to degree 8, scaled by 2^k
inline float pexp2(float z) {
  if (z > 128.0f) return std::numeric_limits<float>::infinity();
  if (z < -151.0f) return 0.0f;
  float k = std::floor(z + 0.5f);
  float t = (z - k) * 0x1.62e430p-1f;
  float q = 1.0f + t * (1.0f + t * (0x1p-1f + t * (0x1.555556p-3f + t * (0x1.555556p-5f + t * (0x1.111112p-7f + t * (0x1.6c16c2p-10f + t * (0x1.a01a02p-13f + t * 0x1.a01a02p-16f)))))));
  return std::ldexp(q, (int)k);
}
// x^y = 2^(y log2 x), all in float
inline float ppow(float x, float y) {
  if (x != x || y != y) return x + y;
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (x < 0.0f) return std::numeric_limits<float>::quiet_NaN();
  if (x == 0.0f) return y > 0.0f ? 0.0f : std::numeric_limits<float>::infinity();
  if (std::isinf(x)) return y > 0.0f ? x : 0.0f;
  int e;
  float lm = pln_m(x, e);
  return pexp2(y * ((float)e + lm * 0x1.715476p+0f));
}
inline float ppow5(float x) { return (x * x) * (x * x) * x; }
// natural log; -inf at 0
inline float plogf(float x) { return x > 0.0f ? pln(x) : -std::numeric_limits<float>::infinity(); }
#else
// The double build of the oracle (converged reference images) takes the C
// library's double functions: it is not held to the GPU's arithmetic.
inline void psincos(double x, double& s, double& c) { s = std::sin(x); c = std::cos(x); }
inline double psin(double x) { return std::sin(x); }
inline double pcos(double x) { return std::cos(x); }
inline double ptan(double x) { return std::tan(x); }
inline double patan(double x) { return std::atan(x); }
inline double patan2(double y, double x) { return std::atan2(y, x); }
inline double pasin(double x) { return std::asin(x); }
inline double pacos(double x) { return std::acos(x); }
inline double ppow(double x, double y) { return std::pow(x, y); }
inline double ppow5(double x) { return (x * x) * (x * x) * x; }
inline double plogf(double x) { return x > 0.0 ? std::log(x) : -std::numeric_limits<double>::infinity(); }
#endif

}  // namespace ptref
