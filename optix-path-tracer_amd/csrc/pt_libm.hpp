// pt_libm.hpp -- elementary functions of the gfx950 kernels.
//
// Own implementations, evaluated with exactly the operations and order of a
// scalar CPU build (no fused multiply-add), so that results are
// reproducible bit for bit on the host:
//   sin, cos, tan: Cody-Waite reduction by pi/2 (three-part constant) and
//     Taylor polynomials of degree 9 / 10 on [-pi/4, pi/4];
//   atan, atan2, asin, acos: reduction to |t| <= tan(pi/12) via
//     atan(a) = pi/6 + atan((a*sqrt3 - 1) / (a + sqrt3)) and the odd Taylor
//     series to t^11 (Cody and Waite, "Software Manual for the Elementary
//     Functions", 1980);
//   pow: 2^(y log2 x) in float (ln by the atanh series of (m - 1) / (m + 1),
//     2^r by reduction to |r| <= 1/2 and Taylor, then exact scaling);
//   pow5(x) = x^5 by products.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

#define PTL __device__ __forceinline__

PTL void psincos(float x, float& s, float& c) {
  if (!(fabsf(x) < 2097152.0f)) {
    s = c = __builtin_nanf("");
    return;
  }
  float kf = floorf(x * 0x1.45f306p-1f + 0.5f);
  float r = ((x - kf * 0x1.92p+0f) - kf * 0x1.fb4p-12f) - kf * 0x1.4442d2p-24f;
  int q = ((int)kf) & 3;
  float r2 = r * r;
  float sp = r + r * (r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f))));
  float cp = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f))));
  if (q == 0) { s = sp; c = cp; }
  else if (q == 1) { s = cp; c = -sp; }
  else if (q == 2) { s = -sp; c = -cp; }
  else { s = -cp; c = sp; }
}
PTL float psin(float x) { float s, c; psincos(x, s, c); return s; }
PTL float pcos(float x) { float s, c; psincos(x, s, c); return c; }
PTL float ptan(float x) { float s, c; psincos(x, s, c); return s / c; }

PTL float patan_unit(float a) {
  float off = 0.0f;
  if (a > 0x1.126146p-2f) {
    a = (a * 0x1.bb67aep+0f - 1.0f) / (a + 0x1.bb67aep+0f);
    off = 0x1.0c1524p-1f;
  }
  float a2 = a * a;
  float p = a + a * (a2 * (-0x1.555556p-2f + a2 * (0x1.99999ap-3f + a2 * (-0x1.24924ap-3f + a2 * (0x1.c71c72p-4f + a2 * -0x1.745d18p-4f)))));
  return off + p;
}
PTL float patan(float x) {
  float ax = fabsf(x);
  if (ax != ax) return x;
  float r = ax > 1.0f ? 0x1.921fb6p+0f - patan_unit(1.0f / ax) : patan_unit(ax);
  return copysignf(r, x);
}
PTL float patan2(float y, float x) {
  if (x != x || y != y) return x + y;
  float ax = fabsf(x), ay = fabsf(y);
  float r;
  if (ay == 0.0f) {
    r = __builtin_signbitf(x) ? 0x1.921fb6p+1f : 0.0f;
  } else if (ax == 0.0f) {
    r = 0x1.921fb6p+0f;
  } else if (__builtin_isinf(ax) && __builtin_isinf(ay)) {
    r = __builtin_signbitf(x) ? 0x1.2d97c8p+1f : 0x1.921fb6p-1f;
  } else {
    r = ay > ax ? 0x1.921fb6p+0f - patan_unit(ax / ay) : patan_unit(ay / ax);
    if (x < 0.0f) r = 0x1.921fb6p+1f - r;
  }
  return copysignf(r, y);
}
PTL float pasin(float x) { return patan2(x, sqrtf((1.0f - x) * (1.0f + x))); }
PTL float pacos(float x) { return patan2(sqrtf((1.0f - x) * (1.0f + x)), x); }

// ln(m) of x = m 2^e, m in [sqrt(1/2), sqrt(2)): 2 atanh(s), s = (m - 1) /
// (m + 1), by the odd series to s^11 (|s| <= 0.1716: the next term is below
// 2^-30 relative); e returned
PTL float pln_m(float x, int& e) {
  uint32_t b = __float_as_uint(x);
  e = 0;
  if (b < 0x00800000u) {  // subnormal: scale by 2^24
    x = x * 16777216.0f;
    b = __float_as_uint(x);
    e = -24;
  }
  e += (int)((b >> 23) & 0xFFu) - 127;
  float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
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
PTL float pln(float x) {
  int e;
  float lm = pln_m(x, e);
  return (float)e * 0x1.62e4p-1f + ((float)e * 0x1.7f7d1cp-20f + lm);
}
// 2^z: z = k + r, |r| <= 1/2, e^(r ln 2) by Taylor to degree 8, scaled by 2^k
PTL float pexp2(float z) {
  if (z > 128.0f) return __builtin_huge_valf();
  if (z < -151.0f) return 0.0f;
  float k = floorf(z + 0.5f);
  float t = (z - k) * 0x1.62e430p-1f;
  float q = 1.0f + t * (1.0f + t * (0x1p-1f + t * (0x1.555556p-3f + t * (0x1.555556p-5f + t * (0x1.111112p-7f + t * (0x1.6c16c2p-10f + t * (0x1.a01a02p-13f + t * 0x1.a01a02p-16f)))))));
  return ldexpf(q, (int)k);
}
// x^y = 2^(y log2 x), all in float
PTL float ppow(float x, float y) {
  if (x != x || y != y) return x + y;
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_huge_valf();
  if (__builtin_isinf(x)) return y > 0.0f ? x : 0.0f;
  int e;
  float lm = pln_m(x, e);
  return pexp2(y * ((float)e + lm * 0x1.715476p+0f));
}
PTL float ppow5(float x) { return (x * x) * (x * x) * x; }
// natural log; -inf at 0
PTL float plogf(float x) { return x > 0.0f ? pln(x) : -__builtin_huge_valf(); }

}  // namespace pt
