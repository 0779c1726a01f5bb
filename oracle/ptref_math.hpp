// ptref_math.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Scalar CPU restatement of the vector helpers, RNG and sampling routines the
// reference's device programs use.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load the oracle, and only as the checker.
//
// Every function cites the reference file:line it restates.  Floating-point
// expressions keep the reference's evaluation order; the oracle is compiled
// with -ffp-contract=off so that no multiply-add is fused.
#pragma once

#include <cmath>
#include <cstdint>

#include "ptref_libm.hpp"

// The oracle is written over `Real`: float for parity with the GPU path
// (the default), double for converged reference images (a translation unit
// that defines PTREF_DOUBLE; the RNG stays 32-bit integer).
namespace ptref {

#ifdef PTREF_DOUBLE
using Real = double;
#else
using Real = float;
#endif

constexpr Real PI_F = 3.141592654f;  // programs/common.hpp:18

struct f3 {
  Real x, y, z;
};
struct f2 {
  Real x, y;
};

inline f3 mk(Real x, Real y, Real z) { return f3{x, y, z}; }
inline f3 mk(Real s) { return f3{s, s, s}; }
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 operator*(f3 a, Real s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(Real s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
// one IEEE reciprocal and three multiplies (as the GPU code)
inline f3 operator/(f3 a, Real s) {
  const Real r = 1.0f / s;
  return {a.x * r, a.y * r, a.z * r};
}
inline f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
inline f3& operator*=(f3& a, f3 b) { a = a * b; return a; }
inline f3& operator*=(f3& a, Real s) { a = a * s; return a; }

inline Real dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Real length(f3 a) { return std::sqrt(dot(a, a)); }
// normalize(v) = v * (1 / |v|)
inline f3 normalize(f3 v) {
  Real inv = 1.0f / std::sqrt(dot(v, v));
  return v * inv;
}
// vec.hpp:21 squared_length = length * length (not dot)
inline Real squared_length(f3 a) { return length(a) * length(a); }
// vec.hpp:26
inline bool isNull(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
// vec.hpp:47
inline f3 unit_vector(f3 v) { return v / length(v); }
// vec.hpp:57-64
inline Real ffmin(Real a, Real b) { return a < b ? a : b; }
inline Real ffmax(Real a, Real b) { return a > b ? a : b; }
// vec.hpp:67-83
inline f3 min_vec(f3 a, f3 b) { return {ffmin(a.x, b.x), ffmin(a.y, b.y), ffmin(a.z, b.z)}; }
inline f3 max_vec(f3 a, f3 b) { return {ffmax(a.x, b.x), ffmax(a.y, b.y), ffmax(a.z, b.z)}; }
inline Real max_component(f3 a) { return ffmax(ffmax(a.x, a.y), a.z); }
inline Real min_component(f3 a) { return ffmin(ffmin(a.x, a.y), a.z); }
// clamp(x, lo, hi) = fmaxf(lo, fminf(x, hi))
inline Real clampf(Real x, Real lo, Real hi) { return std::fmax(lo, std::fmin(x, hi)); }
inline f3 clamp3(f3 v, Real lo, Real hi) {
  return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)};
}
// reflect(i, n) = i - 2 n (n . i)
inline f3 reflect(f3 i, f3 n) { return i - n * (2.0f * dot(n, i)); }
inline Real saturate(Real x) { return ffmax(0.f, ffmin(1.f, x)); }  // host_common.hpp:71

// ------------------------------------------------------------------ RNG
// programs/random.cuh:33-47 (TEA, N rounds)
inline uint32_t tea(uint32_t val0, uint32_t val1, int rounds = 64) {
  uint32_t v0 = val0, v1 = val1, s0 = 0;
  for (int n = 0; n < rounds; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
// programs/random.cuh:50-55
inline uint32_t lcg(uint32_t& prev) {
  prev = 1664525u * prev + 1013904223u;
  return prev & 0x00FFFFFFu;
}
// programs/random.cuh:63-65 -- exact in fp32 (24-bit integer / 2^24)
inline Real rnd(uint32_t& prev) { return (Real)lcg(prev) / (Real)0x01000000; }

// ------------------------------------------------------------- sampling
// programs/sampling.cuh:22-29
inline f3 random_in_unit_disk(uint32_t& seed) {
  Real a = rnd(seed) * 2.f * PI_F;
  f3 xy = mk(psin(a), pcos(a), 0.f);
  xy *= std::sqrt(rnd(seed));
  return xy;
}
// programs/sampling.cuh:31-44
inline f3 random_in_unit_sphere(uint32_t& seed) {
  Real z = rnd(seed) * 2.f - 1.f;
  Real t = rnd(seed) * 2.f * PI_F;
  Real r = std::sqrt((0.f > (1.f - z * z) ? 0.f : (1.f - z * z)));
  Real x = r * pcos(t);
  Real y = r * psin(t);
  f3 res = mk(x, y, z);
  res *= ppow(rnd(seed), 1.f / 3.f);
  return res;
}
// programs/sampling.cuh:46-59
inline f3 random_on_unit_sphere(uint32_t& seed) {
  Real z = rnd(seed) * 2.f - 1.f;
  Real t = rnd(seed) * 2.f * PI_F;
  Real r = std::sqrt((0.f > (1.f - z * z) ? 0.f : (1.f - z * z)));
  Real x = r * pcos(t);
  Real y = r * psin(t);
  f3 res = mk(x, y, z);
  res *= ppow(rnd(seed), 1.f / 3.f);
  return unit_vector(res);
}

// Cosine-weighted hemisphere sample about +z by Malley's method: a uniform
// point on the unit disk (r = sqrt(u1), phi = 2 pi u2) lifted to the
// hemisphere.  Used where the reference calls the OptiX SDK helper
// cosine_sample_hemisphere (lambertian.cuh:13, oren_nayar.cuh:14,
// ashikhmin_shirley.cuh:35); the SDK's own op order is not reproduced.
inline f3 cosine_sample_hemisphere(Real u1, Real u2) {
  Real r = std::sqrt(u1);
  Real phi = 2.0f * PI_F * u2;
  Real x = r * pcos(phi);
  Real y = r * psin(phi);
  Real z = std::sqrt(std::fmax(0.0f, 1.0f - x * x - y * y));
  return mk(x, y, z);
}

// Orthonormal basis around a unit normal n (Duff et al., "Building an
// Orthonormal Basis, Revisited", JCGT 6(1), 2017).  Replaces the OptiX SDK
// Onb used by lambertian.cuh:15-16 and friends: local (x, y, z) maps to
// x*t + y*b + z*n.
struct Onb {
  f3 t, b, n;
  explicit Onb(f3 nn) {
    n = nn;
    Real sign = std::copysign(1.0f, n.z);
    Real a = -1.0f / (sign + n.z);
    Real bb = n.x * n.y * a;
    t = mk(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    b = mk(bb, sign + n.y * n.y * a, -n.y);
  }
  f3 to_world(f3 p) const { return t * p.x + b * p.y + n * p.z; }
};

// programs/math/trigonometric.cuh:5-58 (angles against world +Y)
inline Real CosTheta(f3 w) { return w.y; }
inline Real Cos2Theta(f3 w) { return w.y * w.y; }
inline Real AbsCosTheta(f3 w) { return std::fabs(w.y); }
inline Real Sin2Theta(f3 w) { return std::fmax(0.f, 1.f - Cos2Theta(w)); }
inline Real SinTheta(f3 w) { return std::sqrt(Sin2Theta(w)); }
inline Real TanTheta(f3 w) { return SinTheta(w) / CosTheta(w); }
inline Real CosPhi(f3 w) {
  Real s = SinTheta(w);
  return (s == 0) ? 1.f : clampf(w.x / s, -1.f, 1.f);
}
inline Real SinPhi(f3 w) {
  Real s = SinTheta(w);
  return (s == 0) ? 0.f : clampf(w.z / s, -1.f, 1.f);
}
inline Real Cos2Phi(f3 w) { Real c = CosPhi(w); return c * c; }
inline Real Sin2Phi(f3 w) { Real s = SinPhi(w); return s * s; }
// trigonometric.cuh:124-126
inline f3 Spherical_Vector(Real sintheta, Real costheta, Real phi) {
  return mk(sintheta * pcos(phi), costheta, sintheta * psin(phi));
}

}  // namespace ptref
