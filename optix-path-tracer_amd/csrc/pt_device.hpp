// pt_device.hpp -- device-side math, RNG, intersection, BRDF, light and miss
// functions of the gfx950 path tracer.
//
// Every function cites the reference program it replaces.  Floating-point
// expressions keep the reference's evaluation order and the library is
// built with -ffp-contract=off; transcendentals come from pt_libm.hpp.  The
// results are the same IEEE fp32 values a scalar CPU evaluation of the same
// expressions produces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_layout.hpp"
#include "pt_libm.hpp"

namespace pt {

#define PTD __device__ __forceinline__

constexpr float PI_F = 3.141592654f;  // programs/common.hpp:18
constexpr float FINF = __builtin_huge_valf();
constexpr float SLAB_FAR_SCALE = 1.00000024f;  // 1 + 2^-22, conservative slab test

struct v3 {
  float x, y, z;
};
PTD v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
PTD v3 mk3(float s) { return v3{s, s, s}; }
PTD v3 mk3(float4 a) { return v3{a.x, a.y, a.z}; }
PTD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PTD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PTD v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
PTD v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PTD v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PTD v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
PTD v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
// one IEEE reciprocal and three multiplies (the oracle does the same)
PTD v3 operator/(v3 a, float s) {
  const float r = 1.0f / s;
  return {a.x * r, a.y * r, a.z * r};
}
PTD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PTD v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PTD float length(v3 a) { return sqrtf(dot(a, a)); }
PTD v3 normalize(v3 v) {
  float inv = 1.0f / sqrtf(dot(v, v));
  return v * inv;
}
PTD float squared_length(v3 a) { return length(a) * length(a); }        // vec.hpp:21
PTD bool isNull(v3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }  // vec.hpp:26
PTD v3 unit_vector(v3 v) { return v / length(v); }                        // vec.hpp:47
PTD float ffmin(float a, float b) { return a < b ? a : b; }               // vec.hpp:57
PTD float ffmax(float a, float b) { return a > b ? a : b; }               // vec.hpp:62
PTD v3 min_vec(v3 a, v3 b) { return {ffmin(a.x, b.x), ffmin(a.y, b.y), ffmin(a.z, b.z)}; }
PTD v3 max_vec(v3 a, v3 b) { return {ffmax(a.x, b.x), ffmax(a.y, b.y), ffmax(a.z, b.z)}; }
PTD float max_component(v3 a) { return ffmax(ffmax(a.x, a.y), a.z); }
PTD float min_component(v3 a) { return ffmin(ffmin(a.x, a.y), a.z); }
PTD float clampf(float x, float lo, float hi) { return fmaxf(lo, fminf(x, hi)); }
PTD v3 clamp3(v3 v, float lo, float hi) { return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)}; }
PTD v3 reflect(v3 i, v3 n) { return i - n * (2.0f * dot(n, i)); }
PTD float comp(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// ------------------------------------------------------------------ RNG
// programs/random.cuh:33-47.  Fully unrolled so that every v + s0 is one add
// of a literal: 6 VALU ops per half round, 768 per seed (the path start is
// VALU-bound on this hash).
PTD uint32_t tea64(uint32_t val0, uint32_t val1) {
  uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
  for (int n = 0; n < 64; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
// programs/random.cuh:50-55: the 24-bit LCG output itself
PTD uint32_t lcg24(uint32_t& prev) {
  prev = 1664525u * prev + 1013904223u;
  return prev & 0x00FFFFFFu;
}
// programs/random.cuh:50-65: 24-bit LCG output / 2^24 (exact in fp32)
PTD float rnd(uint32_t& prev) {
  prev = 1664525u * prev + 1013904223u;
  return (float)(prev & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// programs/sampling.cuh:22-29
PTD v3 random_in_unit_disk(uint32_t& seed) {
  float a = rnd(seed) * 2.f * PI_F;
  v3 xy = mk3(psin(a), pcos(a), 0.f);
  return xy * sqrtf(rnd(seed));
}
// programs/sampling.cuh:31-44
PTD v3 random_in_unit_sphere(uint32_t& seed) {
  float z = rnd(seed) * 2.f - 1.f;
  float t = rnd(seed) * 2.f * PI_F;
  float r = sqrtf((0.f > (1.f - z * z) ? 0.f : (1.f - z * z)));
  v3 res = mk3(r * pcos(t), r * psin(t), z);
  return res * ppow(rnd(seed), 1.f / 3.f);
}
// programs/sampling.cuh:46-59
PTD v3 random_on_unit_sphere(uint32_t& seed) {
  float z = rnd(seed) * 2.f - 1.f;
  float t = rnd(seed) * 2.f * PI_F;
  float r = sqrtf((0.f > (1.f - z * z) ? 0.f : (1.f - z * z)));
  v3 res = mk3(r * pcos(t), r * psin(t), z);
  res = res * ppow(rnd(seed), 1.f / 3.f);
  return unit_vector(res);
}
// Malley's method: uniform disk point (r = sqrt(u1), phi = 2 pi u2) lifted to
// the +z hemisphere (replaces the OptiX SDK cosine_sample_hemisphere).
PTD v3 cosine_sample_hemisphere(float u1, float u2) {
  float r = sqrtf(u1);
  float phi = 2.0f * PI_F * u2;
  float x = r * pcos(phi);
  float y = r * psin(phi);
  float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
  return mk3(x, y, z);
}
// Orthonormal basis about a unit normal (Duff et al., JCGT 6(1), 2017);
// replaces the OptiX SDK Onb.  Local (x, y, z) -> x*t + y*b + z*n.
PTD v3 onb_to_world(v3 n, v3 p) {
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float bb = n.x * n.y * a;
  v3 t = mk3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  v3 b = mk3(bb, sign + n.y * n.y * a, -n.y);
  return t * p.x + b * p.y + n * p.z;
}

// programs/math/trigonometric.cuh:5-58 (angles against world +Y)
PTD float Cos2Theta(v3 w) { return w.y * w.y; }
PTD float AbsCosTheta(v3 w) { return fabsf(w.y); }
PTD float SinTheta(v3 w) { return sqrtf(fmaxf(0.f, 1.f - Cos2Theta(w))); }
PTD float TanTheta(v3 w) { return SinTheta(w) / w.y; }
PTD float CosPhi(v3 w) {
  float s = SinTheta(w);
  return (s == 0) ? 1.f : clampf(w.x / s, -1.f, 1.f);
}
PTD float SinPhi(v3 w) {
  float s = SinTheta(w);
  return (s == 0) ? 0.f : clampf(w.z / s, -1.f, 1.f);
}
PTD float Cos2Phi(v3 w) { float c = CosPhi(w); return c * c; }
PTD float Sin2Phi(v3 w) { float s = SinPhi(w); return s * s; }

// ------------------------------------------------------------ transforms
// rows r0..r2 of a 3x4 matrix
PTD v3 xf_point(float4 r0, float4 r1, float4 r2, v3 p) {
  return mk3(r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w, r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w,
             r2.x * p.x + r2.y * p.y + r2.z * p.z + r2.w);
}
PTD v3 xf_vector(float4 r0, float4 r1, float4 r2, v3 v) {
  return mk3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
             r2.x * v.x + r2.y * v.y + r2.z * v.z);
}
// normal by the transpose of world-to-object
PTD v3 xf_normal(float4 w0, float4 w1, float4 w2, v3 n) {
  return mk3(w0.x * n.x + w1.x * n.y + w2.x * n.z, w0.y * n.x + w1.y * n.y + w2.y * n.z,
             w0.z * n.x + w1.z * n.y + w2.z * n.z);
}

// ------------------------------------------------------------ intersectors
// sphere.cu:32-62: the nearer root above tmin is the only one that can win
PTD float isect_sphere(float4 s, v3 o, v3 d, float tmin) {
  v3 oc = o - mk3(s);
  float a = dot(d, d);
  float b = dot(oc, d);
  float c = dot(oc, oc) - s.w * s.w;
  float disc = b * b - a * c;
  if (disc < 0.f) return FINF;
  float sq = sqrtf(disc);
  float t1 = (-b - sq) / a;
  if (t1 > tmin) return t1;
  return (-b + sq) / a;
}
// aarect.cu:20-69 (inclusive bounds); A = (a0, a1, b0, b1), k, axis.  The
// plane distance multiplies by the ray's reciprocal direction `inv`
// (1 / d per component, computed once per instance) where the reference
// divides: one multiply instead of an IEEE division per test.
PTD float isect_rect(float4 A, float k, int axis, v3 o, v3 d, v3 inv) {
  float t, a, b;
  if (axis == 0) { t = (k - o.x) * inv.x; a = o.y + t * d.y; b = o.z + t * d.z; }
  else if (axis == 1) { t = (k - o.y) * inv.y; a = o.x + t * d.x; b = o.z + t * d.z; }
  else { t = (k - o.z) * inv.z; a = o.x + t * d.x; b = o.y + t * d.y; }
  if (a < A.x || a > A.y || b < A.z || b > A.w) return FINF;
  return t;
}
// box.cu:36-53: tmin offered first, tmax when tmin is not above the ray
// tmin; slab distances multiply by the reciprocal direction `inv`
PTD v3 mulv(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
PTD float isect_box(v3 lo, v3 hi, v3 o, v3 inv, float tmin) {
  v3 t0 = mulv(lo - o, inv);
  v3 t1 = mulv(hi - o, inv);
  float tn = max_component(min_vec(t0, t1));
  float tf = min_component(max_vec(t0, t1));
  if (!(tn <= tf)) return FINF;
  return tn > tmin ? tn : tf;
}

// "lowbias32" integer hash (C. Wellons)
PTD uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
// Free-flight sample of a volume.  The reference draws rnd(prd.seed) inside
// the intersection program (volume_sphere.cu:84, volume_box.cu:63), so the
// draw would depend on the traversal's visit order; here xi is a hash of
// the ray's RNG state, the volume and its instance, and the path's own
// stream is left as it is.  Only the low 24 bits of the state enter (the
// bits every rnd() output depends on): in a batch of scenes the top byte
// holds the scene index, and the path draws the same free flights as in a
// render of its scene alone.
PTD float vol_xi(uint32_t seed, uint32_t vol, uint32_t inst) {
  const uint32_t h = mix32((seed & 0xFFFFFFu) ^ mix32(vol * 0x9E3779B9u ^ mix32(inst + 0x632BE5ABu)));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}
// Free-flight seed of query ray i: a hash of the whole 64-bit index (vol_xi
// reads 24 bits of it), so rays 2^24 apart do not share their flights.
PTD uint32_t query_vseed(uint64_t i) { return mix32((uint32_t)i ^ mix32((uint32_t)(i >> 32) + 0x7F4A7C15u)); }
constexpr float FLT_BIG = 3.40282347e+38f;
// hit_boundary of volume_sphere.cu:21-55 and volume_box.cu:20-40: the first
// boundary root in (lo, hi); A = (a, density), B = (b, index | box << 31)
PTD bool vol_boundary(float4 A, float4 B, bool box, v3 o, v3 d, float lo, float hi, float& rec) {
  if (box) {
    v3 t0 = (mk3(A) - o) / d;
    v3 t1 = (mk3(B) - o) / d;
    float ta = max_component(min_vec(t0, t1));
    float tb = min_component(max_vec(t0, t1));
    if (ta > tb) return false;
    if (ta < hi && ta > lo) { rec = ta; return true; }
    if (tb < hi && tb > lo) { rec = tb; return true; }
    return false;
  }
  v3 oc = o - mk3(A);
  float a = dot(d, d);
  float b = dot(oc, d);
  float c = dot(oc, oc) - B.x * B.x;
  float disc = b * b - a * c;
  if (disc < 0.f) return false;
  float temp = (-b - sqrtf(disc)) / a;
  if (temp < hi && temp > lo) { rec = temp; return true; }
  temp = (-b + sqrtf(disc)) / a;
  if (temp < hi && temp > lo) { rec = temp; return true; }
  return false;
}
// volume_sphere.cu:63-110, volume_box.cu:42-85: entry and exit of the
// boundary, then a free flight of -ln(xi) / density; the scatter point, or
// +inf when the flight leaves the medium first (the reference reports the
// point even then; the boundary test of Shirley's constant medium is kept
// here).  The exit is the boundary's own, where the reference clips it to
// the current closest hit: the closest hit is the same either way.
PTD float isect_volume(float4 A, float4 B, v3 o, v3 d, float tmin, float xi) {
  const bool box = (__float_as_uint(B.w) >> 31) != 0;
  float r1 = 0.f, r2 = 0.f;
  if (!vol_boundary(A, B, box, o, d, -FLT_BIG, FLT_BIG, r1)) return FINF;
  if (!vol_boundary(A, B, box, o, d, r1 + 0.0001f, FLT_BIG, r2)) return FINF;
  if (r1 < tmin) r1 = tmin;
  if (r1 >= r2) return FINF;
  if (r1 < 0.f) r1 = 0.f;
  const float len = length(d);
  const float inside = (r2 - r1) * len;
  const float hd = -(1.f / A.w) * plogf(xi);
  if (!(hd < inside)) return FINF;
  return r1 + hd / len;
}

// cylinder.cu:47-66 fixed to its intent (the reference's coefficients
// read b = d.x P.x + d.z + P.z and c = P.x^2 + P.z^2 + R^2, and it rejects
// t < delta): the lateral surface x^2 + z^2 = R^2 around the origin's
// vertical, for 0 <= y - origin.y <= length; the nearer valid root.  Rays
// parallel to the axis (and grazing ones, delta < 1e-8, as the reference)
// miss.  A = (origin, length), B.x = radius.
PTD float isect_cylinder(float4 A, float4 B, v3 o, v3 d, float tmin) {
  v3 P0 = o - mk3(A);
  float a = d.x * d.x + d.z * d.z;
  float b = d.x * P0.x + d.z * P0.z;
  float c = P0.x * P0.x + P0.z * P0.z - B.x * B.x;
  float delta = b * b - a * c;
  if (delta < 1e-8f) return FINF;
  float sq = sqrtf(delta);
  float t = (-b - sq) / a;
  float y = P0.y + t * d.y;
  if (t > tmin && y >= 0.f && y <= A.w) return t;
  t = (-b + sq) / a;
  y = P0.y + t * d.y;
  if (t > tmin && y >= 0.f && y <= A.w) return t;
  return FINF;
}
// moving_sphere.cu center(): center0 + ((time - time0) / (time1 - time0)) (center1 - center0)
PTD v3 moving_center(float4 A, float4 B, float4 C, float time) {
  return mk3(A) + ((time - C.x) / (C.y - C.x)) * (mk3(B) - mk3(A));
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1), 2013).
struct TriRay {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
// a * b = hi + lo exactly (Dekker's product over Veltkamp's 12-bit split; no
// fused multiply-add)
PTD void two_prod(float a, float b, float& hi, float& lo) {
  const float sa = a * 4097.0f, sb = b * 4097.0f;
  const float ah = sa - (sa - a), al = a - ah;
  const float bh = sb - (sb - b), bl = b - bh;
  hi = a * b;
  lo = (((ah * bh - hi) + ah * bl) + al * bh) + al * bl;
}
// a * b - c * d with the exact sign and a nearly correctly rounded value: the
// edge functions of the watertight test when one of them rounds to zero
// (Woop et al. 2013 recompute them at higher precision there)
PTD float diff_of_products(float a, float b, float c, float d) {
  float h1, l1, h2, l2;
  two_prod(a, b, h1, l1);
  two_prod(c, d, h2, l2);
  const float s = h1 - h2, z = s - h1;
  const float t = (h1 - (s - z)) + (-h2 - z);
  return s + (t + (l1 - l2));
}
PTD TriRay tri_setup(v3 d) {
  TriRay r;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
  r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
  if (comp(d, r.kz) < 0.0f) { int t = r.kx; r.kx = r.ky; r.ky = t; }
  r.Sx = comp(d, r.kx) / comp(d, r.kz);
  r.Sy = comp(d, r.ky) / comp(d, r.kz);
  r.Sz = 1.0f / comp(d, r.kz);
  return r;
}
PTD bool tri_test(const TriRay& r, v3 o, v3 v0, v3 v1, v3 v2, float& t, float& b1, float& b2) {
  v3 A = v0 - o, B = v1 - o, C = v2 - o;
  float Ax = comp(A, r.kx) - r.Sx * comp(A, r.kz);
  float Ay = comp(A, r.ky) - r.Sy * comp(A, r.kz);
  float Bx = comp(B, r.kx) - r.Sx * comp(B, r.kz);
  float By = comp(B, r.ky) - r.Sy * comp(B, r.kz);
  float Cx = comp(C, r.kx) - r.Sx * comp(C, r.kz);
  float Cy = comp(C, r.ky) - r.Sy * comp(C, r.kz);
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = diff_of_products(Cx, By, Cy, Bx);
    V = diff_of_products(Ax, Cy, Ay, Cx);
    W = diff_of_products(Bx, Ay, By, Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * comp(A, r.kz);
  float Bz = r.Sz * comp(B, r.kz);
  float Cz = r.Sz * comp(C, r.kz);
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  t = T * rcp;
  b1 = V * rcp;
  b2 = W * rcp;
  return true;
}

// ------------------------------------------------------------ materials
// material.cuh:30-39
PTD float schlick(float cosine, float ref_idx) {
  float r0 = (1.f - ref_idx) / (1.f + ref_idx);
  r0 = r0 * r0;
  return r0 + (1.f - r0) * ppow5((1.f - cosine));
}
PTD v3 schlick3(v3 r0, float cosine) {
  float e = ppow5(1.f - cosine);
  return r0 + (mk3(1.f) - r0) * e;
}
// material.cuh:50-63
PTD bool Refract(v3 v, v3 n, float ni_over_nt, v3& refracted) {
  v3 uv = normalize(v);
  float dt = dot(uv, n);
  float disc = 1.f - ni_over_nt * ni_over_nt * (1.f - dt * dt);
  if (disc > 0.f) {
    refracted = ni_over_nt * (uv - n * dt) - n * sqrtf(disc);
    return true;
  }
  return false;
}
// microfacets.cuh:11-20
PTD void Sample_Quadrant(float nu, float nv, float rx, float ry, float& phi, float& theta) {
  phi = patan(sqrtf((nu + 1.f) / (nv + 1.f)) * ptan(2.f * PI_F * rx));
  float cp = pcos(phi), sp = psin(phi);
  theta = ppow(ry, 1.f / (nu * cp * cp + nv * sp * sp + 1.f));
}
// microfacets.cuh:85-99
PTD float GGX_D(v3 H, float nu, float nv) {
  float c2 = Cos2Theta(H);
  if (c2 <= 0.0f) return 0.f;
  if (nu == nv) {
    float beta = c2 * ((nu * nu) - 1) + 1;
    return (nu * nu) / (PI_F * beta * beta);
  }
  float s2nv = Sin2Phi(H) / (nv * nv);
  float c2nu = (1.f - Sin2Phi(H)) / (nu * nu);
  float beta = (c2 + (1.f - c2) * (c2nu + s2nv));
  return 1.f / (PI_F * nu * nv * beta * beta);
}
// microfacets.cuh:102-168 (visible-slope GGX sample)
PTD v3 GGX_Sample(v3 origin, float rx, float ry, float nu, float nv) {
  bool flip = origin.y < 0;
  v3 so = normalize(mk3(origin.x * nu, origin.y, origin.z * nv));
  float slopeX, slopeY;
  float cosTheta = so.y;
  if (cosTheta > 0.9999f) {
    float r = sqrtf(rx / (1 - rx));
    float phi = 6.28318530718f * ry;
    slopeX = r * pcos(phi);
    slopeY = r * psin(phi);
  } else {
    float sinTheta = sqrtf(fmaxf(0.f, 1.f - cosTheta * cosTheta));
    float tanTheta = sinTheta / cosTheta;
    float a = 1.f / tanTheta;
    float G1 = 2.f / (1.f + sqrtf(1.f + 1.f / (a * a)));
    float A = 2.f * rx / G1 - 1.f;
    float tmp = 1.f / (A * A - 1.f);
    if (tmp > 1e10f) tmp = 1e10f;
    float B = tanTheta;
    float D = sqrtf(fmaxf(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.f));
    float sx1 = B * tmp - D;
    float sx2 = B * tmp + D;
    slopeX = (A < 0 || sx2 > 1.f / tanTheta) ? sx1 : sx2;
    float S;
    if (ry > 0.5f) { S = 1.f; ry = 2.f * (ry - .5f); }
    else { S = -1.f; ry = 2.f * (.5f - ry); }
    float z = (ry * (ry * (ry * 0.27385f - 0.73369f) + 0.46341f)) /
              (ry * (ry * (ry * 0.093073f + 0.309420f) - 1.f) + 0.597999f);
    slopeY = S * z * sqrtf(1.f + slopeX * slopeX);
  }
  float cp = CosPhi(so), sp = SinPhi(so);
  float t = cp * slopeX - sp * slopeY;
  slopeY = sp * slopeX + cp * slopeY;
  slopeX = t;
  slopeX = nu * slopeX;
  slopeY = nv * slopeY;
  v3 H = normalize(mk3(-slopeX, 1.f, -slopeY));
  if (flip) H = H * -1.f;
  return H;
}
// microfacets.cuh:170-195
PTD float GGX_Lambda(v3 V, float nu, float nv) {
  float at = fabsf(TanTheta(V));
  if (__builtin_isinf(at)) return 0.f;
  float alpha = sqrtf(Cos2Phi(V) * nu * nu + Sin2Phi(V) * nv * nv);
  float a2t2 = (alpha * at) * (alpha * at);
  return (-1.f + sqrtf(1.f + a2t2)) / 2.f;
}
PTD float GGX_G(v3 Wo, v3 Wi, float nu, float nv) {
  return 1.f / (1.f + GGX_Lambda(Wo, nu, nv) + GGX_Lambda(Wi, nu, nv));
}

struct Surface {
  int type;
  v3 c0, c1;
  float p0, p1;
};

// Sample(): lambertian.cuh:7-19, oren_nayar.cuh:8-20, diffuse_light.cuh:7-13,
// ashikhmin_shirley.cuh:11-82, torrance_sparrow.cuh:10-35
// Material sets a kernel is compiled for (bit = 1 << PT_MAT_*).
constexpr int MS_ALL = 0x7F;   // the surface materials
constexpr int MS_FULL = 0x1FF; // and the isotropic medium and the normal shader
constexpr int MS_DIFFUSE = (1 << PT_MAT_LAMBERTIAN) | (1 << PT_MAT_OREN_NAYAR) | (1 << PT_MAT_DIFFUSE_LIGHT);
constexpr int MS_SPHERES = (1 << PT_MAT_LAMBERTIAN) | (1 << PT_MAT_METAL) | (1 << PT_MAT_DIELECTRIC) |
                           (1 << PT_MAT_DIFFUSE_LIGHT);
template <int MS>
PTD constexpr bool has_mat(int t) {
  return (MS >> t) & 1;
}

template <int MS = MS_ALL>
PTD v3 brdf_sample(const Surface& f, v3 Wo, v3 N, uint32_t& seed) {
  if ((has_mat<MS>(PT_MAT_LAMBERTIAN) || has_mat<MS>(PT_MAT_OREN_NAYAR)) &&
      (f.type == PT_MAT_LAMBERTIAN || f.type == PT_MAT_OREN_NAYAR)) {
    float u1 = rnd(seed);
    float u2 = rnd(seed);
    return onb_to_world(N, cosine_sample_hemisphere(u1, u2));
  }
  if (has_mat<MS>(PT_MAT_DIFFUSE_LIGHT) && f.type == PT_MAT_DIFFUSE_LIGHT) return random_on_unit_sphere(seed);
  if (has_mat<MS>(PT_MAT_ISOTROPIC) && f.type == PT_MAT_ISOTROPIC) return random_on_unit_sphere(seed);  // isotropic.cuh:7-13
  if (!has_mat<MS>(PT_MAT_ASHIKHMIN_SHIRLEY) && !has_mat<MS>(PT_MAT_TORRANCE_SPARROW)) return mk3(0.f);
  float nu = f.p0, nv = f.p1;
  v3 Nn = normalize(N);
  v3 T = normalize(cross(Nn, mk3(0.f, 1.f, 0.f)));
  v3 B = cross(T, Nn);
  float rx = rnd(seed);
  float ry = rnd(seed);
  v3 H;
  if (has_mat<MS>(PT_MAT_ASHIKHMIN_SHIRLEY) && (!has_mat<MS>(PT_MAT_TORRANCE_SPARROW) || f.type == PT_MAT_ASHIKHMIN_SHIRLEY)) {
    if (rx < 0.5f) {
      rx = fminf(2 * rx, 1.f - 1e-6f);
      return onb_to_world(Nn, cosine_sample_hemisphere(rx, ry));
    }
    rx = fminf(2 * (rx - 0.5f), 1.f - 1e-6f);
    float phi, cos_theta;
    if (nu == nv) {
      phi = 2.f * PI_F * rx;
      cos_theta = ppow(ry, 1.f / (nu + 1.f));
    } else {
      if (rx < 0.25f) { Sample_Quadrant(nu, nv, 4.f * rx, ry, phi, cos_theta); }
      else if (rx < 0.5f) { Sample_Quadrant(nu, nv, 4.f * (0.5f - rx), ry, phi, cos_theta); phi = PI_F - phi; }
      else if (rx < 0.75f) { Sample_Quadrant(nu, nv, 4.f * (rx - 0.5f), ry, phi, cos_theta); phi = PI_F + phi; }
      else { Sample_Quadrant(nu, nv, 4.f * (1.f - rx), ry, phi, cos_theta); phi = 2.f * PI_F - phi; }
    }
    float sin_theta = sqrtf(fmaxf(0.f, 1.f - cos_theta * cos_theta));
    H = normalize(mk3(sin_theta * pcos(phi), cos_theta, sin_theta * psin(phi)));
  } else {
    H = normalize(GGX_Sample(Wo, rx, ry, nu, nv));
  }
  H = H.x * B + H.y * Nn + H.z * T;
  if (dot(H, Wo) < 0.f) H = -H;
  return normalize(-Wo + 2.f * dot(Wo, H) * H);
}

// Evaluate(): lambertian.cuh:21-42, oren_nayar.cuh:22-65,
// diffuse_light.cuh:15-22, ashikhmin_shirley.cuh:84-151,
// torrance_sparrow.cuh:37-64.  pdf keeps its incoming value on early returns.
template <int MS = MS_ALL>
PTD v3 brdf_eval(const Surface& f, v3 Wo, v3 Wi, v3 N, float& pdf) {
  int type = f.type;
  // fold the dispatch for materials outside the compiled set
  if (!has_mat<MS>(type)) return mk3(0.f);
  switch (type) {
    case PT_MAT_LAMBERTIAN: {
      float cosine = dot(normalize(Wi), normalize(N));
      if (cosine < 0.f) return mk3(0.f);
      pdf = dot(normalize(Wi), normalize(N)) * (1.0f / PI_F);  // lambertian.cuh:33-38, 1/pi folded
      return (cosine * f.c0) / PI_F;
    }
    case PT_MAT_OREN_NAYAR: {
      v3 WiN = normalize(Wi);
      float sI = SinTheta(WiN), sO = SinTheta(Wo);
      float maxCos = 0;
      if (sI > 1e-4f && sO > 1e-4f) {
        float sPI = SinPhi(WiN), cPI = CosPhi(WiN);
        float sPO = SinPhi(Wo), cPO = CosPhi(Wo);
        float dCos = cPI * cPO + sPI * sPO;
        maxCos = fmaxf(0.f, dCos);
      }
      float sinAlpha, tanBeta;
      if (AbsCosTheta(WiN) > AbsCosTheta(Wo)) { sinAlpha = sO; tanBeta = sI / AbsCosTheta(WiN); }
      else { sinAlpha = sI; tanBeta = sO / AbsCosTheta(Wo); }
      pdf = dot(normalize(Wi), normalize(N)) / PI_F;
      return f.c0 * (1.f / PI_F) * (f.p0 + f.p1 * maxCos * sinAlpha * tanBeta);
    }
    case PT_MAT_DIFFUSE_LIGHT:
      pdf = 1.f;
      return f.c0;
    case PT_MAT_ISOTROPIC: {
      // isotropic.cuh:15-32 returns pi/4 for both; the phase function and
      // its pdf are 1/(4 pi)
      const float k = 0.25f / PI_F;
      pdf = k;
      return f.c0 * k;
    }
    case PT_MAT_ASHIKHMIN_SHIRLEY: {
      v3 Rd = f.c0, Rs = f.c1;
      float nu = f.p0, nv = f.p1;
      v3 Up = mk3(0.f, 1.f, 0.f);
      v3 Nn = normalize(N);
      v3 T = normalize(cross(Nn, Up));
      v3 B = cross(T, Nn);
      float NdotI = fabsf(dot(Up, Wi)), NdotO = fabsf(dot(Up, Wo));
      v3 diff = Rd * (28.f / (23.f * PI_F));
      diff = diff * (mk3(1.f) - Rs);
      diff = diff * (1.f - ppow5(1.f - NdotI * 0.5f));
      diff = diff * (1.f - ppow5(1.f - NdotO * 0.5f));
      float dpdf = dot(Wi, Nn) / PI_F;
      if (dpdf < 0.f) return mk3(0.f);
      v3 H = Wo + Wi;
      if (isNull(H)) return mk3(0.f);
      H = normalize(H);
      float HdotI = fabsf(dot(H, Wi));
      float HdotN = fabsf(dot(H, Nn)), HdotT = fabsf(dot(H, T)), HdotB = fabsf(dot(H, B));
      float norm, lobe;
      if (nu == nv) {
        norm = (nu + 1.f) / (8.f * PI_F);
        lobe = ppow(HdotN, nu);
      } else {
        norm = sqrtf((nu + 1.f) * (nv + 1.f)) / (8.f * PI_F);
        if (HdotN < 1.f)
          lobe = ppow(HdotN, (nu * HdotT * HdotT + nv * HdotB * HdotB) / (1.f - HdotN * HdotN));
        else
          lobe = 1.f;
      }
      v3 spec = schlick3(Rs, HdotI);
      spec = spec * norm;
      spec = spec * (lobe / (HdotI * fmaxf(NdotI, NdotO)));
      float spdf = HdotI / (norm * lobe);
      pdf = (dpdf + spdf) / 2.f;
      return (diff / dpdf) + (spec / spdf);
    }
    default: {  // Torrance-Sparrow
      v3 Rs = f.c0;
      float nu = f.p0, nv = f.p1;
      v3 Up = mk3(0.f, 1.f, 0.f);
      float NdotI = fmaxf(dot(Up, Wi), 1e-6f), NdotO = fmaxf(dot(Up, Wo), 1e-6f);
      v3 H = Wo + Wi;
      if (isNull(H)) return mk3(0.f);
      H = normalize(H);
      float HdotI = fabsf(dot(H, Wi));
      v3 F = schlick3(Rs, HdotI);
      float G = GGX_G(Wo, Wi, nu, nv);
      float D = GGX_D(H, nu, nv);
      pdf = (D * AbsCosTheta(H)) / (4.f * dot(Wo, H));
      return Rs * D * G * F / (4.f * NdotI * NdotO);
    }
  }
}

// light_sample.cuh:27-33 with numf = numg = 1
PTD float power_heuristic(float fPdf, float gPdf) {
  float f = 1u * fPdf, g = 1u * gPdf;
  return (f * f) / (f * f + g * g);
}

PTD uint32_t sig_mix(uint32_t h, uint32_t v) { return (h ^ v) * 16777619u; }

// ------------------------------------------------------- radiance cache
// The network's outputs y are log(1 + radiance) per channel; the radiance
// it stands for is expm1(min(y, CACHE_Y_MAX)).
constexpr float CACHE_Y_MAX = 20.f;     // e^20 ~ 4.9e8 per unit throughput
constexpr float CACHE_REL_EPS = 0.01f;  // relative-loss denominator floor
// Radiance of a query's answer: non-negative, 0 for a NaN output.
PTD float cache_radiance(float y) {
  const float r = expm1f(y > CACHE_Y_MAX ? CACHE_Y_MAX : y);
  return r > 0.f ? r : 0.f;
}
// Relative L2 loss of Mueller et al. 2021 (neural radiance caching) on one
// output: with yh = expm1(min(y, CACHE_Y_MAX)) and the target radiance t,
//   loss = (yh - t)^2 / (yh^2 + CACHE_REL_EPS),  the denominator held
//   constant, so dloss/dy = 2 (yh - t) (yh + 1) / (yh^2 + CACHE_REL_EPS);
// its minimiser over noisy targets of one input is their mean.  Above
// CACHE_Y_MAX the gradient is taken at the clamp (an over-bright output is
// still pulled down); a NaN output gives a NaN loss.
PTD void cache_loss_terms(float y, float t, float& loss, float& dy) {
  const float yh = expm1f(y > CACHE_Y_MAX ? CACHE_Y_MAX : y);
  const float den = yh * yh + CACHE_REL_EPS;
  const float d = yh - t;
  loss = d * d / den;
  dy = 2.f * d * (yh + 1.f) / den;
}

}  // namespace pt
