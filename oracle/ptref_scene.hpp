// ptref_scene.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Scalar CPU restatement of the reference's geometry path: primitive bounds
// and intersection programs (programs/hitables/*.cu), instance transforms
// (host_includes/transforms.hpp), the acceleration structure and closest-hit
// traversal that OptiX performs inside rtTrace (programs/raygen.cu:77).
//
// OptiX's Trbvh builder and traversal order are closed and absent, so the
// acceleration structure here is the project's LBVH (30-bit Morton codes,
// stable sort, Karras 2012 hierarchy, collapse to 4-wide nodes), restated on
// the CPU so the GPU build can be compared bit for bit.  Closest-hit results
// do not depend on the structure: they are also checked against brute force.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/pt_api.h"
#include "ptref_math.hpp"

namespace ptref {

constexpr Real INF = std::numeric_limits<Real>::infinity();
constexpr int BLAS_LEAF_MAX = 4;
constexpr int TLAS_LEAF_MAX = 1;
// Conservative slab test: tfar is scaled by 1 + 2^-22 before the compare.
constexpr Real SLAB_FAR_SCALE = 1.00000024f;

struct Aabb {
  f3 lo{INF, INF, INF}, hi{-INF, -INF, -INF};
};
inline Aabb merge(const Aabb& a, const Aabb& b) {
  Aabb r;
  r.lo = mk(std::fmin(a.lo.x, b.lo.x), std::fmin(a.lo.y, b.lo.y), std::fmin(a.lo.z, b.lo.z));
  r.hi = mk(std::fmax(a.hi.x, b.hi.x), std::fmax(a.hi.y, b.hi.y), std::fmax(a.hi.z, b.hi.z));
  return r;
}
// half surface area, ((dx*dy) + (dy*dz)) + (dz*dx)
inline Real half_area(const Aabb& b) {
  Real dx = b.hi.x - b.lo.x, dy = b.hi.y - b.lo.y, dz = b.hi.z - b.lo.z;
  return dx * dy + dy * dz + dz * dx;
}

// ------------------------------------------------------------ transforms
// Row-major 3x4 affine transform as uploaded by Transform::setMatrix
// (transforms.hpp:65); rtTransformPoint/Vector/Normal restated.
inline f3 xform_point(const Real* m, f3 p) {
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
            m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
inline f3 xform_vector(const Real* m, f3 v) {
  return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
// normal: transpose of the world-to-object 3x3 (inverse transpose of M)
inline f3 xform_normal(const Real* w2o, f3 n) {
  return mk(w2o[0] * n.x + w2o[4] * n.y + w2o[8] * n.z, w2o[1] * n.x + w2o[5] * n.y + w2o[9] * n.z,
            w2o[2] * n.x + w2o[6] * n.y + w2o[10] * n.z);
}

// ------------------------------------------------------------ primitives
struct Sphere { f3 c; Real r; int mat; };
struct Rect { int axis, flip; Real a0, a1, b0, b1, k; int mat; };
struct Box { f3 lo, hi; int mat; };
// hitables.hpp:175-209 / :316-349: shape PT_VOLUME_SPHERE (a = center,
// b.x = radius) or PT_VOLUME_BOX (a = lo, b = hi)
struct Volume { int shape; f3 a, b; Real density; int mat; };
// hitables.hpp:411-454: lateral surface around the vertical through o
struct Cylinder { f3 o; Real L, R; int mat; };
// hitables.hpp:131-172
struct MovingSphere { f3 c0, c1; Real r, t0, t1; int mat; };
// moving_sphere.cu center()
inline f3 moving_center(const MovingSphere& m, Real time) {
  return m.c0 + ((time - m.t0) / (m.t1 - m.t0)) * (m.c1 - m.c0);
}
struct Mesh {
  std::vector<f3> v, n;
  std::vector<f2> uv;
  std::vector<uint32_t> idx;
  std::vector<int> face_mat;
  std::vector<int> face_idx;  // per-face vector-texture index (empty: 0)
  int mat = 0;
  uint32_t ntri() const { return (uint32_t)(idx.size() / 3); }
};

// bounds programs: sphere.cu:96-100, aarect.cu:117-136, box.cu:84-88,
// triangle.cu:61-83
inline Aabb bounds(const Sphere& s) {
  Aabb b;
  b.lo = s.c - mk(s.r);
  b.hi = s.c + mk(s.r);
  return b;
}
inline Aabb bounds(const Rect& r) {
  Aabb b;
  if (r.axis == PT_AXIS_X) { b.lo = mk(r.k - 0.0001f, r.a0, r.b0); b.hi = mk(r.k + 0.0001f, r.a1, r.b1); }
  else if (r.axis == PT_AXIS_Y) { b.lo = mk(r.a0, r.k - 0.0001f, r.b0); b.hi = mk(r.a1, r.k + 0.0001f, r.b1); }
  else { b.lo = mk(r.a0, r.b0, r.k - 0.0001f); b.hi = mk(r.a1, r.b1, r.k + 0.0001f); }
  return b;
}
inline Aabb bounds(const Box& x) {
  Aabb b;
  b.lo = x.lo - mk(0.001f);
  b.hi = x.hi + mk(0.001f);
  return b;
}
// volume_sphere.cu:111-115 (no padding), volume_box.cu:90-94 (1e-4)
inline Aabb bounds(const Volume& v) {
  Aabb r;
  if (v.shape == PT_VOLUME_BOX) {
    r.lo = mk(v.a.x - 0.0001f, v.a.y - 0.0001f, v.a.z - 0.0001f);
    r.hi = mk(v.b.x + 0.0001f, v.b.y + 0.0001f, v.b.z + 0.0001f);
  } else {
    r.lo = mk(v.a.x - v.b.x, v.a.y - v.b.x, v.a.z - v.b.x);
    r.hi = mk(v.a.x + v.b.x, v.a.y + v.b.x, v.a.z + v.b.x);
  }
  return r;
}
// cylinder.cu:100-107
inline Aabb bounds(const Cylinder& c) {
  f3 o2 = mk(c.o.x + c.L * 0.f, c.o.y + c.L * 1.f, c.o.z + c.L * 0.f);
  Aabb r;
  r.lo = mk(std::fmin(c.o.x, o2.x) - c.R, std::fmin(c.o.y, o2.y) - c.R, std::fmin(c.o.z, o2.z) - c.R);
  r.hi = mk(std::fmax(c.o.x, o2.x) + c.R, std::fmax(c.o.y, o2.y) + c.R, std::fmax(c.o.z, o2.z) + c.R);
  return r;
}
// moving_sphere.cu:137-147: boxes at time0 and time1
inline Aabb bounds(const MovingSphere& m) {
  f3 p = moving_center(m, m.t0), q = moving_center(m, m.t1);
  Aabb r;
  r.lo = mk(std::fmin(p.x - m.r, q.x - m.r), std::fmin(p.y - m.r, q.y - m.r), std::fmin(p.z - m.r, q.z - m.r));
  r.hi = mk(std::fmax(p.x + m.r, q.x + m.r), std::fmax(p.y + m.r, q.y + m.r), std::fmax(p.z + m.r, q.z + m.r));
  return r;
}
inline Aabb bounds_tri(f3 a, f3 bb, f3 c) {
  Aabb r;
  r.lo = mk(std::fmin(std::fmin(a.x, bb.x), c.x) - 0.0001f, std::fmin(std::fmin(a.y, bb.y), c.y) - 0.0001f,
            std::fmin(std::fmin(a.z, bb.z), c.z) - 0.0001f);
  r.hi = mk(std::fmax(std::fmax(a.x, bb.x), c.x) + 0.0001f, std::fmax(std::fmax(a.y, bb.y), c.y) + 0.0001f,
            std::fmax(std::fmax(a.z, bb.z), c.z) + 0.0001f);
  return r;
}

// Every box the traversal tests is widened outward by at least one ulp:
// x -+ (|x| 2^-23 + 2^-126).  A ray lying exactly in a bounding plane with a
// zero direction component then never meets 0 * inf in the slab test, so
// the structure never culls a primitive whose intersection program accepts
// points on its (inclusive) bounds.
inline Real widen_lo(Real x) { return x - (std::fabs(x) * 0x1p-23f + 0x1p-126f); }
inline Real widen_hi(Real x) { return x + (std::fabs(x) * 0x1p-23f + 0x1p-126f); }
inline Aabb widen(const Aabb& b) {
  Aabb r;
  r.lo = mk(widen_lo(b.lo.x), widen_lo(b.lo.y), widen_lo(b.lo.z));
  r.hi = mk(widen_hi(b.hi.x), widen_hi(b.hi.y), widen_hi(b.hi.z));
  return r;
}

// Candidate acceptance shared by every intersector: tmin < t and either
// t < best or (t == best and the (instance, primitive) key is smaller).
struct HitState {
  Real t;
  int inst = -1, prim = -1;
  bool found() const { return inst >= 0; }
  bool offer(Real tc, Real tmin, int ci, int cp) {
    if (!(tc > tmin)) return false;
    if (tc < t || (found() && tc == t && (ci < inst || (ci == inst && cp < prim)))) {
      t = tc; inst = ci; prim = cp;
      return true;
    }
    return false;
  }
};

// sphere.cu:32-62 -- both roots offered in order
inline void isect_sphere(const Sphere& s, f3 o, f3 d, Real tmin, HitState& h, int inst, int prim) {
  f3 oc = o - s.c;
  Real a = dot(d, d);
  Real b = dot(oc, d);
  Real c = dot(oc, oc) - s.r * s.r;
  Real disc = b * b - a * c;
  if (disc < 0.f) return;
  Real t = (-b - std::sqrt(disc)) / a;
  h.offer(t, tmin, inst, prim);
  t = (-b + std::sqrt(disc)) / a;
  h.offer(t, tmin, inst, prim);
}
// aarect.cu:20-69 -- inclusive bounds
// The plane and slab distances of rects and boxes multiply by the ray's
// reciprocal direction `inv` (1 / d per component, once per instance)
// where the reference divides (it was compiled with --use_fast_math, whose
// division is approximate anyway).
inline void isect_rect(const Rect& r, f3 o, f3 d, f3 inv, Real tmin, HitState& h, int inst, int prim) {
  Real t, a, b;
  if (r.axis == PT_AXIS_X) { t = (r.k - o.x) * inv.x; a = o.y + t * d.y; b = o.z + t * d.z; }
  else if (r.axis == PT_AXIS_Y) { t = (r.k - o.y) * inv.y; a = o.x + t * d.x; b = o.z + t * d.z; }
  else { t = (r.k - o.z) * inv.z; a = o.x + t * d.x; b = o.y + t * d.y; }
  if (a < r.a0 || a > r.a1 || b < r.b0 || b > r.b1) return;
  h.offer(t, tmin, inst, prim);
}
// box.cu:36-53 -- slab test, tmin offered first, tmax only if tmin rejected
inline f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline void isect_box(const Box& x, f3 o, f3 inv, Real tmin, HitState& h, int inst, int prim) {
  f3 t0 = mulv(x.lo - o, inv);
  f3 t1 = mulv(x.hi - o, inv);
  Real tn = max_component(min_vec(t0, t1));
  Real tf = min_component(max_vec(t0, t1));
  if (tn <= tf) {
    if (!h.offer(tn, tmin, inst, prim)) h.offer(tf, tmin, inst, prim);
  }
}

// "lowbias32" integer hash (C. Wellons)
inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
// Free-flight sample of volume `vol` in instance `inst` for a ray whose path
// RNG state is `seed`.  The reference draws rnd(prd.seed) inside the
// intersection program (volume_sphere.cu:84), which depends on the visit
// order; the hash leaves the path's stream untouched.  Only the low 24 bits
// of the state enter (the bits every rnd() output depends on), so a path
// draws the same free flights whatever the state's top byte holds.
inline Real vol_xi(uint32_t seed, uint32_t vol, uint32_t inst) {
  const uint32_t h = mix32((seed & 0xFFFFFFu) ^ mix32(vol * 0x9E3779B9u ^ mix32(inst + 0x632BE5ABu)));
  return (Real)(h >> 8) * (1.0f / 16777216.0f);
}
// Free-flight seed of query ray i: a hash of the whole 64-bit index (vol_xi
// reads 24 bits of it), so rays 2^24 apart do not share their flights.
inline uint32_t query_vseed(uint64_t i) { return mix32((uint32_t)i ^ mix32((uint32_t)(i >> 32) + 0x7F4A7C15u)); }
// hit_boundary of volume_sphere.cu:21-55 / volume_box.cu:20-40
inline bool vol_boundary(const Volume& v, f3 o, f3 d, Real lo, Real hi, Real& rec) {
  if (v.shape == PT_VOLUME_BOX) {
    f3 t0 = (v.a - o) / d;
    f3 t1 = (v.b - o) / d;
    Real ta = max_component(min_vec(t0, t1));
    Real tb = min_component(max_vec(t0, t1));
    if (ta > tb) return false;
    if (ta < hi && ta > lo) { rec = ta; return true; }
    if (tb < hi && tb > lo) { rec = tb; return true; }
    return false;
  }
  f3 oc = o - v.a;
  Real a = dot(d, d);
  Real b = dot(oc, d);
  Real c = dot(oc, oc) - v.b.x * v.b.x;
  Real disc = b * b - a * c;
  if (disc < 0.f) return false;
  Real temp = (-b - std::sqrt(disc)) / a;
  if (temp < hi && temp > lo) { rec = temp; return true; }
  temp = (-b + std::sqrt(disc)) / a;
  if (temp < hi && temp > lo) { rec = temp; return true; }
  return false;
}
// volume_sphere.cu:63-110 / volume_box.cu:42-85 with the boundary test of
// Shirley's constant medium: the free flight -ln(xi) / density from the
// entry (clamped to tmin) must end before the exit.
inline void isect_volume(const Volume& v, f3 o, f3 d, Real tmin, Real xi, HitState& h, int inst, int prim) {
  Real r1 = 0.f, r2 = 0.f;
  if (!vol_boundary(v, o, d, -FLT_MAX, FLT_MAX, r1)) return;
  if (!vol_boundary(v, o, d, r1 + (Real)0.0001f, FLT_MAX, r2)) return;
  if (r1 < tmin) r1 = tmin;
  if (r1 >= r2) return;
  if (r1 < 0.f) r1 = 0.f;
  const Real len = length(d);
  const Real inside = (r2 - r1) * len;
  const Real hd = -(1.f / v.density) * plogf(xi);
  if (!(hd < inside)) return;
  h.offer(r1 + hd / len, tmin, inst, prim);
}

// cylinder.cu:47-66 fixed to intent (its b and c terms read
// d.x P.x + d.z + P.z and P.x^2 + P.z^2 + R^2, and it rejects t < delta):
// x^2 + z^2 = R^2 around the origin's vertical with 0 <= y - o.y <= L; the
// nearer valid root; delta < 1e-8 misses as in the reference.
inline void isect_cylinder(const Cylinder& cy, f3 o, f3 d, Real tmin, HitState& h, int inst, int prim) {
  f3 P0 = o - cy.o;
  Real a = d.x * d.x + d.z * d.z;
  Real b = d.x * P0.x + d.z * P0.z;
  Real c = P0.x * P0.x + P0.z * P0.z - cy.R * cy.R;
  Real delta = b * b - a * c;
  if (delta < 1e-8f) return;
  Real sq = std::sqrt(delta);
  Real t = (-b - sq) / a;
  Real y = P0.y + t * d.y;
  if (t > tmin && y >= 0.f && y <= cy.L) { h.offer(t, tmin, inst, prim); return; }
  t = (-b + sq) / a;
  y = P0.y + t * d.y;
  if (t > tmin && y >= 0.f && y <= cy.L) h.offer(t, tmin, inst, prim);
}

// Watertight ray/triangle test (Woop, Benthin, Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013), replacing the RTX hardware
// test and the Moller-Trumbore-style triangle.cu:26-58.  Edge functions that
// are exactly zero are recomputed in double precision.  Returns barycentric
// weights of v1 (b1) and v2 (b2).
struct TriRay {
  int kx, ky, kz;
  Real Sx, Sy, Sz;
};
inline Real comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
inline TriRay tri_ray_setup(f3 d) {
  TriRay r;
  Real ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
  r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
  if (comp(d, r.kz) < 0.0f) std::swap(r.kx, r.ky);
  r.Sx = comp(d, r.kx) / comp(d, r.kz);
  r.Sy = comp(d, r.ky) / comp(d, r.kz);
  r.Sz = 1.0f / comp(d, r.kz);
  return r;
}
// a * b = hi + lo exactly (Dekker's product over Veltkamp's 12-bit split; no
// fused multiply-add)
inline void two_prod(float a, float b, float& hi, float& lo) {
  const float sa = a * 4097.0f, sb = b * 4097.0f;
  const float ah = sa - (sa - a), al = a - ah;
  const float bh = sb - (sb - b), bl = b - bh;
  hi = a * b;
  lo = (((ah * bh - hi) + ah * bl) + al * bh) + al * bl;
}
// a * b - c * d with the exact sign and a nearly correctly rounded value: the
// edge functions of the watertight test when one of them rounds to zero
// (Woop et al. 2013 recompute them at higher precision there; the double
// build of the oracle needs no second pass)
inline float diff_of_products(float a, float b, float c, float d) {
  float h1, l1, h2, l2;
  two_prod(a, b, h1, l1);
  two_prod(c, d, h2, l2);
  const float s = h1 - h2, z = s - h1;
  const float t = (h1 - (s - z)) + (-h2 - z);
  return s + (t + (l1 - l2));
}
inline bool tri_test(const TriRay& r, f3 o, f3 v0, f3 v1, f3 v2, Real& t, Real& b1, Real& b2) {
  f3 A = v0 - o, B = v1 - o, C = v2 - o;
  Real Ax = comp(A, r.kx) - r.Sx * comp(A, r.kz);
  Real Ay = comp(A, r.ky) - r.Sy * comp(A, r.kz);
  Real Bx = comp(B, r.kx) - r.Sx * comp(B, r.kz);
  Real By = comp(B, r.ky) - r.Sy * comp(B, r.kz);
  Real Cx = comp(C, r.kx) - r.Sx * comp(C, r.kz);
  Real Cy = comp(C, r.ky) - r.Sy * comp(C, r.kz);
  Real U = Cx * By - Cy * Bx;
  Real V = Ax * Cy - Ay * Cx;
  Real W = Bx * Ay - By * Ax;
#ifndef PTREF_DOUBLE
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = diff_of_products(Cx, By, Cy, Bx);
    V = diff_of_products(Ax, Cy, Ay, Cx);
    W = diff_of_products(Bx, Ay, By, Ax);
  }
#endif
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  Real det = U + V + W;
  if (det == 0.0f) return false;
  Real Az = r.Sz * comp(A, r.kz);
  Real Bz = r.Sz * comp(B, r.kz);
  Real Cz = r.Sz * comp(C, r.kz);
  Real T = U * Az + V * Bz + W * Cz;
  Real rcp = 1.0f / det;
  t = T * rcp;
  b1 = V * rcp;
  b2 = W * rcp;
  return true;
}

// ------------------------------------------------------------------ LBVH
// 10-bit spread for a 30-bit Morton code (x in bit 2, y in bit 1, z in bit 0)
inline uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
inline uint32_t quantize10(Real c, Real lo, Real scale) {
  Real f = (c - lo) * scale;
  return f >= 1023.0f ? 1023u : (uint32_t)f;
}

struct Bvh4 {
  std::vector<uint32_t> words;   // 32 words per 128-byte node
  std::vector<uint32_t> order;   // sorted position -> local primitive
  std::vector<uint32_t> morton;  // codes in sorted order
  Aabb root;
  uint32_t num_nodes() const { return (uint32_t)(words.size() / 32); }
};

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
// Node boxes are stored as float: a wider Real (the double build) rounds
// them outward; for Real = float both are the identity.
inline float f_down(Real x) {
  float f = (float)x;
  return ((Real)f > x) ? std::nextafter(f, -std::numeric_limits<float>::infinity()) : f;
}
inline float f_up(Real x) {
  float f = (float)x;
  return ((Real)f < x) ? std::nextafter(f, std::numeric_limits<float>::infinity()) : f;
}

// Build: bounds -> centroids -> Morton codes -> stable sort -> Karras binary
// hierarchy -> bottom-up bounds -> breadth-first collapse to 4-wide nodes.
inline Bvh4 build_bvh4(const std::vector<Aabb>& prim, int leaf_max) {
  Bvh4 out;
  const uint32_t n = (uint32_t)prim.size();
  if (n == 0) throw std::runtime_error("empty bottom-level structure");
  std::vector<f3> cen(n);
  Aabb cb;
  for (uint32_t i = 0; i < n; i++) {
    cen[i] = mk((prim[i].lo.x + prim[i].hi.x) * 0.5f, (prim[i].lo.y + prim[i].hi.y) * 0.5f,
                (prim[i].lo.z + prim[i].hi.z) * 0.5f);
    Aabb p; p.lo = cen[i]; p.hi = cen[i];
    cb = merge(cb, p);
    out.root = merge(out.root, prim[i]);
  }
  Real ex = cb.hi.x - cb.lo.x, ey = cb.hi.y - cb.lo.y, ez = cb.hi.z - cb.lo.z;
  Real sx = ex > 0.0f ? 1024.0f / ex : 0.0f;
  Real sy = ey > 0.0f ? 1024.0f / ey : 0.0f;
  Real sz = ez > 0.0f ? 1024.0f / ez : 0.0f;
  std::vector<uint32_t> code(n);
  for (uint32_t i = 0; i < n; i++) {
    code[i] = (spread10(quantize10(cen[i].x, cb.lo.x, sx)) << 2) |
              (spread10(quantize10(cen[i].y, cb.lo.y, sy)) << 1) |
              spread10(quantize10(cen[i].z, cb.lo.z, sz));
  }
  out.order.resize(n);
  for (uint32_t i = 0; i < n; i++) out.order[i] = i;
  std::stable_sort(out.order.begin(), out.order.end(),
                   [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
  out.morton.resize(n);
  for (uint32_t i = 0; i < n; i++) out.morton[i] = code[out.order[i]];

  // Binary hierarchy over augmented keys (code << 32 | sorted index).
  // Node refs: internal m -> m, leaf l -> (1u << 31) | l.
  const uint32_t LEAF = 1u << 31;
  uint32_t ni = n - 1;
  std::vector<uint32_t> left(ni), right(ni), first(ni), last(ni);
  auto key = [&](int64_t i) { return ((uint64_t)out.morton[i] << 32) | (uint64_t)i; };
  auto delta = [&](int64_t i, int64_t j) -> int {
    if (j < 0 || j > (int64_t)n - 1) return -1;
    return __builtin_clzll(key(i) ^ key(j));
  };
  for (int64_t i = 0; i < (int64_t)ni; i++) {
    int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(i, i - d);
    int64_t lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (delta(i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(i, j);
    int64_t s = 0;
    int64_t div = 2;
    int64_t t;
    do {
      t = (l + div - 1) / div;
      if (delta(i, i + (s + t) * d) > dnode) s += t;
      div *= 2;
    } while (t > 1);
    int64_t gamma = i + s * d + std::min(d, 0);
    int64_t lo = std::min(i, j), hi = std::max(i, j);
    left[i] = (lo == gamma) ? (LEAF | (uint32_t)gamma) : (uint32_t)gamma;
    right[i] = (hi == gamma + 1) ? (LEAF | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    first[i] = (uint32_t)lo;
    last[i] = (uint32_t)hi;
  }
  // bounds of every binary node (min/max are exact, order-independent)
  std::vector<Aabb> nb(ni);
  std::vector<Aabb> lb(n);
  for (uint32_t i = 0; i < n; i++) lb[i] = prim[out.order[i]];
  auto ref_box = [&](uint32_t r) -> const Aabb& { return (r & LEAF) ? lb[r & ~LEAF] : nb[r]; };
  // children of internal node i have larger index or are leaves?  Not in
  // general; compute by post-order recursion with an explicit stack.
  if (ni > 0) {
    std::vector<uint32_t> st;
    std::vector<uint8_t> done(ni, 0);
    st.push_back(0);
    while (!st.empty()) {
      uint32_t m = st.back();
      bool ready = true;
      for (uint32_t c : {left[m], right[m]})
        if (!(c & LEAF) && !done[c]) { st.push_back(c); ready = false; }
      if (!ready) continue;
      st.pop_back();
      if (done[m]) continue;
      nb[m] = merge(ref_box(left[m]), ref_box(right[m]));
      done[m] = 1;
    }
  }
  auto ref_first = [&](uint32_t r) { return (r & LEAF) ? (r & ~LEAF) : first[r]; };
  auto ref_size = [&](uint32_t r) { return (r & LEAF) ? 1u : (last[r] - first[r] + 1); };

  // Breadth-first collapse.
  auto empty_node = [&]() {
    size_t base = out.words.size();
    out.words.resize(base + 32, 0);
    for (int c = 0; c < 4; c++) {
      out.words[base + 0 + c] = fbits(INF);
      out.words[base + 4 + c] = fbits(-INF);
      out.words[base + 8 + c] = fbits(INF);
      out.words[base + 12 + c] = fbits(-INF);
      out.words[base + 16 + c] = fbits(INF);
      out.words[base + 20 + c] = fbits(-INF);
      out.words[base + 24 + c] = 0xFFFFFFFFu;
      out.words[base + 28 + c] = 0;
    }
    return base;
  };
  auto set_slot = [&](size_t base, int c, const Aabb& b, uint32_t child, uint32_t count) {
    out.words[base + 0 + c] = fbits(f_down(b.lo.x));
    out.words[base + 4 + c] = fbits(f_up(b.hi.x));
    out.words[base + 8 + c] = fbits(f_down(b.lo.y));
    out.words[base + 12 + c] = fbits(f_up(b.hi.y));
    out.words[base + 16 + c] = fbits(f_down(b.lo.z));
    out.words[base + 20 + c] = fbits(f_up(b.hi.z));
    out.words[base + 24 + c] = child;
    out.words[base + 28 + c] = count;
  };
  if (n == 1 || ref_size(0) <= (uint32_t)leaf_max) {
    size_t base = empty_node();
    set_slot(base, 0, out.root, 0, n);
    return out;
  }
  std::vector<uint32_t> queue{0};  // binary node of each wide node, BFS order
  for (size_t w = 0; w < queue.size(); w++) {
    uint32_t m = queue[w];
    std::vector<uint32_t> slots{left[m], right[m]};
    while (slots.size() < 4) {
      int best = -1;
      Real best_area = -1.0f;
      for (size_t s = 0; s < slots.size(); s++) {
        uint32_t r = slots[s];
        if ((r & LEAF) || ref_size(r) <= (uint32_t)leaf_max) continue;
        Real a = half_area(nb[r]);
        if (a > best_area) { best_area = a; best = (int)s; }
      }
      if (best < 0) break;
      uint32_t r = slots[best];
      slots[best] = left[r];
      slots.insert(slots.begin() + best + 1, right[r]);
    }
    size_t base = empty_node();
    for (size_t s = 0; s < slots.size(); s++) {
      uint32_t r = slots[s];
      if ((r & LEAF) || ref_size(r) <= (uint32_t)leaf_max) {
        set_slot(base, (int)s, ref_box(r), ref_first(r), ref_size(r));
      } else {
        uint32_t child = (uint32_t)queue.size();
        queue.push_back(r);
        set_slot(base, (int)s, nb[r], child, 0);
      }
    }
  }
  return out;
}

// ------------------------------------------------------------------ scene
struct Instance {
  int kind;
  uint32_t first, count;
  bool xf;
  Real o2w[12], w2o[12];
  int blas;
};
struct Blas {
  int kind;
  uint32_t first, count;  // prims of the kind (mesh: mesh index, count tris)
  Bvh4 bvh;
};

struct Material { int type; int tex[2]; Real p[4]; };
struct Texture { int type; f3 rgb; uint32_t w = 0, h = 0; std::vector<float> texels; };
struct Light { int type, axis; Real a0, a1, b0, b1, k; f3 c; Real r; f3 e; };

struct Counts {
  uint64_t nodes = 0, prims = 0, insts = 0;
};

// Importance-sampling tables of an environment map (no reference
// counterpart; the reference only looks the map up on a miss, miss.cu:64-95).
// Texel weight = luminance x sin(theta) in 64-bit fixed point; the tables are
// inclusive prefix sums per row (cond) and over row totals (marg).  Integer
// prefix sums are exact, so any summation order gives the same tables.
struct EnvCdf {
  uint32_t W = 0, H = 0;
  std::vector<uint64_t> cond;  // H x W
  std::vector<uint64_t> marg;  // H
  uint64_t total = 0;
  int tex = -1;
};

// Texel weight of row j (theta at the row centre), FP64 luminance (Rec. 709
// weights) times sin(theta) from the shared Real libm.
inline double env_texel_weight(const float* t, uint32_t j, uint32_t H) {
  double lum = 0.2126 * (double)t[0] + 0.7152 * (double)t[1] + 0.0722 * (double)t[2];
  Real st = psin(PI_F * (((Real)j + 0.5f) / (Real)H));
  double w = lum * (double)st;
  return (w > 0.0) ? w : 0.0;  // NaN and negative texels get no weight
}
// Bits per quantised texel weight: 40, fewer for maps of more than 2^23
// texels, so that the total of W * H weights stays below 2^63.
inline int env_weight_bits(uint32_t W, uint32_t H) {
  const uint64_t n = (uint64_t)W * H;
  int lg = 0;
  while (lg < 63 && (1ull << lg) < n) lg++;
  return std::min(40, 63 - lg);
}
// weights are scaled by 2^(B - e), e the binary exponent of the largest
// weight and B = env_weight_bits, so every texel fits in B bits and the
// total below 2^63
inline void build_env_cdf(const Texture& t, int tex, EnvCdf& e) {
  e.W = t.w;
  e.H = t.h;
  e.tex = tex;
  double mx = 0.0;
  for (uint32_t j = 0; j < e.H; j++)
    for (uint32_t i = 0; i < e.W; i++) mx = std::fmax(mx, env_texel_weight(&t.texels[4 * ((size_t)j * e.W + i)], j, e.H));
  int ex = 0;
  std::frexp(mx, &ex);
  double scale = std::ldexp(1.0, env_weight_bits(e.W, e.H) - ex);
  e.cond.assign((size_t)e.W * e.H, 0);
  e.marg.assign(e.H, 0);
  uint64_t run = 0;
  for (uint32_t j = 0; j < e.H; j++) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < e.W; i++) {
      double w = env_texel_weight(&t.texels[4 * ((size_t)j * e.W + i)], j, e.H);
      acc += (mx > 0.0) ? (uint64_t)(w * scale) : 0;
      e.cond[(size_t)j * e.W + i] = acc;
    }
    run += acc;
    e.marg[j] = run;
  }
  e.total = run;
}

struct Scene {
  std::vector<Sphere> spheres;
  std::vector<Rect> rects;
  std::vector<Box> boxes;
  std::vector<Volume> volumes;
  std::vector<Cylinder> cylinders;
  std::vector<MovingSphere> moving;
  std::vector<Mesh> meshes;
  std::vector<Instance> insts;
  std::vector<Blas> blas;
  std::vector<Material> mats;
  std::vector<Texture> texs;
  std::vector<Light> lights;
  std::vector<Light> lights_plain;  // without environment lights
  pt_miss_desc miss{};
  EnvCdf env;
  Bvh4 tlas;

  explicit Scene(const pt_scene_desc& d) {
    for (uint32_t i = 0; i < d.num_spheres; i++) {
      auto& s = d.spheres[i];
      spheres.push_back({mk(s.center[0], s.center[1], s.center[2]), s.radius, s.material});
    }
    for (uint32_t i = 0; i < d.num_rects; i++) {
      auto& r = d.rects[i];
      rects.push_back({r.axis, r.flip, r.a0, r.a1, r.b0, r.b1, r.k, r.material});
    }
    for (uint32_t i = 0; i < d.num_boxes; i++) {
      auto& b = d.boxes[i];
      boxes.push_back({mk(b.pmin[0], b.pmin[1], b.pmin[2]), mk(b.pmax[0], b.pmax[1], b.pmax[2]), b.material});
    }
    for (uint32_t i = 0; i < d.num_volumes; i++) {
      auto& v = d.volumes[i];
      volumes.push_back({v.shape, mk(v.a[0], v.a[1], v.a[2]), mk(v.b[0], v.b[1], v.b[2]), v.density, v.material});
    }
    for (uint32_t i = 0; i < d.num_cylinders; i++) {
      auto& c = d.cylinders[i];
      cylinders.push_back({mk(c.origin[0], c.origin[1], c.origin[2]), c.length, c.radius, c.material});
    }
    for (uint32_t i = 0; i < d.num_moving_spheres; i++) {
      auto& m = d.moving_spheres[i];
      moving.push_back({mk(m.center0[0], m.center0[1], m.center0[2]), mk(m.center1[0], m.center1[1], m.center1[2]),
                        m.radius, m.time0, m.time1, m.material});
    }
    for (uint32_t i = 0; i < d.num_meshes; i++) {
      auto& m = d.meshes[i];
      Mesh me;
      for (uint32_t v = 0; v < m.num_vertices; v++) {
        me.v.push_back(mk(m.vertices[3 * v], m.vertices[3 * v + 1], m.vertices[3 * v + 2]));
        if (m.normals) me.n.push_back(mk(m.normals[3 * v], m.normals[3 * v + 1], m.normals[3 * v + 2]));
        if (m.uvs) me.uv.push_back(f2{m.uvs[2 * v], m.uvs[2 * v + 1]});
      }
      me.idx.assign(m.indices, m.indices + 3 * (size_t)m.num_triangles);
      if (m.face_materials) me.face_mat.assign(m.face_materials, m.face_materials + m.num_triangles);
      if (m.face_indices) me.face_idx.assign(m.face_indices, m.face_indices + m.num_triangles);
      me.mat = m.material;
      meshes.push_back(std::move(me));
    }
    for (uint32_t i = 0; i < d.num_materials; i++) {
      auto& m = d.materials[i];
      mats.push_back({m.type, {m.tex[0], m.tex[1]}, {m.param[0], m.param[1], m.param[2], m.param[3]}});
    }
    for (uint32_t i = 0; i < d.num_textures; i++) {
      auto& t = d.textures[i];
      Texture tx;
      tx.type = t.type;
      tx.rgb = mk(t.rgb[0], t.rgb[1], t.rgb[2]);
      if (t.type == PT_TEX_IMAGE) {
        tx.w = t.width; tx.h = t.height;
        tx.texels.assign(t.texels, t.texels + 4 * (size_t)t.width * t.height);
      } else if (t.type == PT_TEX_CHECKER) {
        tx.w = t.width; tx.h = t.height;  // odd, even
      } else if (t.type == PT_TEX_NOISE) {
        tx.texels.assign(t.texels, t.texels + PT_NOISE_FLOATS);
      } else if (t.type == PT_TEX_GRADIENT) {
        tx.texels.assign(t.texels, t.texels + 12);
      } else if (t.type == PT_TEX_VECTOR) {
        tx.w = t.width;  // child indices
        tx.texels.assign(t.texels, t.texels + t.width);
      }
      texs.push_back(std::move(tx));
    }
    for (uint32_t i = 0; i < d.num_lights; i++) {
      auto& l = d.lights[i];
      lights.push_back({l.type, l.axis, l.a0, l.a1, l.b0, l.b1, l.k, mk(l.center[0], l.center[1], l.center[2]),
                        l.radius, mk(l.emission[0], l.emission[1], l.emission[2])});
    }
    miss = d.miss;
    for (const Light& l : lights) {
      if (l.type != PT_LIGHT_ENV) { lights_plain.push_back(l); continue; }
      if (miss.type != PT_MISS_HDR || !miss.spherical || miss.texture < 0 ||
          miss.texture >= (int)texs.size() || texs[miss.texture].type != PT_TEX_IMAGE)
        throw std::runtime_error("an environment light needs a spherical HDR miss texture");
    }
    if (lights.size() != lights_plain.size()) build_env_cdf(texs[miss.texture], miss.texture, env);
    // bottom-level structures, shared between identical (kind, first, count)
    for (uint32_t i = 0; i < d.num_instances; i++) {
      auto& in = d.instances[i];
      Instance I;
      I.kind = in.kind; I.first = in.first; I.count = in.count; I.xf = in.has_transform != 0;
      for (int k = 0; k < 12; k++) {
        I.o2w[k] = in.obj_to_world[k];
        I.w2o[k] = in.world_to_obj[k];
      }
      if (I.kind == PT_GEOM_MESH) I.count = meshes.at(I.first).ntri();
      I.blas = -1;
      for (size_t b = 0; b < blas.size(); b++)
        if (blas[b].kind == I.kind && blas[b].first == I.first && blas[b].count == I.count) I.blas = (int)b;
      if (I.blas < 0) {
        Blas B;
        B.kind = I.kind; B.first = I.first; B.count = I.count;
        std::vector<Aabb> pb(I.count);
        for (uint32_t p = 0; p < I.count; p++) pb[p] = prim_bounds(I.kind, I.first, p);
        B.bvh = build_bvh4(pb, BLAS_LEAF_MAX);
        I.blas = (int)blas.size();
        blas.push_back(std::move(B));
      }
      insts.push_back(I);
    }
    if (insts.empty()) throw std::runtime_error("scene has no instances");
    std::vector<Aabb> ib(insts.size());
    for (size_t i = 0; i < insts.size(); i++) ib[i] = instance_bounds(insts[i]);
    tlas = build_bvh4(ib, TLAS_LEAF_MAX);
  }

  Aabb prim_bounds(int kind, uint32_t first, uint32_t p) const {
    switch (kind) {
      case PT_GEOM_SPHERES: return widen(bounds(spheres.at(first + p)));
      case PT_GEOM_RECTS: return widen(bounds(rects.at(first + p)));
      case PT_GEOM_BOXES: return widen(bounds(boxes.at(first + p)));
      case PT_GEOM_VOLUMES: return widen(bounds(volumes.at(first + p)));
      case PT_GEOM_CYLINDERS: return widen(bounds(cylinders.at(first + p)));
      case PT_GEOM_MOVING_SPHERES: return widen(bounds(moving.at(first + p)));
      default: {
        const Mesh& m = meshes.at(first);
        return widen(bounds_tri(m.v.at(m.idx[3 * p]), m.v.at(m.idx[3 * p + 1]), m.v.at(m.idx[3 * p + 2])));
      }
    }
  }
  // world bounds of an instance: its root box's 8 corners transformed
  Aabb instance_bounds(const Instance& I) const {
    const Aabb& r = blas[I.blas].bvh.root;
    if (!I.xf) return r;
    Aabb w;
    for (int c = 0; c < 8; c++) {
      f3 p = mk((c & 1) ? r.hi.x : r.lo.x, (c & 2) ? r.hi.y : r.lo.y, (c & 4) ? r.hi.z : r.lo.z);
      f3 q = xform_point(I.o2w, p);
      Aabb pb; pb.lo = q; pb.hi = q;
      w = merge(w, pb);
    }
    return widen(w);
  }

  // object-space ray of an instance (identity instances are not transformed)
  void to_object(const Instance& I, f3 o, f3 d, f3& oo, f3& od) const {
    if (!I.xf) { oo = o; od = d; return; }
    oo = xform_point(I.w2o, o);
    od = xform_vector(I.w2o, d);
  }

  // `vseed`: the ray's RNG state, for volume free flights; `time`: the
  // ray's time, for moving spheres
  void isect_prim(const Instance& I, int inst, uint32_t local, f3 o, f3 d, const TriRay& tr, Real tmin,
                  HitState& h, uint32_t vseed, Real time) const {
    uint32_t g = I.first + local;
    switch (I.kind) {
      case PT_GEOM_SPHERES: isect_sphere(spheres[g], o, d, tmin, h, inst, (int)g); break;
      case PT_GEOM_RECTS: isect_rect(rects[g], o, d, mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z), tmin, h, inst, (int)g); break;
      case PT_GEOM_BOXES: isect_box(boxes[g], o, mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z), tmin, h, inst, (int)g); break;
      case PT_GEOM_VOLUMES:
        isect_volume(volumes[g], o, d, tmin, vol_xi(vseed, g, (uint32_t)inst), h, inst, (int)g);
        break;
      case PT_GEOM_CYLINDERS: isect_cylinder(cylinders[g], o, d, tmin, h, inst, (int)g); break;
      case PT_GEOM_MOVING_SPHERES: {
        const MovingSphere& m = moving[g];
        isect_sphere(Sphere{moving_center(m, time), m.r, m.mat}, o, d, tmin, h, inst, (int)g);
        break;
      }
      default: {
        const Mesh& m = meshes[I.first];
        Real t, b1, b2;
        if (tri_test(tr, o, m.v[m.idx[3 * local]], m.v[m.idx[3 * local + 1]], m.v[m.idx[3 * local + 2]], t, b1, b2))
          h.offer(t, tmin, inst, (int)local);
      }
    }
  }

  // slab test of one child slot; returns entry distance or +inf on a miss
  static Real slab(const uint32_t* w, int c, f3 o, f3 inv, Real tmin, Real tmax) {
    Real t0x = (bitsf(w[0 + c]) - o.x) * inv.x, t1x = (bitsf(w[4 + c]) - o.x) * inv.x;
    Real t0y = (bitsf(w[8 + c]) - o.y) * inv.y, t1y = (bitsf(w[12 + c]) - o.y) * inv.y;
    Real t0z = (bitsf(w[16 + c]) - o.z) * inv.z, t1z = (bitsf(w[20 + c]) - o.z) * inv.z;
    Real tn = std::fmax(std::fmax(std::fmin(t0x, t1x), std::fmin(t0y, t1y)), std::fmax(std::fmin(t0z, t1z), tmin));
    Real tf = std::fmin(std::fmin(std::fmax(t0x, t1x), std::fmax(t0y, t1y)), std::fmin(std::fmax(t0z, t1z), tmax));
    return (tn <= tf * SLAB_FAR_SCALE) ? tn : INF;
  }

  // Closest hit through the two-level BVH4.  Children are visited nearest
  // first (entry distance, then slot); popped entries farther than the
  // current hit are skipped.
  HitState closest(f3 o, f3 d, Real tmin, Real tmax, Counts* cnt = nullptr, uint32_t vseed = 0,
                   Real time = 0.f) const {
    HitState h;
    h.t = tmax;
    struct Entry { uint32_t ref; Real tn; };
    const uint32_t LEAF = 1u << 31;
    std::vector<Entry> st;
    st.reserve(128);
    st.push_back({0, tmin});
    int cur = -1;       // current instance (-1: top level)
    size_t base_sp = 0; // stack depth when the instance was entered
    f3 wo = o, wd = d, winv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    f3 ro = wo, rd = wd, rinv = winv;
    TriRay tr{};
    while (true) {
      if (cur >= 0 && st.size() == base_sp) { cur = -1; ro = wo; rd = wd; rinv = winv; }
      if (st.empty()) break;
      Entry e = st.back();
      st.pop_back();
      if (e.tn > h.t) continue;
      const Bvh4& bvh = cur < 0 ? tlas : blas[insts[cur].blas].bvh;
      if (e.ref & LEAF) {
        uint32_t first = e.ref & 0x0FFFFFFFu;
        uint32_t count = ((e.ref >> 28) & 7u) + 1;
        if (cur < 0) {
          // top-level leaf: enter the instance
          uint32_t inst = tlas.order[first];
          const Instance& I = insts[inst];
          if (cnt) cnt->insts++;
          to_object(I, wo, wd, ro, rd);
          rinv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
          if (I.kind == PT_GEOM_MESH) tr = tri_ray_setup(rd);
          cur = (int)inst;
          base_sp = st.size();
          st.push_back({0, e.tn});
        } else {
          const Instance& I = insts[cur];
          for (uint32_t k = 0; k < count; k++) {
            if (cnt) cnt->prims++;
            isect_prim(I, cur, bvh.order[first + k], ro, rd, tr, tmin, h, vseed, time);
          }
        }
        continue;
      }
      if (cnt) cnt->nodes++;
      const uint32_t* w = &bvh.words[(size_t)e.ref * 32];
      Entry hitc[4];
      int nh = 0;
      for (int c = 0; c < 4; c++) {
        if (w[24 + c] == 0xFFFFFFFFu) continue;
        Real tn = slab(w, c, ro, rinv, tmin, h.t);
        if (tn == INF) continue;
        uint32_t cnt_c = w[28 + c];
        uint32_t ref = cnt_c ? (LEAF | ((cnt_c - 1) << 28) | w[24 + c]) : w[24 + c];
        // insertion by (tn, slot): stable
        int p = nh;
        while (p > 0 && hitc[p - 1].tn > tn) { hitc[p] = hitc[p - 1]; p--; }
        hitc[p] = {ref, tn};
        nh++;
      }
      for (int k = nh - 1; k >= 0; k--) st.push_back(hitc[k]);
    }
    return h;
  }

  // Brute force over every instance and primitive.
  HitState brute(f3 o, f3 d, Real tmin, Real tmax, uint32_t vseed = 0, Real time = 0.f) const {
    HitState h;
    h.t = tmax;
    for (size_t i = 0; i < insts.size(); i++) {
      const Instance& I = insts[i];
      f3 ro, rd;
      to_object(I, o, d, ro, rd);
      TriRay tr{};
      if (I.kind == PT_GEOM_MESH) tr = tri_ray_setup(rd);
      for (uint32_t p = 0; p < I.count; p++) isect_prim(I, (int)i, p, ro, rd, tr, tmin, h, vseed, time);
    }
    return h;
  }

  int material_of(int inst, int prim) const {
    const Instance& I = insts[inst];
    switch (I.kind) {
      case PT_GEOM_SPHERES: return spheres[prim].mat;
      case PT_GEOM_RECTS: return rects[prim].mat;
      case PT_GEOM_BOXES: return boxes[prim].mat;
      case PT_GEOM_VOLUMES: return volumes[prim].mat;
      case PT_GEOM_CYLINDERS: return cylinders[prim].mat;
      case PT_GEOM_MOVING_SPHERES: return moving[prim].mat;
      default: {
        const Mesh& m = meshes[I.first];
        return m.face_mat.empty() ? m.mat : m.face_mat[prim];
      }
    }
  }
};

}  // namespace ptref
