// ptref_render.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Scalar CPU restatement of the reference's shading and integrator programs:
// hit-record reconstruction (programs/hitables/*.cu Get_HitRecord), the BRDF
// sample/eval/pdf of programs/materials/*, light sampling and MIS
// (programs/materials/light_sample.cuh, programs/pdfs/*.cu), miss programs
// (programs/miss.cu), the bounce loop (programs/raygen.cu:66-123), per-pixel
// accumulation and the display mapping (raygen.cu:126-168).
//
// Decisions shared with the GPU path where the reference is unspecified:
//  * rnd() calls inside one argument list are evaluated left to right.
//  * Hit records are rebuilt in object space from the object-space ray and
//    then transformed (fixes the double transform of box.cu / sphere.cu).
//  * Shadow rays: lit iff the closest hit in (1e-3, inf) is a diffuse light
//    (and dot(N, Wi) > 0), or there is no hit (light_sample.cuh:61-72,
//    hit.cu:16-24 made order-independent).
//  * A pdf the reference leaves unset (Evaluate returning early) reads as 1.
//  * Russian roulette applies when depth >= rr_min_depth (reference: 11).
#pragma once

#include "ptref_scene.hpp"

namespace ptref {

struct HitRec {
  f3 P, N, Wo;
  Real u = 0.f, v = 0.f;
  int mat;
  int fidx = 0;  // the face index a vector texture selects by (0 off meshes)
};

// sphere.cu:65-93, aarect.cu:71-115, box.cu:24-33 + 56-81, triangle.cu:85-149
inline HitRec hit_record(const Scene& s, int inst, int prim, Real t, f3 o, f3 d, Real time = 0.f) {
  const Instance& I = s.insts[inst];
  f3 ro, rd;
  s.to_object(I, o, d, ro, rd);
  HitRec r;
  r.Wo = normalize(-d);
  r.mat = s.material_of(inst, prim);
  f3 hp, n;
  switch (I.kind) {
    case PT_GEOM_SPHERES: {
      const Sphere& sp = s.spheres[prim];
      hp = ro + t * rd;
      n = (hp - sp.c) / sp.r;
      Real phi = patan2(n.z, n.x);
      Real theta = pasin(n.y);
      r.u = 1.f - (phi + PI_F) / (2.f * PI_F);
      r.v = (theta + PI_F / 2.f) / PI_F;
      break;
    }
    case PT_GEOM_RECTS: {
      const Rect& q = s.rects[prim];
      hp = ro + t * rd;
      if (q.axis == PT_AXIS_X) { n = mk(1.f, 0.f, 0.f); r.u = (hp.y - q.a0) / (q.a1 - q.a0); r.v = (hp.z - q.b0) / (q.b1 - q.b0); }
      else if (q.axis == PT_AXIS_Y) { n = mk(0.f, 1.f, 0.f); r.u = (hp.x - q.a0) / (q.a1 - q.a0); r.v = (hp.z - q.b0) / (q.b1 - q.b0); }
      else { n = mk(0.f, 0.f, 1.f); r.u = (hp.x - q.a0) / (q.a1 - q.a0); r.v = (hp.y - q.b0) / (q.b1 - q.b0); }
      if (q.flip) n = -n;
      break;
    }
    case PT_GEOM_BOXES: {
      const Box& b = s.boxes[prim];
      hp = ro + t * rd;
      f3 rinv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
      f3 t0 = mulv(b.lo - ro, rinv);
      f3 t1 = mulv(b.hi - ro, rinv);
      f3 neg = mk(t == t0.x ? 1.f : 0.f, t == t0.y ? 1.f : 0.f, t == t0.z ? 1.f : 0.f);
      f3 pos = mk(t == t1.x ? 1.f : 0.f, t == t1.y ? 1.f : 0.f, t == t1.z ? 1.f : 0.f);
      n = pos - neg;
      break;
    }
    case PT_GEOM_CYLINDERS: {  // cylinder.cu:70-97: radial normal, no uv
      const Cylinder& cy = s.cylinders[prim];
      hp = ro + t * rd;
      n = normalize(hp - mk(cy.o.x, hp.y, cy.o.z));
      break;
    }
    case PT_GEOM_MOVING_SPHERES: {  // moving_sphere.cu:78-91, in object space at the ray's time
      const MovingSphere& m = s.moving[prim];
      f3 c = moving_center(m, time);
      hp = ro + t * rd;
      n = (hp - c) / m.r;
      Real phi = patan2(n.z, n.x);
      Real theta = pasin(n.y);
      r.u = 1.f - (phi + PI_F) / (2.f * PI_F);
      r.v = (theta + PI_F / 2.f) / PI_F;
      break;
    }
    case PT_GEOM_VOLUMES: {  // volume_sphere.cu:88-100: normal +X, no uv
      hp = ro + t * rd;
      n = mk(1.f, 0.f, 0.f);
      break;
    }
    default: {
      const Mesh& m = s.meshes[I.first];
      uint32_t i0 = m.idx[3 * prim], i1 = m.idx[3 * prim + 1], i2 = m.idx[3 * prim + 2];
      f3 a = m.v[i0], b = m.v[i1], c = m.v[i2];
      TriRay tr = tri_ray_setup(rd);
      Real tt = 0.f, b1 = 0.f, b2 = 0.f;
      tri_test(tr, ro, a, b, c, tt, b1, b2);
      Real b0 = 1.f - b1 - b2;
      hp = a * b0 + b * b1 + c * b2;
      n = cross(b - a, c - a);
      if (!m.n.empty()) {
        f3 ns = m.n[i0] * b0 + m.n[i1] * b1 + m.n[i2] * b2;
        n = ns;
      }
      if (!m.uv.empty()) {
        r.u = m.uv[i0].x * b0 + m.uv[i1].x * b1 + m.uv[i2].x * b2;
        r.v = m.uv[i0].y * b0 + m.uv[i1].y * b1 + m.uv[i2].y * b2;
      }
      if (!m.face_idx.empty()) r.fidx = m.face_idx[prim];
      break;
    }
  }
  if (I.xf) {
    r.P = xform_point(I.o2w, hp);
    r.N = normalize(xform_normal(I.w2o, n));
  } else {
    r.P = hp;
    r.N = normalize(n);
  }
  return r;
}

// normal_shader.cu:27-31: the geometric normal differs from the hit
// record's only on meshes with vertex normals
inline f3 geometric_normal(const Scene& s, int inst, int prim, f3 n_shading) {
  const Instance& I = s.insts[inst];
  if (I.kind != PT_GEOM_MESH) return n_shading;
  const Mesh& m = s.meshes[I.first];
  if (m.n.empty()) return n_shading;
  f3 a = m.v[m.idx[3 * prim]], b = m.v[m.idx[3 * prim + 1]], c = m.v[m.idx[3 * prim + 2]];
  f3 n = cross(b - a, c - a);
  if (I.xf) n = xform_normal(I.w2o, n);
  return normalize(n);
}

// Bilinear lookup with repeat wrap, texel centres at (i + 0.5) / w
// (textures.hpp:177-230 sampler settings; image_texture.cu:5-7).
// Coordinates that are NaN or beyond 2^62 texels (a NaN direction reaching
// a miss program) give a NaN colour, which de_nan removes and the render
// statistics count.
inline f3 tex_image(const Texture& t, Real u, Real v) {
  Real x = u * (Real)t.w - 0.5f, y = v * (Real)t.h - 0.5f;
  if (!(std::fabs(x) < (Real)0x1p62f && std::fabs(y) < (Real)0x1p62f))
    return mk(std::numeric_limits<Real>::quiet_NaN());
  Real x0 = std::floor(x), y0 = std::floor(y);
  Real fx = x - x0, fy = y - y0;
  auto wrap = [](int64_t i, int64_t n) { int64_t r = i % n; return r < 0 ? r + n : r; };
  int64_t i0 = wrap((int64_t)x0, t.w), i1 = wrap((int64_t)x0 + 1, t.w);
  int64_t j0 = wrap((int64_t)y0, t.h), j1 = wrap((int64_t)y0 + 1, t.h);
  auto px = [&](int64_t i, int64_t j) {
    const float* p = &t.texels[4 * (size_t)(j * t.w + i)];
    return mk(p[0], p[1], p[2]);
  };
  f3 top = px(i0, j0) * (1.f - fx) + px(i1, j0) * fx;
  f3 bot = px(i0, j1) * (1.f - fx) + px(i1, j1) * fx;
  return top * (1.f - fy) + bot * fy;
}
// noise_texture.cu:13-54 perlin_interp / noise: tables as pt_api.h gives
// them (256 gradient vectors (x, y, z, 0), then perm_x, perm_y, perm_z)
inline Real perlin_noise(const std::vector<float>& tab, f3 p) {
  Real u = p.x - std::floor(p.x), v = p.y - std::floor(p.y), w = p.z - std::floor(p.z);
  int i = (int)std::floor(p.x), j = (int)std::floor(p.y), k = (int)std::floor(p.z);
  auto perm = [&](int axis, int idx) { return (int)tab[1024 + 256 * axis + idx]; };
  Real uu = u * u * (3 - 2 * u), vv = v * v * (3 - 2 * v), ww = w * w * (3 - 2 * w);
  Real accum = 0;
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++)
      for (int dk = 0; dk < 2; dk++) {
        int h = perm(0, (i + di) & 255) ^ perm(1, (j + dj) & 255) ^ perm(2, (k + dk) & 255);
        f3 c = mk(tab[4 * h], tab[4 * h + 1], tab[4 * h + 2]);
        f3 wv = mk(u - (Real)di, v - (Real)dj, w - (Real)dk);
        accum += ((Real)di * uu + (Real)(1 - di) * (1 - uu)) * ((Real)dj * vv + (Real)(1 - dj) * (1 - vv)) *
                 ((Real)dk * ww + (Real)(1 - dk) * (1 - ww)) * dot(c, wv);
      }
  return accum;
}
// noise_texture.cu:56-68 turb
inline Real perlin_turb(const std::vector<float>& tab, f3 p) {
  Real accum = 0, weight = 1.0f;
  f3 tp = p;
  for (int i = 0; i < 7; i++) {
    accum += weight * perlin_noise(tab, tp);
    weight *= 0.5f;
    tp = tp * 2.f;
  }
  return std::fabs(accum);
}
inline f3 tex_leaf(const Texture& t, Real u, Real v, f3 P) {
  switch (t.type) {
    case PT_TEX_CONSTANT: return t.rgb;
    case PT_TEX_NOISE: {
      // noise_texture.cu:70-84: the axis switch falls through to the Z case
      Real sv = psin(t.rgb.x * P.z + 5 * perlin_turb(t.texels, t.rgb.x * P));
      return mk(1.f) * 0.5f * (1 + sv);
    }
    case PT_TEX_GRADIENT: {  // gradient_texture.cu:7-14
      const std::vector<float>& c = t.texels;
      f3 ud = normalize(P);
      Real x = std::fabs(ud.x), y = std::fabs(ud.y), z = std::fabs(ud.z);
      return x * mk(c[0], c[1], c[2]) + y * mk(c[4], c[5], c[6]) + z * mk(c[8], c[9], c[10]);
    }
    default: return tex_image(t, u, v);
  }
}
// the texture at (u, v) and the world hit point P
// vector.cu:7-12: child `fidx` of a vector texture (children see index 0),
// black when it is out of range
inline f3 tex_eval(const Scene& s, int tex, Real u, Real v, f3 P = mk(0.f), int fidx = 0) {
  if (s.texs[tex].type == PT_TEX_VECTOR) {
    const Texture& tv = s.texs[tex];
    if (fidx < 0 || (uint32_t)fidx >= tv.w) return mk(0.f);
    tex = (int)tv.texels[fidx];
  }
  const Texture& t = s.texs[tex];
  if (t.type == PT_TEX_CHECKER) {
    // checkered_texture.cu:8-15: odd texture in w, even in h
    Real sines = psin(10 * P.x) * psin(10 * P.y) * psin(10 * P.z);
    return tex_leaf(s.texs[sines < 0 ? t.w : t.h], u, v, P);
  }
  return tex_leaf(t, u, v, P);
}

// ----------------------------------------------------------------- BRDFs
struct Surface {
  int type;
  f3 c0, c1;    // albedo / diffuse ; specular (AS)
  Real p0, p1; // fuzz | ior | nu,nv | A,B
};

inline Surface make_surface(const Scene& s, const HitRec& r) {
  const Material& m = s.mats[r.mat];
  Surface f;
  f.type = m.type;
  f.c0 = m.tex[0] >= 0 ? tex_eval(s, m.tex[0], r.u, r.v, r.P, r.fidx) : mk(0.f);  // the normal shader has none
  f.c1 = (m.type == PT_MAT_ASHIKHMIN_SHIRLEY) ? tex_eval(s, m.tex[1], r.u, r.v, r.P, r.fidx) : mk(0.f);
  f.p0 = m.p[0];
  f.p1 = m.p[1];
  return f;
}

// material.cuh:30-39
inline Real schlick(Real cosine, Real ref_idx) {
  Real r0 = (1.f - ref_idx) / (1.f + ref_idx);
  r0 = r0 * r0;
  return r0 + (1.f - r0) * ppow5((1.f - cosine));
}
inline f3 schlick3(f3 r0, Real cosine) {
  Real e = ppow5(1.f - cosine);
  return r0 + (mk(1.f) - r0) * e;
}
// material.cuh:50-63
inline bool Refract(f3 v, f3 n, Real ni_over_nt, f3& refracted) {
  f3 uv = normalize(v);
  Real dt = dot(uv, n);
  Real disc = 1.f - ni_over_nt * ni_over_nt * (1.f - dt * dt);
  if (disc > 0.f) {
    refracted = ni_over_nt * (uv - n * dt) - n * std::sqrt(disc);
    return true;
  }
  return false;
}
// microfacets.cuh:11-20
inline void Sample_Quadrant(Real nu, Real nv, Real rx, Real ry, Real& phi, Real& theta) {
  phi = patan(std::sqrt((nu + 1.f) / (nv + 1.f)) * ptan(2.f * PI_F * rx));
  Real cp = pcos(phi), sp = psin(phi);
  theta = ppow(ry, 1.f / (nu * cp * cp + nv * sp * sp + 1.f));
}
// microfacets.cuh:85-99
inline Real GGX_D(f3 H, Real nu, Real nv) {
  Real c2 = Cos2Theta(H);
  if (c2 <= 0.0f) return 0.f;
  if (nu == nv) {
    Real beta = c2 * ((nu * nu) - 1) + 1;
    return (nu * nu) / (PI_F * beta * beta);
  }
  Real s2nv = Sin2Phi(H) / (nv * nv);
  Real c2nu = (1.f - Sin2Phi(H)) / (nu * nu);
  Real beta = (c2 + (1.f - c2) * (c2nu + s2nv));
  return 1.f / (PI_F * nu * nv * beta * beta);
}
// microfacets.cuh:102-168
inline f3 GGX_Sample(f3 origin, Real rx, Real ry, Real nu, Real nv) {
  bool flip = origin.y < 0;
  f3 so = normalize(mk(origin.x * nu, origin.y, origin.z * nv));
  Real slopeX, slopeY;
  Real cosTheta = CosTheta(so);
  if (cosTheta > 0.9999f) {
    Real r = std::sqrt(rx / (1 - rx));
    Real phi = (Real)6.28318530718f * ry;
    slopeX = r * pcos(phi);
    slopeY = r * psin(phi);
  } else {
    Real sinTheta = std::sqrt(std::fmax(0.f, 1.f - cosTheta * cosTheta));
    Real tanTheta = sinTheta / cosTheta;
    Real a = 1.f / tanTheta;
    Real G1 = 2.f / (1.f + std::sqrt(1.f + 1.f / (a * a)));
    Real A = 2.f * rx / G1 - 1.f;
    Real tmp = 1.f / (A * A - 1.f);
    if (tmp > 1e10f) tmp = 1e10f;
    Real B = tanTheta;
    Real D = std::sqrt(std::fmax(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.f));
    Real sx1 = B * tmp - D;
    Real sx2 = B * tmp + D;
    slopeX = (A < 0 || sx2 > 1.f / tanTheta) ? sx1 : sx2;
    Real S;
    if (ry > 0.5f) { S = 1.f; ry = 2.f * (ry - .5f); }
    else { S = -1.f; ry = 2.f * (.5f - ry); }
    Real z = (ry * (ry * (ry * 0.27385f - 0.73369f) + 0.46341f)) /
              (ry * (ry * (ry * 0.093073f + 0.309420f) - 1.f) + 0.597999f);
    slopeY = S * z * std::sqrt(1.f + slopeX * slopeX);
  }
  Real t = CosPhi(so) * slopeX - SinPhi(so) * slopeY;
  slopeY = SinPhi(so) * slopeX + CosPhi(so) * slopeY;
  slopeX = t;
  slopeX = nu * slopeX;
  slopeY = nv * slopeY;
  f3 H = normalize(mk(-slopeX, 1.f, -slopeY));
  if (flip) H *= -1.f;
  return H;
}
// microfacets.cuh:170-195
inline Real GGX_Lambda(f3 V, Real nu, Real nv) {
  Real at = std::fabs(TanTheta(V));
  if (std::isinf(at)) return 0.f;
  Real alpha = std::sqrt(Cos2Phi(V) * nu * nu + Sin2Phi(V) * nv * nv);
  Real a2t2 = (alpha * at) * (alpha * at);
  return (-1.f + std::sqrt(1.f + a2t2)) / 2.f;
}
inline Real GGX_G(f3 Wo, f3 Wi, Real nu, Real nv) {
  return 1.f / (1.f + GGX_Lambda(Wo, nu, nv) + GGX_Lambda(Wi, nu, nv));
}

// metal.cu:48-49: mirror direction of the incoming ray plus fuzz times a
// point uniform in the unit ball (sampling.cuh:31-44)
inline f3 metal_scatter(f3 Wo, f3 N, Real fuzz, uint32_t& seed) {
  f3 refl = reflect(-Wo, N);
  return refl + fuzz * random_in_unit_sphere(seed);
}

// Sample(surface, P, Wo, N, seed) of the materials with NEE or a pdf:
// lambertian.cuh:7-19, oren_nayar.cuh:8-20, diffuse_light.cuh:7-13,
// ashikhmin_shirley.cuh:11-82, torrance_sparrow.cuh:10-35
inline f3 brdf_sample(const Surface& f, f3 Wo, f3 N, uint32_t& seed) {
  switch (f.type) {
    case PT_MAT_LAMBERTIAN:
    case PT_MAT_OREN_NAYAR: {
      Real u1 = rnd(seed);
      Real u2 = rnd(seed);
      return Onb(N).to_world(cosine_sample_hemisphere(u1, u2));
    }
    case PT_MAT_DIFFUSE_LIGHT:
    case PT_MAT_ISOTROPIC:  // isotropic.cuh:7-13
      return random_on_unit_sphere(seed);
    case PT_MAT_ASHIKHMIN_SHIRLEY: {
      Real nu = f.p0, nv = f.p1;
      f3 Nn = normalize(N);
      f3 T = normalize(cross(Nn, mk(0.f, 1.f, 0.f)));
      f3 B = cross(T, Nn);
      Real rx = rnd(seed);
      Real ry = rnd(seed);
      if (rx < 0.5f) {
        rx = std::fmin(2 * rx, 1.f - 1e-6f);
        return Onb(Nn).to_world(cosine_sample_hemisphere(rx, ry));
      }
      rx = std::fmin(2 * (rx - 0.5f), 1.f - 1e-6f);
      Real phi, cos_theta;
      if (nu == nv) {
        phi = 2.f * PI_F * rx;
        cos_theta = ppow(ry, 1.f / (nu + 1.f));
      } else {
        if (rx < 0.25f) { Sample_Quadrant(nu, nv, 4.f * rx, ry, phi, cos_theta); }
        else if (rx < 0.5f) { Sample_Quadrant(nu, nv, 4.f * (0.5f - rx), ry, phi, cos_theta); phi = PI_F - phi; }
        else if (rx < 0.75f) { Sample_Quadrant(nu, nv, 4.f * (rx - 0.5f), ry, phi, cos_theta); phi = PI_F + phi; }
        else { Sample_Quadrant(nu, nv, 4.f * (1.f - rx), ry, phi, cos_theta); phi = 2.f * PI_F - phi; }
      }
      Real sin_theta = std::sqrt(std::fmax(0.f, 1.f - cos_theta * cos_theta));
      f3 H = normalize(Spherical_Vector(sin_theta, cos_theta, phi));
      H = H.x * B + H.y * Nn + H.z * T;
      if (dot(H, Wo) < 0.f) H = -H;
      return normalize(-Wo + 2.f * dot(Wo, H) * H);
    }
    default: {  // Torrance-Sparrow
      Real nu = f.p0, nv = f.p1;
      f3 Nn = normalize(N);
      f3 T = normalize(cross(Nn, mk(0.f, 1.f, 0.f)));
      f3 B = cross(T, Nn);
      Real rx = rnd(seed);
      Real ry = rnd(seed);
      f3 H = normalize(GGX_Sample(Wo, rx, ry, nu, nv));
      H = H.x * B + H.y * Nn + H.z * T;
      if (dot(H, Wo) < 0.f) H = -H;
      return normalize(-Wo + 2.f * dot(Wo, H) * H);
    }
  }
}

// Evaluate(surface, P, Wo, Wi, N, pdf): lambertian.cuh:21-42,
// oren_nayar.cuh:22-65, diffuse_light.cuh:15-22, ashikhmin_shirley.cuh:84-151,
// torrance_sparrow.cuh:37-64.  pdf keeps its incoming value on early returns.
inline f3 brdf_eval(const Surface& f, f3 Wo, f3 Wi, f3 N, Real& pdf) {
  switch (f.type) {
    case PT_MAT_LAMBERTIAN: {
      Real cosine = dot(normalize(Wi), normalize(N));
      if (cosine < 0.f) return mk(0.f);
      pdf = dot(normalize(Wi), normalize(N)) * (1.0f / PI_F);  // lambertian.cuh:33-38, 1/pi folded
      return (cosine * f.c0) / PI_F;
    }
    case PT_MAT_OREN_NAYAR: {
      f3 WiN = normalize(Wi);
      Real sI = SinTheta(WiN), sO = SinTheta(Wo);
      Real maxCos = 0;
      if (sI > 1e-4f && sO > 1e-4f) {
        Real sPI = SinPhi(WiN), cPI = CosPhi(WiN);
        Real sPO = SinPhi(Wo), cPO = CosPhi(Wo);
        Real dCos = cPI * cPO + sPI * sPO;
        maxCos = std::fmax(0.f, dCos);
      }
      Real sinAlpha, tanBeta;
      if (AbsCosTheta(WiN) > AbsCosTheta(Wo)) { sinAlpha = sO; tanBeta = sI / AbsCosTheta(WiN); }
      else { sinAlpha = sI; tanBeta = sO / AbsCosTheta(Wo); }
      pdf = dot(normalize(Wi), normalize(N)) / PI_F;
      return f.c0 * (1.f / PI_F) * (f.p0 + f.p1 * maxCos * sinAlpha * tanBeta);
    }
    case PT_MAT_DIFFUSE_LIGHT:
      pdf = 1.f;
      return f.c0;
    case PT_MAT_ISOTROPIC: {
      // isotropic.cuh:15-32 returns pi/4 for both; phase function and pdf
      // are 1/(4 pi)
      const Real k = 0.25f / PI_F;
      pdf = k;
      return f.c0 * k;
    }
    case PT_MAT_ASHIKHMIN_SHIRLEY: {
      f3 Rd = f.c0, Rs = f.c1;
      Real nu = f.p0, nv = f.p1;
      f3 Up = mk(0.f, 1.f, 0.f);
      f3 Nn = normalize(N);
      f3 T = normalize(cross(Nn, Up));
      f3 B = cross(T, Nn);
      Real NdotI = std::fabs(dot(Up, Wi)), NdotO = std::fabs(dot(Up, Wo));
      f3 diff = Rd * (28.f / (23.f * PI_F));
      diff *= (mk(1.f) - Rs);
      diff *= (1.f - ppow5(1.f - NdotI * 0.5f));
      diff *= (1.f - ppow5(1.f - NdotO * 0.5f));
      Real dpdf = dot(Wi, Nn) / PI_F;
      if (dpdf < 0.f) return mk(0.f);
      f3 H = Wo + Wi;
      if (isNull(H)) return mk(0.f);
      H = normalize(H);
      Real HdotI = std::fabs(dot(H, Wi));
      Real HdotN = std::fabs(dot(H, Nn)), HdotT = std::fabs(dot(H, T)), HdotB = std::fabs(dot(H, B));
      Real norm, lobe;
      if (nu == nv) {
        norm = (nu + 1.f) / (8.f * PI_F);
        lobe = ppow(HdotN, nu);
      } else {
        norm = std::sqrt((nu + 1.f) * (nv + 1.f)) / (8.f * PI_F);
        if (HdotN < 1.f)
          lobe = ppow(HdotN, (nu * HdotT * HdotT + nv * HdotB * HdotB) / (1.f - HdotN * HdotN));
        else
          lobe = 1.f;
      }
      f3 spec = schlick3(Rs, HdotI);
      spec *= norm;
      spec *= lobe / (HdotI * std::fmax(NdotI, NdotO));
      Real spdf = HdotI / (norm * lobe);
      pdf = (dpdf + spdf) / 2.f;
      return (diff / dpdf) + (spec / spdf);
    }
    default: {  // Torrance-Sparrow
      f3 Rs = f.c0;
      Real nu = f.p0, nv = f.p1;
      f3 Up = mk(0.f, 1.f, 0.f);
      Real NdotI = std::fmax(dot(Up, Wi), 1e-6f), NdotO = std::fmax(dot(Up, Wo), 1e-6f);
      f3 H = Wo + Wi;
      if (isNull(H)) return mk(0.f);
      H = normalize(H);
      Real HdotI = std::fabs(dot(H, Wi));
      f3 F = schlick3(Rs, HdotI);
      Real G = GGX_G(Wo, Wi, nu, nv);
      Real D = GGX_D(H, nu, nv);
      pdf = (D * AbsCosTheta(H)) / (4.f * dot(Wo, H));
      return Rs * D * G * F / (4.f * NdotI * NdotO);
    }
  }
}

// ---------------------------------------------------------------- lights
// rect_pdf.cu:11-146, sphere_pdf.cu:8-74
inline f3 light_sample(const Light& l, f3 P, uint32_t& seed) {
  if (l.type == PT_LIGHT_RECT) {
    Real ra = rnd(seed);
    Real ca = l.a0 + ra * (l.a1 - l.a0);
    Real rb = rnd(seed);
    Real cb = l.b0 + rb * (l.b1 - l.b0);
    f3 pt = l.axis == PT_AXIS_X ? mk(l.k, ca, cb) : (l.axis == PT_AXIS_Y ? mk(ca, l.k, cb) : mk(ca, cb, l.k));
    return pt - P;
  }
  // sphere: cone sample, then the reference's basis built around Wi itself
  Real r1 = rnd(seed);
  Real r2 = rnd(seed);
  Real dsq = squared_length(l.c - P);
  Real z = 1.f + r2 * (std::sqrt(1.f - l.r * l.r / dsq) - 1.f);
  Real phi = 2.f * PI_F * r1;
  Real x = pcos(phi) * std::sqrt(1.f - z * z);
  Real y = psin(phi) * std::sqrt(1.f - z * z);
  f3 Wi = mk(x, y, z);
  return Onb(normalize(Wi)).to_world(Wi);
}
inline Real light_pdf(const Light& l, f3 P, f3 Wi) {
  if (l.type == PT_LIGHT_RECT) {
    Real t, a, b;
    f3 rn;
    if (l.axis == PT_AXIS_X) { t = (l.k - P.x) / Wi.x; a = P.y + t * Wi.y; b = P.z + t * Wi.z; rn = mk(1.f, 0.f, 0.f); }
    else if (l.axis == PT_AXIS_Y) { t = (l.k - P.y) / Wi.y; a = P.x + t * Wi.x; b = P.z + t * Wi.z; rn = mk(0.f, 1.f, 0.f); }
    else { t = (l.k - P.z) / Wi.z; a = P.x + t * Wi.x; b = P.y + t * Wi.y; rn = mk(0.f, 0.f, 1.f); }
    if (a < l.a0 || a > l.a1 || b < l.b0 || b > l.b1) return 0.f;
    if (!(t < FLT_MAX && t > 0.001f)) return 0.f;
    Real dsq = t * t * squared_length(Wi);
    Real cosine = std::fabs(dot(Wi, rn) / length(Wi));
    Real area = (l.a1 - l.a0) * (l.b1 - l.b0);
    return dsq / (cosine * area);
  }
  // sphere
  f3 oc = P - l.c;
  Real a = dot(Wi, Wi), b = dot(oc, Wi), c = dot(oc, oc) - l.r * l.r;
  Real disc = b * b - a * c;
  if (disc < 0.f) return 0.f;
  Real tt = (-b - std::sqrt(disc)) / a;
  bool hit = (tt < FLT_MAX && tt > 0.001f);
  if (!hit) {
    tt = (-b + std::sqrt(disc)) / a;
    hit = (tt < FLT_MAX && tt > 0.001f);
  }
  if (!hit) return 0.f;
  Real dsq = squared_length(l.c - P);
  Real ctm = std::sqrt(1.f - l.r * l.r / dsq);
  Real sa = 2.f * PI_F * (1.f - ctm);
  return 1.f / sa;
}

// ------------------------------------------------------- environment
// Spherical HDR mapping of miss.cu:64-95: u = atan2(z, x) / 2pi,
// v = acos(y / |d|) / pi (u wraps through the texture's repeat mode).
inline void env_uv(f3 d, Real& u, Real& v) {
  Real r = length(d);
  Real lon = patan2(d.z, d.x);
  Real lat = pacos(d.y / r);
  u = lon * (1.f / (PI_F * 2.f));
  v = lat * (1.f / PI_F);
}
// radiance the miss program returns for direction d (x2, miss.cu:93)
inline f3 env_radiance(const Scene& s, f3 d) {
  Real u, v;
  env_uv(d, u, v);
  return 2.f * tex_eval(s, s.miss.texture, u, v);
}
// floor(total * r / 2^24) for a 24-bit r, exactly
inline uint64_t scale24(uint64_t total, uint32_t r) {
  return (total >> 24) * r + (((total & 0xFFFFFFull) * r) >> 24);
}
// first index in [0, n) whose inclusive prefix exceeds target
inline uint32_t upper_index(const uint64_t* a, uint32_t n, uint64_t target) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > target) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}
// density of env_sample per unit solid angle: p(u, v) W H / (2 pi^2 sin(theta))
inline Real env_pdf(const Scene& s, f3 d) {
  const EnvCdf& e = s.env;
  if (e.total == 0) return 0.f;
  Real u, v;
  env_uv(d, u, v);
  Real uw = u - std::floor(u);
  if (!(uw == uw && v == v)) return 0.f;  // a NaN direction
  int64_t i = (int64_t)(uw * (Real)e.W), j = (int64_t)(v * (Real)e.H);
  if (i >= (int64_t)e.W) i = e.W - 1;
  if (j >= (int64_t)e.H) j = e.H - 1;
  if (i < 0) i = 0;
  if (j < 0) j = 0;
  const uint64_t* row = &e.cond[(size_t)j * e.W];
  uint64_t w = row[i] - (i ? row[i - 1] : 0);
  Real st = psin(v * PI_F);
  if (w == 0 || !(st > 0.f)) return 0.f;
  const Real p = ((Real)w / (Real)e.total) * ((Real)e.W * (Real)e.H);
  return p / ((2.0f * PI_F * PI_F) * st);
}
// two draws: row from the marginal table, column from the row's table,
// continuous offsets inside the texel
inline f3 env_sample(const Scene& s, uint32_t& seed) {
  const EnvCdf& e = s.env;
  uint32_t r1 = lcg(seed);
  uint32_t r2 = lcg(seed);
  if (e.total == 0) return mk(0.f, 1.f, 0.f);
  uint64_t tr = scale24(e.total, r1);
  uint32_t j = upper_index(e.marg.data(), e.H, tr);
  uint64_t rprev = j ? e.marg[j - 1] : 0;
  uint64_t rw = e.marg[j] - rprev;
  const uint64_t* row = &e.cond[(size_t)j * e.W];
  uint64_t tc = scale24(rw, r2);
  uint32_t i = upper_index(row, e.W, tc);
  uint64_t cprev = i ? row[i - 1] : 0;
  // offsets inside the texel, below 1 (a ratio of two roundings may reach it)
  const Real fu = std::fmin((Real)(tc - cprev) / (Real)(row[i] - cprev), (Real)0x1.fffffep-1f);
  const Real fv = std::fmin((Real)(tr - rprev) / (Real)rw, (Real)0x1.fffffep-1f);
  Real u = ((Real)i + fu) / (Real)e.W;
  Real v = ((Real)j + fv) / (Real)e.H;
  Real lon = 2.f * PI_F * u, lat = PI_F * v;
  Real sl = psin(lat);
  return mk(sl * pcos(lon), pcos(lat), sl * psin(lon));
}

// light_sample.cuh:27-33
inline Real PowerHeuristic(unsigned int nf, Real fPdf, unsigned int ng, Real gPdf) {
  Real f = nf * fPdf, g = ng * gPdf;
  return (f * f) / (f * f + g * g);
}

inline uint32_t mix(uint32_t h, uint32_t v) { return (h ^ v) * 16777619u; }

struct PathCounters {
  uint64_t ext = 0, shadow = 0;
};

// light_sample.cuh:35-103 with the shadow ray resolved by the shared rule.
// `lights` is the scene's table, without environment lights unless
// environment MIS is on.  An environment light is lit iff the shadow ray
// hits nothing; its BRDF-sampled half is left to the continuation ray, whose
// miss is MIS-weighted against env_pdf.
inline f3 direct_light(const Scene& s, const std::vector<Light>& lights, const Surface& f, f3 P, f3 Wo, f3 N,
                       bool isLight, uint32_t& seed, uint32_t& sig, PathCounters& pc, Real time) {
  f3 direct = mk(0.f);
  int nl = (int)lights.size();
  if (nl == 0) return mk(0.f);
  int index = ((int)(rnd(seed) * nl)) % nl;
  if (isLight && nl == 1) return mk(0.f);
  const Light& L = lights[index];
  const bool env = L.type == PT_LIGHT_ENV;
  f3 Wi = env ? env_sample(s, seed) : light_sample(L, P, seed);
  Real lightPDF = env ? env_pdf(s, Wi) : light_pdf(L, P, Wi);
  f3 emission = env ? env_radiance(s, Wi) : L.e;
  // the isotropic medium's +X normal is not a surface's: no hemisphere tests
  // (light_sample.cuh:60 and the shadow any-hit's :71)
  const bool iso = f.type == PT_MAT_ISOTROPIC;
  if (!iso && dot(Wi, N) < 0.f) { sig = mix(sig, 0x33333333u); return mk(0.f); }
  pc.shadow++;
  HitState h = s.closest(P, Wi, 1e-3f, INF, nullptr, seed, time);
  bool occluded = false;
  if (h.found()) {
    int m = s.material_of(h.inst, h.prim);
    occluded = env || !(s.mats[m].type == PT_MAT_DIFFUSE_LIGHT && (iso || dot(N, Wi) > 0.f));
  }
  sig = mix(sig, occluded ? 0xA5A5A5A5u : 0x5A5A5A5Au);
  if (occluded) return mk(0.f);
  if (lightPDF != 0.f && !isNull(emission)) {
    Real matPDF = 1.f;
    f3 mv = brdf_eval(f, Wo, Wi, N, matPDF);
    if (matPDF != 0.f && !isNull(mv)) {
      Real w = PowerHeuristic(1, lightPDF, 1, matPDF);
      direct += mv * emission * w / lightPDF;
    }
  }
  if (env) return direct;
  Wi = brdf_sample(f, Wo, N, seed);
  Real matPDF = 1.f;
  f3 mv = brdf_eval(f, Wo, Wi, N, matPDF);
  if (matPDF != 0.f && !isNull(mv)) {
    lightPDF = light_pdf(L, P, Wi);
    if (!lightPDF || isNull(emission)) return direct;
    Real w = PowerHeuristic(1, matPDF, 1, lightPDF);
    direct += mv * emission * w / matPDF;
  }
  return direct;
}

// miss.cu:30-95
inline f3 miss_color(const Scene& s, f3 d) {
  const pt_miss_desc& m = s.miss;
  switch (m.type) {
    case PT_MISS_GRADIENT: {
      f3 ud = normalize(d);
      Real t = 0.5f * (ud.y + 1.f);
      f3 c = (1.f - t) * mk(m.color1[0], m.color1[1], m.color1[2]);
      c += t * mk(m.color2[0], m.color2[1], m.color2[2]);
      return c;
    }
    case PT_MISS_CONSTANT:
      return mk(m.color1[0], m.color1[1], m.color1[2]);
    case PT_MISS_IMAGE: {
      const Real PI = 3.14159265358979323846f, INV_PI = 0.318309886183790671538f;
      Real theta = patan2(d.x, d.z);
      Real phi = PI * 0.5f - pacos(d.y);
      Real u = (theta + PI) * (0.5f * INV_PI);
      Real v = 0.5f * (1.f + psin(phi));
      return tex_eval(s, m.texture, u, v);
    }
    default: {
      Real u, v;
      if (m.spherical) {
        env_uv(d, u, v);
      } else {
        Real theta = patan2(d.x, d.z);
        theta = theta < 0.f ? theta + (2.f * PI_F) : theta;
        Real phi = pacos(d.y);
        u = 1.f - (theta / (2.f * PI_F));
        v = phi / PI_F;
      }
      return 2.f * tex_eval(s, m.texture, u, v);
    }
  }
}

struct Params {
  uint32_t W, H, max_depth;
  bool russian;
  uint32_t rr_min_depth;
  bool env_mis;  // PT_FLAG_ENV_MIS: environment lights sampled, unclamped MIS-weighted misses
};

struct PathResult {
  f3 L;
  uint32_t sig;
  uint32_t bounces;
};

// raygen.cu:148-168 (seed, jitter, camera ray) and raygen.cu:66-123 (color)
inline PathResult trace_path(const Scene& s, const pt_camera& cam, const Params& p, uint32_t x, uint32_t y,
                             uint32_t frame, PathCounters& pc) {
  PathResult res{mk(0.f), 0x811C9DC5u, 0};
  uint32_t seed = tea(p.W * y + x, frame, 64);
  Real u = ((Real)x + rnd(seed)) / (Real)p.W;
  Real v = ((Real)y + rnd(seed)) / (Real)p.H;
  f3 cu = mk(cam.u[0], cam.u[1], cam.u[2]), cv = mk(cam.v[0], cam.v[1], cam.v[2]);
  f3 rd = cam.lens_radius * random_in_unit_disk(seed);
  f3 lens = cu * rd.x + cv * rd.y;
  f3 origin = mk(cam.origin[0], cam.origin[1], cam.origin[2]) + lens;
  f3 dir = mk(cam.lower_left[0], cam.lower_left[1], cam.lower_left[2]) +
           u * mk(cam.horizontal[0], cam.horizontal[1], cam.horizontal[2]) +
           v * mk(cam.vertical[0], cam.vertical[1], cam.vertical[2]) - origin;
  Real tmin = 1e-6f;
  const Real time = cam.time0 + rnd(seed) * (cam.time1 - cam.time0);  // prd.time (raygen.cu:69)
  const std::vector<Light>& lights = p.env_mis ? s.lights : s.lights_plain;
  const bool env_light = p.env_mis && s.env.total > 0 && lights.size() != s.lights_plain.size();
  f3 T = mk(1.f), L = mk(0.f);
  bool prevSpec = false;
  Real prevPdf = 1.f;  // pdf of the last continuation sample (Lambertian / Oren-Nayar)
  for (uint32_t depth = 0; depth < p.max_depth; depth++) {
    pc.ext++;
    res.bounces = depth + 1;
    HitState h = s.closest(origin, dir, tmin, INF, nullptr, seed, time);
    res.sig = mix(res.sig, h.found() ? (((uint32_t)h.inst << 20) ^ (uint32_t)h.prim) : 0xFFFFFFFFu);
    if (!h.found()) {
      if (p.env_mis && s.miss.type == PT_MISS_HDR) {
        // unclamped environment radiance, MIS-weighted against the light
        // sample of the previous (diffuse) bounce
        f3 Le = env_radiance(s, dir);
        Real w = 1.f;
        if (env_light && depth > 0 && !prevSpec) w = PowerHeuristic(1, prevPdf, 1, env_pdf(s, dir));
        T *= Le * w;
        res.L = L + T;
        return res;
      }
      T *= miss_color(s, dir);
      res.L = L + clamp3(T, 0.f, 1.f);
      return res;
    }
    HitRec r = hit_record(s, h.inst, h.prim, h.t, origin, dir, time);
    Surface f = make_surface(s, r);
    f3 P = r.P, Wo = r.Wo, N = r.N;
    bool spec = false;
    switch (f.type) {
      case PT_MAT_DIFFUSE_LIGHT: {  // diffuse_light.cu:30-48
        f3 dl = direct_light(s, lights, f, P, Wo, N, true, seed, res.sig, pc, time);
        L += T * dl;
        if (dot(N, Wo) < 0.f) T *= f.c0;
        if (depth == 0 || prevSpec) L += T;
        res.L = L;
        return res;
      }
      case PT_MAT_LAMBERTIAN:
      case PT_MAT_OREN_NAYAR: {  // lambertian.cu:49-73, oren_nayar.cu:46-69
        f3 dl = direct_light(s, lights, f, P, Wo, N, false, seed, res.sig, pc, time);
        L += T * dl;
        f3 Wi = brdf_sample(f, Wo, N, seed);
        Real pdf = 1.f;
        f3 att = brdf_eval(f, Wo, Wi, N, pdf);
        origin = P;
        dir = Wi;
        T *= clamp3(att / pdf, 0.f, 1.f);
        prevPdf = pdf;
        break;
      }
      case PT_MAT_ISOTROPIC: {  // isotropic.cu:27-51
        f3 dl = direct_light(s, lights, f, P, Wo, N, false, seed, res.sig, pc, time);
        L += T * dl;
        f3 Wi = brdf_sample(f, Wo, N, seed);
        Real pdf = 1.f;
        f3 att = brdf_eval(f, Wo, Wi, N, pdf);
        origin = P;
        dir = Wi;
        T *= att / pdf;
        prevPdf = pdf;
        break;
      }
      case PT_MAT_NORMAL: {  // normal_shader.cu:22-34: radiance = 0.5 N + 0.5, cancelled
        f3 n = f.p0 != 0.f ? N : geometric_normal(s, h.inst, h.prim, N);
        res.L = n * 0.5f + mk(0.5f);
        return res;
      }
      case PT_MAT_METAL: {  // metal.cu:38-56
        dir = metal_scatter(Wo, N, f.p0, seed);
        origin = P;
        T *= f.c0;
        spec = true;
        break;
      }
      case PT_MAT_DIELECTRIC: {  // dielectric.cu:46-96
        Real ior = f.p0;
        f3 W = -Wo;
        f3 Nd = N;
        Real ni_over_nt;
        Real cosine = dot(W, Nd);
        if (cosine > 0.f) {
          Nd = -Nd;
          ni_over_nt = ior;
          cosine = ior * cosine / length(W);
        } else {
          ni_over_nt = 1.f / ior;
          cosine = -cosine / length(W);
        }
        f3 refracted = mk(0.f);
        Real reflect_prob;
        if (Refract(W, Nd, ni_over_nt, refracted)) reflect_prob = schlick(cosine, ior);
        else reflect_prob = 1.f;
        bool refl = rnd(seed) < reflect_prob;
        res.sig = mix(res.sig, refl ? 0x11111111u : 0x22222222u);
        if (refl) dir = reflect(W, Nd);
        else dir = normalize(refracted);
        origin = P;
        T *= f.c0;  // absorption disabled (dielectric.cu:67)
        spec = true;
        break;
      }
      case PT_MAT_ASHIKHMIN_SHIRLEY: {  // ashikhmin_shirley.cu:48-68
        f3 Wi = brdf_sample(f, Wo, N, seed);
        Real pdf = 1.f;
        f3 att = brdf_eval(f, Wo, Wi, N, pdf);
        origin = P;
        dir = Wi;
        T *= clamp3(att, 0.f, 1.f);
        spec = true;
        break;
      }
      default: {  // Torrance-Sparrow, torrance_sparrow.cu:39-59
        f3 Wi = brdf_sample(f, Wo, N, seed);
        Real pdf = 1.f;
        f3 att = brdf_eval(f, Wo, Wi, N, pdf);
        origin = P;
        dir = Wi;
        T *= clamp3(att / pdf, 0.f, 1.f);
        spec = true;
        break;
      }
    }
    tmin = 1e-3f;
    prevSpec = spec;
    if (p.russian) {  // raygen.cu:110-118
      Real prob = max_component(T);
      if (depth >= p.rr_min_depth) {
        bool term = rnd(seed) >= prob;
        res.sig = mix(res.sig, term ? 0x44444444u : 0x55555555u);
        if (term) {
          res.L = L + T;
          return res;
        }
        T *= 1.f / prob;
      }
    }
  }
  res.L = mk(0.f);  // raygen.cu:121-122
  return res;
}

// raygen.cu:126-134
inline f3 de_nan(f3 c) {
  if (!(c.x == c.x)) c.x = 0.f;
  if (!(c.y == c.y)) c.y = 0.f;
  if (!(c.z == c.z)) c.z = 0.f;
  return c;
}

// raygen.cu:136-146 on the FP64 accumulation
inline uint8_t display_channel(double acc, uint32_t frames) {
  Real avg = (Real)(acc / (double)frames);
  Real s = std::sqrt(avg);
  s = clampf(s, 0.f, 1.f);
  return (uint8_t)(int)(255.99 * (double)s);
}

}  // namespace ptref
