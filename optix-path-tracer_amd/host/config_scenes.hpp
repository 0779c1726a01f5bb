// config_scenes.hpp -- the synthetic benchmark scenes, written against the
// host scene API exactly as the reference's scenes.hpp functions are
// (setRayGenerationProgram, setMissProgram, Group + Hitable_List, Camera).
//
//  0  random spheres: ground r=1000 at (0,-1000,0), three r=1 spheres, a
//     22x22 grid of r=0.2 spheres (a + 0.9u, 0.2, b + 0.9u) skipped within
//     0.9 of (4,0.2,0); camera (13,2,3) -> 0, vfov 20; sky gradient.
//  1  Cornell box: 555 box of rects (green x=555 on the image left, red x=0),
//     light rect y=554 emission 15, tall box 165x330x165 rotated 15 deg and
//     translated (265,0,295), short box 165^3 rotated -18 deg and translated
//     (130,0,65); camera (278,278,-800) -> (278,278,0), vfov 40.
//  2  triangle meshes: 8 octaspheres of 8*4^7 = 2^17 triangles each, displaced
//     along the normal by 0.05 fBm (4 octaves, noise seed = mesh index),
//     scaled U[0.5,2], rotated about Y U[0,360), translated U[-10,10]^3
//     (scene RNG seed 2), on a 40x40 ground quad at the lowest mesh point;
//     materials cycle Torrance-Sparrow (roughness U[0.05,0.6]),
//     Ashikhmin-Shirley (nu, nv U[10,1000]), Oren-Nayar (sigma U[0,30 deg]),
//     dielectric 1.5; a synthetic 2048x1024 HDRI (sky gradient + Gaussian
//     sun of peak 5e4 and 0.5 deg width at elevation 35 deg, azimuth 40 deg)
//     as miss program and importance-sampled environment light; camera
//     (20,4,0) -> origin, vfov 45.
#pragma once

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <memory>
#include <numeric>
#include <random>
#include <unordered_map>
#include <vector>

#include "camera.hpp"
#include "hitables.hpp"
#include "mesh.hpp"
#include "pdfs.hpp"

struct SceneArena {
  std::vector<std::unique_ptr<Texture>> textures;
  std::vector<std::unique_ptr<BRDF>> materials;
  std::vector<std::unique_ptr<Hitable>> hitables;
  std::vector<std::unique_ptr<Mesh>> meshes;
  template <class T, class... A>
  T* tex(A&&... a) {
    textures.emplace_back(new T(std::forward<A>(a)...));
    return static_cast<T*>(textures.back().get());
  }
  template <class T, class... A>
  T* mat(A&&... a) {
    materials.emplace_back(new T(std::forward<A>(a)...));
    return static_cast<T*>(materials.back().get());
  }
  template <class T, class... A>
  T* hit(A&&... a) {
    hitables.emplace_back(new T(std::forward<A>(a)...));
    return static_cast<T*>(hitables.back().get());
  }
};

struct ConfigDefaults {
  int W, H, spp, depth;
  bool russian;
  int rr_min_depth;
  uint32_t flags = PT_FLAG_MATERIAL_SORT;  // PT_FLAG_* the workload renders with
};

// random-spheres scene, scene RNG seeded with `seed`
inline ConfigDefaults RandomSpheres(App_State& app, SceneArena& A, uint32_t seed = 0) {
  rnd_reseed(seed);
  Light_Sampler lights;
  setRayGenerationProgram(app.context, lights);
  setMissProgram(app.context, GRADIENT, make_float3(1.f), make_float3(0.5f, 0.7f, 1.f));
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  Hitable_List list;
  BRDF* ground = A.mat<Lambertian>(A.tex<Constant_Texture>(0.5f));
  list.push(A.hit<Sphere>(make_float3(0.f, -1000.f, 0.f), 1000.f, ground));
  for (int a = -11; a < 11; a++) {
    for (int b = -11; b < 11; b++) {
      float choose_mat = rnd();
      float cx = (float)a + 0.9f * rnd();
      float cz = (float)b + 0.9f * rnd();
      float3 center = make_float3(cx, 0.2f, cz);
      if (length(center - make_float3(4.f, 0.2f, 0.f)) <= 0.9f) continue;
      BRDF* m;
      if (choose_mat < 0.8f) {
        float r = rnd() * rnd();
        float g = rnd() * rnd();
        float bl = rnd() * rnd();
        m = A.mat<Lambertian>(A.tex<Constant_Texture>(r, g, bl));
      } else if (choose_mat < 0.95f) {
        float r = 0.5f * (1.f + rnd());
        float g = 0.5f * (1.f + rnd());
        float bl = 0.5f * (1.f + rnd());
        float fuzz = 0.5f * rnd();
        m = A.mat<Metal>(A.tex<Constant_Texture>(r, g, bl), fuzz);
      } else {
        Texture* one = A.tex<Constant_Texture>(1.f);
        m = A.mat<Dielectric>(one, one, 1.5f, 0.f);
      }
      list.push(A.hit<Sphere>(center, 0.2f, m));
    }
  }
  Texture* one = A.tex<Constant_Texture>(1.f);
  list.push(A.hit<Sphere>(make_float3(0.f, 1.f, 0.f), 1.f, A.mat<Dielectric>(one, one, 1.5f, 0.f)));
  list.push(A.hit<Sphere>(make_float3(-4.f, 1.f, 0.f), 1.f,
                          A.mat<Lambertian>(A.tex<Constant_Texture>(0.4f, 0.2f, 0.1f))));
  list.push(A.hit<Sphere>(make_float3(4.f, 1.f, 0.f), 1.f,
                          A.mat<Metal>(A.tex<Constant_Texture>(0.7f, 0.6f, 0.5f), 0.f)));
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  Camera camera(make_float3(13.f, 2.f, 3.f), make_float3(0.f), make_float3(0.f, 1.f, 0.f), 20.f,
                float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  return ConfigDefaults{400, 267, 16, 50, true, 11};
}

inline ConfigDefaults CornellBox(App_State& app, SceneArena& A) {
  Light_Sampler lights;
  Rectangle_PDF rect_pdf(213.f, 343.f, 227.f, 332.f, 554.f, Y_AXIS);
  lights.pdf.push_back(rect_pdf.createPDF(app.context));
  lights.sample.push_back(rect_pdf.createSample(app.context));
  lights.emissions.push_back(make_float3(15.f));
  setRayGenerationProgram(app.context, lights);
  setMissProgram(app.context, CONSTANT);
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  BRDF* red = A.mat<Lambertian>(A.tex<Constant_Texture>(0.65f, 0.05f, 0.05f));
  BRDF* white = A.mat<Lambertian>(A.tex<Constant_Texture>(0.73f));
  BRDF* green = A.mat<Lambertian>(A.tex<Constant_Texture>(0.12f, 0.45f, 0.15f));
  BRDF* light = A.mat<Diffuse_Light>(A.tex<Constant_Texture>(15.f));
  Hitable_List list;
  list.push(A.hit<AARect>(0.f, 555.f, 0.f, 555.f, 555.f, true, X_AXIS, green));
  list.push(A.hit<AARect>(0.f, 555.f, 0.f, 555.f, 0.f, false, X_AXIS, red));
  list.push(A.hit<AARect>(213.f, 343.f, 227.f, 332.f, 554.f, true, Y_AXIS, light));
  list.push(A.hit<AARect>(0.f, 555.f, 0.f, 555.f, 555.f, true, Y_AXIS, white));
  list.push(A.hit<AARect>(0.f, 555.f, 0.f, 555.f, 0.f, false, Y_AXIS, white));
  list.push(A.hit<AARect>(0.f, 555.f, 0.f, 555.f, 555.f, true, Z_AXIS, white));
  Box* tall = A.hit<Box>(make_float3(0.f), make_float3(165.f, 330.f, 165.f), white);
  tall->translate(make_float3(265.f, 0.f, 295.f));
  tall->rotate(15.f, Y_AXIS);
  list.push(tall);
  Box* small = A.hit<Box>(make_float3(0.f), make_float3(165.f, 165.f, 165.f), white);
  small->translate(make_float3(130.f, 0.f, 65.f));
  small->rotate(-18.f, Y_AXIS);
  list.push(small);
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  Camera camera(make_float3(278.f, 278.f, -800.f), make_float3(278.f, 278.f, 0.f), make_float3(0.f, 1.f, 0.f), 40.f,
                float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  return ConfigDefaults{512, 512, 64, 16, true, 3};
}

// scenes.hpp:95-186: the random-spheres layout on a checkered ground, lit by
// a rect light under a dark sky, with a Perlin-noise metal sphere.  The
// scene RNG is seeded with `seed`; colours are drawn r, g, b in that order.
// The reference's Earth sphere samples assets/other_textures/map.jpg, which
// the reference tree does not ship: a Gradient_Texture stands in for it.
inline ConfigDefaults MovingSpheres(App_State& app, SceneArena& A, uint32_t seed = 0) {
  rnd_reseed(seed);
  Light_Sampler lights;
  Rectangle_PDF rect_pdf(3.f, 5.f, 1.f, 3.f, -0.5f, Z_AXIS);
  lights.pdf.push_back(rect_pdf.createPDF(app.context));
  lights.sample.push_back(rect_pdf.createSample(app.context));
  lights.emissions.push_back(make_float3(4.f));
  setRayGenerationProgram(app.context, lights);
  setMissProgram(app.context, CONSTANT);
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  Hitable_List list;
  Texture* groundTx = A.tex<Checker_Texture>(A.tex<Constant_Texture>(0.2f, 0.3f, 0.1f),
                                             A.tex<Constant_Texture>(0.9f, 0.9f, 0.9f));
  list.push(A.hit<Sphere>(make_float3(0.f, -1000.f, -1.f), 1000.f, A.mat<Lambertian>(groundTx)));
  auto tint = [&]() {
    const float r = 0.5f * (1.f + rnd());
    const float g = 0.5f * (1.f + rnd());
    const float b = 0.5f * (1.f + rnd());
    return A.tex<Constant_Texture>(r, g, b);
  };
  for (int a = -11; a < 11; a++) {
    for (int b = -11; b < 11; b++) {
      const float choose_mat = rnd();
      const float cx = (float)a + rnd();
      const float cz = (float)b + rnd();
      const float3 center = make_float3(cx, 0.2f, cz);
      rnd();  // center2's offset: the reference draws it and renders static spheres
      BRDF* m;
      if (choose_mat < (1.f / 3)) {
        m = A.mat<Lambertian>(tint());
      } else if (choose_mat < (2.f / 3)) {
        Texture* t = tint();
        m = A.mat<Metal>(t, 0.5f * rnd());
      } else {
        Texture* t = tint();
        m = A.mat<Dielectric>(t, t, 1.5f, 0.f);
      }
      list.push(A.hit<Sphere>(center, 0.2f, m));
    }
  }
  Texture* earth = A.tex<Gradient_Texture>(make_float3(0.1f, 0.3f, 0.8f), make_float3(0.2f, 0.7f, 0.2f),
                                           make_float3(0.9f, 0.9f, 0.95f));
  list.push(A.hit<Sphere>(make_float3(-4.f, 1.f, 2.f), 1.f, A.mat<Lambertian>(earth)));
  Texture* gtx1 = A.tex<Constant_Texture>(1.f);
  const float gr = rnd();
  const float gg = rnd();
  const float gb = rnd();
  list.push(A.hit<Sphere>(make_float3(4.f, 1.f, 1.f), 1.f,
                          A.mat<Dielectric>(gtx1, A.tex<Constant_Texture>(gr, gg, gb), 1.5f, 0.f)));
  list.push(A.hit<Sphere>(make_float3(0.f, 1.f, 1.5f), 1.f, A.mat<Metal>(A.tex<Noise_Texture>(4.f), 0.f)));
  list.push(A.hit<AARect>(3.f, 5.f, 1.f, 3.f, -0.5f, false, Z_AXIS, A.mat<Diffuse_Light>(A.tex<Constant_Texture>(4.f))));
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  Camera camera(make_float3(13.f, 2.f, 3.f), make_float3(0.f), make_float3(0.f, 1.f, 0.f), 20.f,
                float(app.W) / float(app.H), 0.1f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  return ConfigDefaults{400, 267, 16, 50, true, 11};
}

// scenes.hpp:263-375: 400 random-height ground boxes, a rect light, glass,
// metal and brown spheres, a blue medium inside a glass sphere, a thin white
// medium around everything, a Perlin sphere and 1000 small spheres in one
// rotated, translated group (Hitable_List::addListTo).  The scene RNG is
// seeded with `seed`; the Earth's map.jpg is not in the reference tree, so
// a Gradient_Texture stands in for it.
inline ConfigDefaults NextWeek(App_State& app, SceneArena& A, uint32_t seed = 0) {
  rnd_reseed(seed);
  Light_Sampler lights;
  Rectangle_PDF rect_pdf(113.f, 443.f, 127.f, 432.f, 554.f, Y_AXIS);
  lights.pdf.push_back(rect_pdf.createPDF(app.context));
  lights.sample.push_back(rect_pdf.createSample(app.context));
  lights.emissions.push_back(make_float3(7.f));
  setRayGenerationProgram(app.context, lights);
  setMissProgram(app.context, CONSTANT);
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  Hitable_List list;
  BRDF* ground = A.mat<Lambertian>(A.tex<Constant_Texture>(0.48f, 0.83f, 0.53f));
  for (int i = 0; i < 20; i++) {
    for (int j = 0; j < 20; j++) {
      const float w = 100.f;
      const float x0 = -1000 + i * w, z0 = -1000 + j * w, y0 = 0.f;
      const float x1 = x0 + w, y1 = 100 * (rnd() + 0.01f), z1 = z0 + w;
      list.push(A.hit<Box>(make_float3(x0, y0, z0), make_float3(x1, y1, z1), ground));
    }
  }
  list.push(A.hit<AARect>(113.f, 443.f, 127.f, 432.f, 554.f, true, Y_AXIS,
                          A.mat<Diffuse_Light>(A.tex<Constant_Texture>(7.f))));
  list.push(A.hit<Sphere>(make_float3(400.f, 400.f, 200.f), 50.f,
                          A.mat<Lambertian>(A.tex<Constant_Texture>(0.7f, 0.3f, 0.1f))));
  Texture* glassTx1 = A.tex<Constant_Texture>(1.f);
  BRDF* glass = A.mat<Dielectric>(glassTx1, glassTx1, 1.5f, 0.f);
  list.push(A.hit<Sphere>(make_float3(260.f, 150.f, 45.f), 50.f, glass));
  list.push(A.hit<Sphere>(make_float3(0.f, 150.f, 145.f), 50.f,
                          A.mat<Metal>(A.tex<Constant_Texture>(0.8f, 0.8f, 0.9f), 10.f)));
  list.push(A.hit<Sphere>(make_float3(360.f, 150.f, 45.f), 70.f, glass));
  list.push(A.hit<Volumetric_Sphere>(make_float3(360.f, 150.f, 45.f), 70.f, 0.2f,
                                     A.mat<Isotropic>(A.tex<Constant_Texture>(0.2f, 0.4f, 0.9f))));
  list.push(A.hit<Volumetric_Sphere>(make_float3(0.f), 5000.f, 0.0001f, A.mat<Isotropic>(glassTx1)));
  Texture* earth = A.tex<Gradient_Texture>(make_float3(0.1f, 0.3f, 0.8f), make_float3(0.2f, 0.7f, 0.2f),
                                           make_float3(0.9f, 0.9f, 0.95f));
  list.push(A.hit<Sphere>(make_float3(400.f, 200.f, 400.f), 100.f, A.mat<Lambertian>(earth)));
  list.push(A.hit<Sphere>(make_float3(220.f, 280.f, 300.f), 80.f, A.mat<Lambertian>(A.tex<Noise_Texture>(0.1f))));
  Hitable_List spheres;
  BRDF* whiteMt = A.mat<Lambertian>(A.tex<Constant_Texture>(0.73f));
  for (int j = 0; j < 1000; j++) {
    const float x = 165 * rnd();
    const float y = 165 * rnd();
    const float z = 165 * rnd();
    spheres.push(A.hit<Sphere>(make_float3(x, y, z), 10.f, whiteMt));
  }
  spheres.translate(make_float3(-100.f, 270.f, 395.f));
  spheres.rotate(15.f, Y_AXIS);
  spheres.addListTo(group, app.context);
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  Camera camera(make_float3(478.f, 278.f, -600.f), make_float3(278.f, 278.f, 0.f), make_float3(0.f, 1.f, 0.f), 40.f,
                float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  return ConfigDefaults{500, 500, 64, 50, true, 11};
}

// ------------------------------------------------------------ meshes
// Perlin's improved gradient noise (Perlin, "Improving Noise", SIGGRAPH
// 2002) over a permutation shuffled by mt19937(seed), and its fBm sum.
struct GradientNoise {
  int perm[512];
  explicit GradientNoise(uint32_t seed) {
    std::mt19937 g(seed);
    int p[256];
    std::iota(p, p + 256, 0);
    for (int i = 255; i > 0; i--) std::swap(p[i], p[g() % (uint32_t)(i + 1)]);
    for (int i = 0; i < 512; i++) perm[i] = p[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    double u = h < 8 ? x : y;
    double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double noise(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx;
    y -= fy;
    z -= fz;
    double u = fade(x), v = fade(y), w = fade(z);
    int A = perm[X] + Y, AA = perm[A] + Z, AB = perm[A + 1] + Z;
    int B = perm[X + 1] + Y, BA = perm[B] + Z, BB = perm[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                     lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))));
  }
  double fbm(double x, double y, double z, int octaves) const {
    double sum = 0.0, amp = 0.5, f = 1.0;
    for (int o = 0; o < octaves; o++) {
      sum += amp * noise(x * f, y * f, z * f);
      amp *= 0.5;
      f *= 2.0;
    }
    return sum;
  }
};

// Octahedron on the unit sphere subdivided `levels` times (8 * 4^levels
// faces), outward winding; vertices in double precision.
inline void octasphere(int levels, std::vector<std::array<double, 3>>& V, std::vector<uint3>& F) {
  V = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  F = {make_uint3(0, 2, 4), make_uint3(4, 2, 1), make_uint3(1, 2, 5), make_uint3(5, 2, 0),
       make_uint3(4, 3, 0), make_uint3(1, 3, 4), make_uint3(5, 3, 1), make_uint3(0, 3, 5)};
  for (int l = 0; l < levels; l++) {
    std::unordered_map<uint64_t, uint32_t> mid;
    mid.reserve(F.size() * 2);
    auto midpoint = [&](uint32_t a, uint32_t b) {
      uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      std::array<double, 3> m{V[a][0] + V[b][0], V[a][1] + V[b][1], V[a][2] + V[b][2]};
      double n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      V.push_back({m[0] / n, m[1] / n, m[2] / n});
      uint32_t idx = (uint32_t)V.size() - 1;
      mid.emplace(key, idx);
      return idx;
    };
    std::vector<uint3> nf;
    nf.reserve(F.size() * 4);
    for (const uint3& t : F) {
      uint32_t ab = midpoint(t.x, t.y), bc = midpoint(t.y, t.z), ca = midpoint(t.z, t.x);
      nf.push_back(make_uint3(t.x, ab, ca));
      nf.push_back(make_uint3(t.y, bc, ab));
      nf.push_back(make_uint3(t.z, ca, bc));
      nf.push_back(make_uint3(ab, bc, ca));
    }
    F.swap(nf);
  }
}

// Displaced octasphere: p (1 + amp fBm(2p)), with area-weighted vertex
// normals of the displaced surface.
inline void displaced_sphere(int levels, uint32_t seed, double amp, std::vector<float3>& v, std::vector<uint3>& f,
                             std::vector<float3>& n) {
  std::vector<std::array<double, 3>> V;
  octasphere(levels, V, f);
  GradientNoise noise(seed);
  std::vector<std::array<double, 3>> P(V.size());
  for (size_t i = 0; i < V.size(); i++) {
    double s = 1.0 + amp * noise.fbm(2.0 * V[i][0], 2.0 * V[i][1], 2.0 * V[i][2], 4);
    P[i] = {V[i][0] * s, V[i][1] * s, V[i][2] * s};
  }
  std::vector<std::array<double, 3>> N(V.size(), {0, 0, 0});
  for (const uint3& t : f) {
    const auto &a = P[t.x], &b = P[t.y], &c = P[t.z];
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    for (uint32_t k : {t.x, t.y, t.z})
      for (int c3 = 0; c3 < 3; c3++) N[k][c3] += cr[c3];
  }
  v.resize(V.size());
  n.resize(V.size());
  for (size_t i = 0; i < V.size(); i++) {
    v[i] = make_float3((float)P[i][0], (float)P[i][1], (float)P[i][2]);
    double l = std::sqrt(N[i][0] * N[i][0] + N[i][1] * N[i][1] + N[i][2] * N[i][2]);
    n[i] = make_float3((float)(N[i][0] / l), (float)(N[i][1] / l), (float)(N[i][2] / l));
  }
}

// Synthetic equirectangular sky (spherical mapping of miss.cu:64-95: row j
// at theta = pi (j + 0.5) / H from +Y, column i at phi = 2 pi (i + 0.5) / W
// from +X toward +Z): horizon-to-zenith gradient, darker ground, and a
// Gaussian sun lobe of peak `sun_peak` and standard deviation `sun_deg`.
inline std::vector<float> synthetic_sky(uint32_t W, uint32_t H, double sun_el_deg, double sun_az_deg, double sun_peak,
                                        double sun_deg) {
  std::vector<float> img((size_t)W * H * 4);
  const double PI = 3.14159265358979323846;
  double el = sun_el_deg * PI / 180.0, az = sun_az_deg * PI / 180.0;
  double sd[3] = {std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)};
  double sig = sun_deg * PI / 180.0;
  for (uint32_t j = 0; j < H; j++) {
    double th = PI * (j + 0.5) / H;
    for (uint32_t i = 0; i < W; i++) {
      double ph = 2.0 * PI * (i + 0.5) / W;
      double d[3] = {std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
      double c[3];
      if (d[1] >= 0.0) {
        double t = std::sqrt(d[1]);
        const double hz[3] = {1.0, 1.0, 1.05}, zn[3] = {0.25, 0.45, 1.0};
        for (int k = 0; k < 3; k++) c[k] = (1.0 - t) * hz[k] + t * zn[k];
      } else {
        const double gr[3] = {0.20, 0.18, 0.16};
        for (int k = 0; k < 3; k++) c[k] = gr[k];
      }
      double cs = std::fmax(-1.0, std::fmin(1.0, d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2]));
      double a = std::acos(cs);
      double sun = sun_peak * std::exp(-0.5 * (a / sig) * (a / sig));
      const double tint[3] = {1.0, 0.95, 0.85};
      float* o = &img[4 * ((size_t)j * W + i)];
      for (int k = 0; k < 3; k++) o[k] = (float)(c[k] + sun * tint[k]);
      o[3] = 0.f;
    }
  }
  return img;
}

// Triangle-mesh scene.  `levels` is the octasphere subdivision (7 gives
// 2^17 triangles per mesh); smaller values give the same layout with coarser
// meshes for quick checks.
inline ConfigDefaults MeshScene(App_State& app, SceneArena& A, int levels = 7, uint32_t sky_w = 2048,
                                uint32_t sky_h = 1024) {
  Light_Sampler lights;
  Environment_PDF env;
  lights.pdf.push_back(env.createPDF(app.context));
  lights.sample.push_back(env.createSample(app.context));
  lights.emissions.push_back(make_float3(1.f));
  setRayGenerationProgram(app.context, lights);
  HDR_Texture* sky = A.tex<HDR_Texture>(sky_w, sky_h, synthetic_sky(sky_w, sky_h, 35.0, 40.0, 5e4, 0.5));
  setMissProgram(app.context, HDR, sky, true);
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  rnd_reseed(2);
  Hitable_List list;
  Mesh_List meshes;
  double ymin = 1e30;
  const double PI = 3.14159265358979323846;
  for (uint32_t i = 0; i < 8; i++) {
    float s = 0.5f + 1.5f * rnd();
    float ang = 360.f * rnd();
    float tx = -10.f + 20.f * rnd();
    float ty = -10.f + 20.f * rnd();
    float tz = -10.f + 20.f * rnd();
    BRDF* m;
    switch (i % 4) {
      case 0: {
        float r = 0.05f + 0.55f * rnd();
        float c0 = 0.4f + 0.6f * rnd(), c1 = 0.4f + 0.6f * rnd(), c2 = 0.4f + 0.6f * rnd();
        m = A.mat<Torrance_Sparrow>(A.tex<Constant_Texture>(c0, c1, c2), r, r);
        break;
      }
      case 1: {
        float nu = 10.f + 990.f * rnd(), nv = 10.f + 990.f * rnd();
        float d0 = 0.1f + 0.8f * rnd(), d1 = 0.1f + 0.8f * rnd(), d2 = 0.1f + 0.8f * rnd();
        float sp = 0.04f + 0.26f * rnd();
        m = A.mat<Ashikhmin_Shirley>(A.tex<Constant_Texture>(d0, d1, d2), A.tex<Constant_Texture>(sp), nu, nv);
        break;
      }
      case 2: {
        float sigma = (float)(30.0 * PI / 180.0) * rnd();
        float c0 = 0.2f + 0.7f * rnd(), c1 = 0.2f + 0.7f * rnd(), c2 = 0.2f + 0.7f * rnd();
        m = A.mat<Oren_Nayar>(A.tex<Constant_Texture>(c0, c1, c2), sigma);
        break;
      }
      default: {
        Texture* one = A.tex<Constant_Texture>(1.f);
        m = A.mat<Dielectric>(one, one, 1.5f, 0.f);
        break;
      }
    }
    std::vector<float3> v, n;
    std::vector<uint3> f;
    displaced_sphere(levels, i, 0.05, v, f, n);
    // world y of every vertex under T(t) R_y(ang) S(s) (scenes.hpp:428-430 order)
    for (const float3& p : v) ymin = std::fmin(ymin, (double)ty + (double)s * (double)p.y);
    A.meshes.emplace_back(new Mesh(std::move(v), std::move(f), m, std::move(n)));
    Mesh* mesh = A.meshes.back().get();
    mesh->scale(make_float3(s));
    mesh->rotate(ang, Y_AXIS);
    mesh->translate(make_float3(tx, ty, tz));
    meshes.push(mesh);
  }
  meshes.addElementsTo(group, app.context);
  BRDF* ground = A.mat<Lambertian>(A.tex<Constant_Texture>(0.5f));
  list.push(A.hit<AARect>(-20.f, 20.f, -20.f, 20.f, (float)ymin, false, Y_AXIS, ground));
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  Camera camera(make_float3(20.f, 4.f, 0.f), make_float3(0.f), make_float3(0.f, 1.f, 0.f), 45.f,
                float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  ConfigDefaults d{1920, 1080, 256, 12, true, 11};
  d.flags = PT_FLAG_MATERIAL_SORT | PT_FLAG_ENV_MIS;
  return d;
}

// ---------------------------------------------------------------- C3
// The 64-scene batch workload: scenes 0..31 are sphere scenes of 50..5000
// spheres, 32..63 mesh scenes of 10k..2M triangles (both log-uniform), each
// at a square resolution of side log-uniform in [256, 1024] and 4, 16 or 64
// samples per pixel.  The parameters come from one mt19937(3) stream (three
// 24-bit draws per scene, in scene order); scene i's content from the scene
// RNG reseeded with 3000 + i.
struct BatchSceneSpec {
  bool spheres;
  uint32_t count;  // spheres or triangles asked for
  uint32_t side;   // image width = height
  uint32_t spp;
};
constexpr uint32_t BATCH_SCENES = 64;
inline BatchSceneSpec batch_scene_spec(uint32_t i) {
  std::mt19937 g(3);
  auto u = [&g]() { return (double)(g() >> 8) * (1.0 / 16777216.0); };
  BatchSceneSpec s{};
  for (uint32_t k = 0; k <= i; k++) {
    const double a = u(), b = u(), c = u();
    s.spheres = k < BATCH_SCENES / 2;
    s.count = s.spheres ? (uint32_t)std::lround(50.0 * std::pow(100.0, a))
                        : (uint32_t)std::lround(1e4 * std::pow(200.0, a));
    s.side = (uint32_t)std::lround(256.0 * std::pow(4.0, b));
    const uint32_t spps[3] = {4, 16, 64};
    s.spp = spps[std::min(2, (int)(c * 3.0))];
  }
  return s;
}

// Sphere scene of `count` spheres: the random-spheres recipe (ground sphere,
// 80 % Lambertian, 15 % metal, 5 % glass, gradient sky) on a jittered
// ceil(sqrt(count))^2 grid, camera pulled back with the grid.
inline ConfigDefaults BatchSpheres(App_State& app, SceneArena& A, uint32_t seed, uint32_t count) {
  rnd_reseed(seed);
  Light_Sampler lights;
  setRayGenerationProgram(app.context, lights);
  setMissProgram(app.context, GRADIENT, make_float3(1.f), make_float3(0.5f, 0.7f, 1.f));
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  Hitable_List list;
  list.push(A.hit<Sphere>(make_float3(0.f, -1000.f, 0.f), 1000.f,
                          A.mat<Lambertian>(A.tex<Constant_Texture>(0.5f))));
  const int K = (int)std::ceil(std::sqrt((double)count));
  uint32_t made = 0;
  for (int a = -K / 2; a < K - K / 2 && made < count; a++) {
    for (int b = -K / 2; b < K - K / 2 && made < count; b++, made++) {
      float choose_mat = rnd();
      float3 center = make_float3((float)a + 0.9f * rnd(), 0.2f, (float)b + 0.9f * rnd());
      BRDF* m;
      if (choose_mat < 0.8f) {
        float r = rnd() * rnd(), g = rnd() * rnd(), bl = rnd() * rnd();
        m = A.mat<Lambertian>(A.tex<Constant_Texture>(r, g, bl));
      } else if (choose_mat < 0.95f) {
        float r = 0.5f * (1.f + rnd()), g = 0.5f * (1.f + rnd()), bl = 0.5f * (1.f + rnd());
        m = A.mat<Metal>(A.tex<Constant_Texture>(r, g, bl), 0.5f * rnd());
      } else {
        Texture* one = A.tex<Constant_Texture>(1.f);
        m = A.mat<Dielectric>(one, one, 1.5f, 0.f);
      }
      list.push(A.hit<Sphere>(center, 0.2f, m));
    }
  }
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  const float pull = std::max(1.f, (float)K / 22.f);
  Camera camera(make_float3(13.f * pull, 2.f * pull, 3.f * pull), make_float3(0.f), make_float3(0.f, 1.f, 0.f),
                20.f, float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  ConfigDefaults d{app.W, app.H, 16, 50, true, 11};
  d.flags = PT_FLAG_MATERIAL_SORT;
  return d;
}

// Displaced octasphere of subdivision `levels` and noise seed `seed`,
// generated once per (levels, seed) and shared by every scene that uses it.
inline void cached_displaced_sphere(int levels, uint32_t seed, std::vector<float3>& v, std::vector<uint3>& f,
                                    std::vector<float3>& n) {
  struct Entry {
    std::vector<float3> v, n;
    std::vector<uint3> f;
  };
  static std::map<std::pair<int, uint32_t>, Entry> cache;
  auto key = std::make_pair(levels, seed);
  auto it = cache.find(key);
  if (it == cache.end()) {
    Entry e;
    displaced_sphere(levels, seed, 0.05, e.v, e.f, e.n);
    it = cache.emplace(key, std::move(e)).first;
  }
  v = it->second.v;
  f = it->second.f;
  n = it->second.n;
}

// Mesh scene of about `count` triangles: M displaced octaspheres of
// subdivision L (8 * 4^L triangles each, L in [2, 7], M in [1, 32]) with
// the mesh-scene materials and random similarity transforms, a ground quad,
// and a 256x128 synthetic sky importance-sampled as an environment light.
inline ConfigDefaults BatchMeshes(App_State& app, SceneArena& A, uint32_t seed, uint32_t count) {
  Light_Sampler lights;
  Environment_PDF env;
  lights.pdf.push_back(env.createPDF(app.context));
  lights.sample.push_back(env.createSample(app.context));
  lights.emissions.push_back(make_float3(1.f));
  setRayGenerationProgram(app.context, lights);
  HDR_Texture* sky = A.tex<HDR_Texture>(256u, 128u, synthetic_sky(256, 128, 35.0, 40.0, 5e4, 0.5));
  setMissProgram(app.context, HDR, sky, true);
  setExceptionProgram(app.context);
  Group group = app.context->createGroup();
  group->setAcceleration(app.context->createAcceleration("Trbvh"));
  int L = (int)std::floor(std::log((double)count / 64.0) / std::log(4.0));
  L = std::max(2, std::min(7, L));
  const double per = 8.0 * std::pow(4.0, L);
  const uint32_t M = (uint32_t)std::max(1.0, std::min(32.0, std::round((double)count / per)));
  rnd_reseed(seed);
  Hitable_List list;
  Mesh_List meshes;
  double ymin = 1e30;
  const double PI = 3.14159265358979323846;
  const float spread = 10.f * std::sqrt((float)M / 8.f);
  for (uint32_t i = 0; i < M; i++) {
    float s = 0.5f + 1.5f * rnd();
    float ang = 360.f * rnd();
    float tx = spread * (2.f * rnd() - 1.f), ty = spread * (2.f * rnd() - 1.f), tz = spread * (2.f * rnd() - 1.f);
    BRDF* m;
    switch (i % 4) {
      case 0: {
        float r = 0.05f + 0.55f * rnd();
        float c0 = 0.4f + 0.6f * rnd(), c1 = 0.4f + 0.6f * rnd(), c2 = 0.4f + 0.6f * rnd();
        m = A.mat<Torrance_Sparrow>(A.tex<Constant_Texture>(c0, c1, c2), r, r);
        break;
      }
      case 1: {
        float nu = 10.f + 990.f * rnd(), nv = 10.f + 990.f * rnd();
        float d0 = 0.1f + 0.8f * rnd(), d1 = 0.1f + 0.8f * rnd(), d2 = 0.1f + 0.8f * rnd();
        float sp = 0.04f + 0.26f * rnd();
        m = A.mat<Ashikhmin_Shirley>(A.tex<Constant_Texture>(d0, d1, d2), A.tex<Constant_Texture>(sp), nu, nv);
        break;
      }
      case 2: {
        float sigma = (float)(30.0 * PI / 180.0) * rnd();
        float c0 = 0.2f + 0.7f * rnd(), c1 = 0.2f + 0.7f * rnd(), c2 = 0.2f + 0.7f * rnd();
        m = A.mat<Oren_Nayar>(A.tex<Constant_Texture>(c0, c1, c2), sigma);
        break;
      }
      default: {
        Texture* one = A.tex<Constant_Texture>(1.f);
        m = A.mat<Dielectric>(one, one, 1.5f, 0.f);
        break;
      }
    }
    std::vector<float3> v, n;
    std::vector<uint3> f;
    cached_displaced_sphere(L, i, v, f, n);
    for (const float3& p : v) ymin = std::fmin(ymin, (double)ty + (double)s * (double)p.y);
    A.meshes.emplace_back(new Mesh(std::move(v), std::move(f), m, std::move(n)));
    Mesh* mesh = A.meshes.back().get();
    mesh->scale(make_float3(s));
    mesh->rotate(ang, Y_AXIS);
    mesh->translate(make_float3(tx, ty, tz));
    meshes.push(mesh);
  }
  meshes.addElementsTo(group, app.context);
  const float g = 2.f * spread;
  list.push(A.hit<AARect>(-g, g, -g, g, (float)ymin, false, Y_AXIS, A.mat<Lambertian>(A.tex<Constant_Texture>(0.5f))));
  list.addElementsTo(group, app.context);
  app.context["world"]->set(group);
  Camera camera(make_float3(2.f * spread, 0.4f * spread, 0.f), make_float3(0.f), make_float3(0.f, 1.f, 0.f), 45.f,
                float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
  camera.set(app.context);
  ConfigDefaults d{app.W, app.H, 16, 12, true, 11};
  d.flags = PT_FLAG_MATERIAL_SORT | PT_FLAG_ENV_MIS;
  return d;
}

// Scene i of the batch workload at its own resolution (or W x H if given).
inline ConfigDefaults BatchScene(App_State& app, SceneArena& A, uint32_t i) {
  const BatchSceneSpec s = batch_scene_spec(i);
  ConfigDefaults d = s.spheres ? BatchSpheres(app, A, 3000 + i, s.count) : BatchMeshes(app, A, 3000 + i, s.count);
  d.spp = (int)s.spp;
  return d;
}
