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
#pragma once

#include <memory>
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
