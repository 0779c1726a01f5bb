// pdfs.hpp -- light sample / pdf programs of the host scene API
// (host_includes/pdfs.hpp:17-110).  The Programs carry the light geometry;
// setRayGenerationProgram turns them into the scene's light table.
#pragma once

#include "host_common.hpp"
#include "programs.hpp"

struct PDF {
  virtual ~PDF() = default;
  virtual Program createSample(Context& g_context) const = 0;
  virtual Program createPDF(Context& g_context) const = 0;
};

inline Program light_program(ptfacade::ProgramObj::Kind kind, const pt_light_desc& l) {
  auto p = std::make_shared<ptfacade::ProgramObj>();
  p->kind = kind;
  p->light = l;
  return Program(p);
}

// pdfs.hpp:17-80
struct Rectangle_PDF : public PDF {
  Rectangle_PDF(const float aa0, const float aa1, const float bb0, const float bb1, const float kk, const AXIS aax)
      : a0(aa0), a1(aa1), b0(bb0), b1(bb1), k(kk), ax(aax) {}
  pt_light_desc desc() const {
    pt_light_desc l{};
    l.type = PT_LIGHT_RECT;
    l.axis = (int)ax;
    l.a0 = a0; l.a1 = a1; l.b0 = b0; l.b1 = b1; l.k = k;
    return l;
  }
  Program createSample(Context&) const override { return light_program(ptfacade::ProgramObj::LIGHT_SAMPLE, desc()); }
  Program createPDF(Context&) const override { return light_program(ptfacade::ProgramObj::LIGHT_PDF, desc()); }
  float a0, a1, b0, b1, k;
  AXIS ax;
};

// pdfs.hpp:82-110
struct Sphere_PDF : public PDF {
  Sphere_PDF(const float3 c, const float r) : radius(r), center(c) {}
  pt_light_desc desc() const {
    pt_light_desc l{};
    l.type = PT_LIGHT_SPHERE;
    l.center[0] = center.x; l.center[1] = center.y; l.center[2] = center.z;
    l.radius = radius;
    return l;
  }
  Program createSample(Context&) const override { return light_program(ptfacade::ProgramObj::LIGHT_SAMPLE, desc()); }
  Program createPDF(Context&) const override { return light_program(ptfacade::ProgramObj::LIGHT_PDF, desc()); }
  float radius;
  float3 center;
};

// Importance sampling of the environment map set by setMissProgram(HDR, ...).
// No counterpart in the reference, which only looks the map up on a miss
// (miss.cu:64-95): with PT_FLAG_ENV_MIS the renderer samples the map in
// proportion to luminance x sin(theta) and weights misses by MIS.  The
// radiance is the miss program's own (x2); the emission entry is unused.
struct Environment_PDF : public PDF {
  pt_light_desc desc() const {
    pt_light_desc l{};
    l.type = PT_LIGHT_ENV;
    return l;
  }
  Program createSample(Context&) const override { return light_program(ptfacade::ProgramObj::LIGHT_SAMPLE, desc()); }
  Program createPDF(Context&) const override { return light_program(ptfacade::ProgramObj::LIGHT_PDF, desc()); }
};
