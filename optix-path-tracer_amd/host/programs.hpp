// programs.hpp -- ray generation, miss and exception program set-up of the
// host scene API (host_includes/programs.hpp:17-97).
#pragma once

#include <string>

#include "buffers.hpp"
#include "host_common.hpp"
#include "textures.hpp"

// programs.hpp:17-30: binds the light table (Light_Sample, Light_PDF,
// Light_Emissions, numLights)
inline void setRayGenerationProgram(Context& g_context, Light_Sampler& lights) {
  auto& L = g_context->rec.lights;
  L.clear();
  if (lights.sample.size() != lights.emissions.size() || lights.pdf.size() != lights.emissions.size())
    throw "Light_Sampler: sample, pdf and emission lists differ in length";
  for (size_t i = 0; i < lights.emissions.size(); i++) {
    if (!lights.sample[i] || lights.sample[i]->kind != ptfacade::ProgramObj::LIGHT_SAMPLE)
      throw "Light_Sampler: sample program is not a light sample program";
    pt_light_desc l = lights.sample[i]->light;
    l.emission[0] = lights.emissions[i].x;
    l.emission[1] = lights.emissions[i].y;
    l.emission[2] = lights.emissions[i].z;
    L.push_back(l);
  }
  g_context["numLights"]->setInt((int)L.size());
  g_context->invalidate();
}

// programs.hpp:32
typedef enum { GRADIENT, CONSTANT, IMG, HDR } Miss_Programs;

// programs.hpp:35-63: image backgrounds from a file
// (the texture is recorded without an identity key: the loaded image is
// not shared with any texture object)
inline void setMissProgram(Context& g_context, Miss_Programs id, std::string fileName, bool isSpherical = true) {
  if (id != HDR && id != IMG) throw "Parameters invalid, miss program unknown or not yet implemented";
  pt_texture_desc t{};
  t.type = PT_TEX_IMAGE;
  auto texels = id == HDR ? hdr_texels(fileName, t.width, t.height) : image_texels(fileName, t.width, t.height);
  int ti = g_context->rec.add_texture(nullptr, t, texels);
  // miss.cu:64-99 environmental_mapping (HDR) / :51-59 image_background (IMG)
  g_context->rec.miss = pt_miss_desc{id == HDR ? PT_MISS_HDR : PT_MISS_IMAGE, {0, 0, 0}, {0, 0, 0}, ti,
                                     isSpherical ? 1 : 0};
  g_context->invalidate();
}

// Environment map from an in-memory HDR texture (same miss program as the
// file form above).
inline void setMissProgram(Context& g_context, Miss_Programs id, const Texture* tex, bool isSpherical = true) {
  if (id != HDR || !tex) throw "Parameters invalid, miss program unknown or not yet implemented";
  int t = tex->assignTo(g_context)->texture;
  g_context->rec.miss = pt_miss_desc{PT_MISS_HDR, {0, 0, 0}, {0, 0, 0}, t, isSpherical ? 1 : 0};
  g_context->invalidate();
}

// programs.hpp:65-92: color backgrounds
inline void setMissProgram(Context& g_context, Miss_Programs id, float3 colorValue1 = make_float3(0.f),
                           float3 colorValue2 = make_float3(0.f)) {
  pt_miss_desc m{};
  m.texture = -1;
  m.spherical = 1;
  m.color1[0] = colorValue1.x; m.color1[1] = colorValue1.y; m.color1[2] = colorValue1.z;
  m.color2[0] = colorValue2.x; m.color2[1] = colorValue2.y; m.color2[2] = colorValue2.z;
  if (id == GRADIENT) m.type = PT_MISS_GRADIENT;
  else if (id == CONSTANT) m.type = PT_MISS_CONSTANT;
  else throw "Parameters invalid, miss program unknown or not yet implemented";
  g_context->rec.miss = m;
  g_context->invalidate();
}

// programs.hpp:94-97: exceptions are reported as status codes by the library
inline void setExceptionProgram(Context& /*g_context*/) {}
