// textures.hpp -- host texture classes (host_includes/textures.hpp).
// Constant, Checker, Noise, Image, HDR (file or in-memory float data),
// Gradient and Vector textures, all evaluated by the GPU path.
#pragma once

#include <fstream>
#include <utility>

#include "buffers.hpp"
#include "host_common.hpp"
#include "image_load.hpp"
#include "rgbe.hpp"

struct Texture {
  virtual ~Texture() = default;
  virtual Program assignTo(Context& g_context) const = 0;
};

inline Program texture_program(int index) {
  auto p = std::make_shared<ptfacade::ProgramObj>();
  p->kind = ptfacade::ProgramObj::TEXTURE;
  p->texture = index;
  return Program(p);
}

// textures.hpp:20-37
struct Constant_Texture : public Texture {
  Constant_Texture(const float3& c) : color(c) {}
  Constant_Texture(const float& rgb) : color(make_float3(rgb)) {}
  Constant_Texture(const float& r, const float& g, const float& b) : color(make_float3(r, g, b)) {}
  Program assignTo(Context& g_context) const override {
    pt_texture_desc t{};
    t.type = PT_TEX_CONSTANT;
    t.rgb[0] = color.x;
    t.rgb[1] = color.y;
    t.rgb[2] = color.z;
    return texture_program(g_context->rec.add_texture(this, t, nullptr));
  }
  const float3 color;
};

// textures.hpp:177-230: float RGB(A) image, bilinear, repeat wrap.  The
// reference loads a Radiance .hdr file; an in-memory image (width x height
// float4, row 0 first) may be given instead.
// float RGBA texels of a Radiance .hdr file, rows in file order
inline std::shared_ptr<std::vector<float>> hdr_texels(const std::string& fileName, uint32_t& width, uint32_t& height) {
  RGBEImage img;
  std::string err;
  if (!rgbe_load(fileName, img, err)) throw std::string("HDR Image is invalid or hasn't been found: ") + err;
  width = img.width;
  height = img.height;
  auto data = std::make_shared<std::vector<float>>((size_t)width * height * 4);
  for (size_t i = 0; i < (size_t)width * height; i++) {
    (*data)[4 * i] = img.rgb[3 * i];
    (*data)[4 * i + 1] = img.rgb[3 * i + 1];
    (*data)[4 * i + 2] = img.rgb[3 * i + 2];
    (*data)[4 * i + 3] = 0.f;
  }
  return data;
}
struct HDR_Texture : public Texture {
  HDR_Texture(const std::string f) : fileName(f) {}
  HDR_Texture(uint32_t w, uint32_t h, std::vector<float> rgba) : width(w), height(h) {
    data = std::make_shared<std::vector<float>>(std::move(rgba));
  }
  Program assignTo(Context& g_context) const override {
    if (!data) data = hdr_texels(fileName, width, height);
    if (data->size() != (size_t)width * height * 4) throw "HDR texture data has the wrong size";
    pt_texture_desc t{};
    t.type = PT_TEX_IMAGE;
    t.width = width;
    t.height = height;
    return texture_program(g_context->rec.add_texture(this, t, data));
  }
  const std::string fileName;
  mutable uint32_t width = 0, height = 0;
  mutable std::shared_ptr<std::vector<float>> data;
};

// textures.hpp:114-175: an 8-bit image read as normalised float (byte /
// 255), bilinear, repeat wrap; rows flipped at load so that row 0 is the
// image's bottom row (textures.hpp:141-142).  Decoded by image_load.hpp
// (PNG, binary PPM/PGM) instead of stb_image.
inline std::shared_ptr<std::vector<float>> image_texels(const std::string& fileName, uint32_t& width,
                                                        uint32_t& height) {
  LdrImage img;
  std::string err;
  if (!load_image(fileName, img, err)) throw std::string("Image is invalid or hasn't been found: ") + err;
  width = img.width;
  height = img.height;
  auto data = std::make_shared<std::vector<float>>((size_t)width * height * 4);
  for (uint32_t j = 0; j < height; j++)
    for (uint32_t i = 0; i < width; i++) {
      const uint8_t* s = &img.rgba[4 * ((size_t)(height - 1 - j) * width + i)];
      float* d = &(*data)[4 * ((size_t)j * width + i)];
      for (int c = 0; c < 4; c++) d[c] = (float)s[c] * (1.0f / 255.0f);
    }
  return data;
}
struct Image_Texture : public Texture {
  Image_Texture(const std::string f) : fileName(f) {}
  Program assignTo(Context& g_context) const override {
    if (!data) data = image_texels(fileName, width, height);
    pt_texture_desc t{};
    t.type = PT_TEX_IMAGE;
    t.width = width;
    t.height = height;
    return texture_program(g_context->rec.add_texture(this, t, data));
  }
  const std::string fileName;
  mutable uint32_t width = 0, height = 0;
  mutable std::shared_ptr<std::vector<float>> data;
};

// textures.hpp:39-56: the odd texture where sin(10x) sin(10y) sin(10z) of
// the hit point is negative, else the even one.  The two sub-texture
// indices travel in width (odd) and height (even); neither may itself be a
// checker.
struct Checker_Texture : public Texture {
  Checker_Texture(const Texture* o, const Texture* e) : odd(o), even(e) {}
  Program assignTo(Context& g_context) const override {
    if (!odd || !even) throw "Checker_Texture needs two textures";
    pt_texture_desc t{};
    t.type = PT_TEX_CHECKER;
    t.width = (uint32_t)odd->assignTo(g_context)->texture;
    t.height = (uint32_t)even->assignTo(g_context)->texture;
    return texture_program(g_context->rec.add_texture(this, t, nullptr));
  }
  const Texture* odd;
  const Texture* even;
};

// textures.hpp:56-112: Perlin noise from 256 random unit vectors and three
// permutations of 0..255, drawn from rnd() at every assignTo as the
// reference does (a texture shared by two materials gets two tables).  The
// three components of a vector are drawn x, y, z (the reference passes the
// draws as function arguments, whose evaluation order C++ leaves open).
// The axis is kept for the interface only: the sampled value always uses z
// (noise_texture.cu).
struct Noise_Texture : public Texture {
  Noise_Texture(const float s, const AXIS ax = X_AXIS) : scale(s), ax(ax) {}
  static void permute(float* p) {
    for (int i = 0; i < 256; i++) p[i] = (float)i;
    for (int i = 256 - 1; i > 0; i--) {
      const int target = int(rnd() * (float)(i + 1));
      std::swap(p[i], p[target]);
    }
  }
  Program assignTo(Context& g_context) const override {
    auto data = std::make_shared<std::vector<float>>(PT_NOISE_FLOATS, 0.f);
    float* v = data->data();
    for (int i = 0; i < 256; i++) {
      const float x = -1.f + 2.f * rnd();
      const float y = -1.f + 2.f * rnd();
      const float z = -1.f + 2.f * rnd();
      const float l = std::sqrt(x * x + y * y + z * z);
      v[4 * i] = x / l;
      v[4 * i + 1] = y / l;
      v[4 * i + 2] = z / l;
    }
    for (int a = 0; a < 3; a++) permute(v + 1024 + 256 * a);
    pt_texture_desc t{};
    t.type = PT_TEX_NOISE;
    t.rgb[0] = scale;
    return texture_program(g_context->rec.add_texture(nullptr, t, data));
  }
  const float scale;
  const AXIS ax;
};

// textures.hpp:233-251: colours A, B, C weighted by |n| of the normalised
// hit point.
struct Gradient_Texture : public Texture {
  Gradient_Texture(const float3& a, const float3& b, const float3& c) : colorA(a), colorB(b), colorC(c) {}
  Program assignTo(Context& g_context) const override {
    auto data = std::make_shared<std::vector<float>>(std::vector<float>{
        colorA.x, colorA.y, colorA.z, 0.f, colorB.x, colorB.y, colorB.z, 0.f, colorC.x, colorC.y, colorC.z, 0.f});
    pt_texture_desc t{};
    t.type = PT_TEX_GRADIENT;
    return texture_program(g_context->rec.add_texture(this, t, data));
  }
  const float3 colorA, colorB, colorC;
};

// textures.hpp:253-272: the list a mesh's faces pick from by their
// material-buffer index (vector.cu); the children are assigned in order
// and must not be vectors themselves.
struct Vector_Texture : public Texture {
  Vector_Texture(const std::vector<Texture*>& tv) : texture_vector(tv) {}
  Program assignTo(Context& g_context) const override {
    if (texture_vector.empty()) throw "Vector_Texture needs at least one texture";
    auto data = std::make_shared<std::vector<float>>();
    for (const Texture* t : texture_vector) {
      if (!t) throw "Vector_Texture holds a NULL texture";
      data->push_back((float)t->assignTo(g_context)->texture);
    }
    pt_texture_desc d{};
    d.type = PT_TEX_VECTOR;
    d.width = (uint32_t)texture_vector.size();
    return texture_program(g_context->rec.add_texture(this, d, data));
  }
  const std::vector<Texture*> texture_vector;
};

// textures.hpp:274-290
struct Texture_List {
  Texture_List() {}
  int push(Texture* t) {
    int index = (int)texList.size();
    texList.push_back(t);
    return index;
  }
  Texture* operator[](const int i) { return texList[i]; }
  std::vector<Texture*> texList;
};
