// textures.hpp -- host texture classes (host_includes/textures.hpp).
// Constant_Texture and HDR_Texture (file or in-memory float data) are
// supported by the GPU path; the procedural textures of the reference
// (Checker, Noise, Gradient, Vector) throw until their device programs exist.
#pragma once

#include <fstream>

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

// Texture kinds of the reference without a device program yet.
struct Unsupported_Texture : public Texture {
  explicit Unsupported_Texture(const char* what) : what(what) {}
  Program assignTo(Context&) const override { throw std::string(what) + " is not supported by the GPU path yet"; }
  const char* what;
};
struct Checker_Texture : Unsupported_Texture {
  Checker_Texture(const Texture*, const Texture*) : Unsupported_Texture("Checker_Texture") {}
};
struct Noise_Texture : Unsupported_Texture {
  Noise_Texture(const float, const AXIS = X_AXIS) : Unsupported_Texture("Noise_Texture") {}
};
struct Gradient_Texture : Unsupported_Texture {
  Gradient_Texture(const float3&, const float3&, const float3&) : Unsupported_Texture("Gradient_Texture") {}
};
struct Vector_Texture : Unsupported_Texture {
  Vector_Texture(const std::vector<Texture*>&) : Unsupported_Texture("Vector_Texture") {}
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
