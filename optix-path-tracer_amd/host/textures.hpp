// textures.hpp -- host texture classes (host_includes/textures.hpp).
// Constant_Texture and HDR_Texture (file or in-memory float data) are
// supported by the GPU path; the procedural textures of the reference
// (Checker, Noise, Gradient, Vector) throw until their device programs exist.
#pragma once

#include <fstream>

#include "buffers.hpp"
#include "host_common.hpp"
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
struct HDR_Texture : public Texture {
  HDR_Texture(const std::string f) : fileName(f) {}
  HDR_Texture(uint32_t w, uint32_t h, std::vector<float> rgba) : width(w), height(h) {
    data = std::make_shared<std::vector<float>>(std::move(rgba));
  }
  Program assignTo(Context& g_context) const override {
    if (!data) {
      RGBEImage img;
      std::string err;
      if (!rgbe_load(fileName, img, err)) throw std::string("HDR Image is invalid or hasn't been found: ") + err;
      width = img.width;
      height = img.height;
      data = std::make_shared<std::vector<float>>((size_t)width * height * 4);
      for (size_t i = 0; i < (size_t)width * height; i++) {
        (*data)[4 * i] = img.rgb[3 * i];
        (*data)[4 * i + 1] = img.rgb[3 * i + 1];
        (*data)[4 * i + 2] = img.rgb[3 * i + 2];
        (*data)[4 * i + 3] = 0.f;
      }
    }
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
struct Image_Texture : Unsupported_Texture {
  Image_Texture(const std::string) : Unsupported_Texture("Image_Texture") {}
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
