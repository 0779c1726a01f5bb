// materials.hpp -- host BRDF classes (host_includes/materials.hpp:27-243),
// recorded as pt_material_desc entries.  Parameter mappings are the
// reference's: Ashikhmin-Shirley clamps nu, nv to >= 1 (:173-174),
// Oren-Nayar maps R to A, B through sigma = saturate(R) (:192-198),
// Torrance-Sparrow maps roughness r to alpha = max(r, 1e-3)^2 (:235-238).
#pragma once

#include "host_common.hpp"
#include "textures.hpp"

struct BRDF {
  virtual ~BRDF() = default;
  virtual Material assignTo(Context& g_context) const = 0;

 protected:
  Material record(Context& g_context, const pt_material_desc& d) const {
    auto m = std::make_shared<ptfacade::MaterialObj>();
    m->index = g_context->rec.add_material(this, d);
    return Material(m);
  }
  static int tex_index(const Texture* t, Context& g_context) {
    if (!t) throw "Material texture is NULL";
    return t->assignTo(g_context)->texture;
  }
};

// materials.hpp:43-58
struct Lambertian : public BRDF {
  Lambertian(const Texture* t) : texture(t) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_LAMBERTIAN, {tex_index(texture, g_context), -1}, {0, 0, 0, 0}};
    return record(g_context, d);
  }
  const Texture* texture;
};

// materials.hpp:61-78
struct Metal : public BRDF {
  Metal(const Texture* t, const float fuzz) : texture(t), fuzz(fuzz) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_METAL, {tex_index(texture, g_context), -1}, {fuzz, 0, 0, 0}};
    return record(g_context, d);
  }
  const Texture* texture;
  const float fuzz;
};

// materials.hpp:81-103
struct Dielectric : public BRDF {
  Dielectric(const Texture* baseTex, const Texture* extTex, const float ref_idx, const float density = 0.f)
      : baseTex(baseTex), extTex(extTex), ref_idx(ref_idx), density(density) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_DIELECTRIC, {tex_index(baseTex, g_context), tex_index(extTex, g_context)},
                       {ref_idx, density, 0, 0}};
    return record(g_context, d);
  }
  const Texture *baseTex, *extTex;
  const float ref_idx;
  const float density;
};

// materials.hpp:106-122
struct Diffuse_Light : public BRDF {
  Diffuse_Light(const Texture* t) : texture(t) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_DIFFUSE_LIGHT, {tex_index(texture, g_context), -1}, {0, 0, 0, 0}};
    return record(g_context, d);
  }
  const Texture* texture;
};

// materials.hpp:162-188
struct Ashikhmin_Shirley : public BRDF {
  Ashikhmin_Shirley(const Texture* diffuse_tex, const Texture* specular_tex, const float nu, const float nv)
      : specular_tex(specular_tex), diffuse_tex(diffuse_tex), nu(nu), nv(nv) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_ASHIKHMIN_SHIRLEY,
                       {tex_index(diffuse_tex, g_context), tex_index(specular_tex, g_context)},
                       {std::fmax(1.f, nu), std::fmax(1.f, nv), 0, 0}};
    return record(g_context, d);
  }
  float roughnessToAlpha(float roughness) const { return 2.0f / (roughness * roughness) - 2.0f; }
  const Texture* specular_tex;
  const Texture* diffuse_tex;
  const float nu, nv;
};

// materials.hpp:191-215
struct Oren_Nayar : public BRDF {
  Oren_Nayar(const Texture* t, const float R) : texture(t) {
    float sigma = saturate(R);
    float div = 1.f / (PI_F + ((3.f * PI_F - 4.f) / 6.f) * sigma);
    rA = 1.0f * div;
    rB = sigma * div;
  }
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_OREN_NAYAR, {tex_index(texture, g_context), -1}, {rA, rB, 0, 0}};
    return record(g_context, d);
  }
  const Texture* texture;
  float rA, rB;
};

// materials.hpp:218-242
struct Torrance_Sparrow : public BRDF {
  Torrance_Sparrow(const Texture* texture, const float nu, const float nv) : texture(texture), nu(nu), nv(nv) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_TORRANCE_SPARROW, {tex_index(texture, g_context), -1},
                       {roughnessToAlpha(nu), roughnessToAlpha(nv), 0, 0}};
    return record(g_context, d);
  }
  float roughnessToAlpha(float roughness) const {
    float R = std::fmax(roughness, 1e-3f);
    return R * R;
  }
  const Texture* texture;
  const float nu, nv;
};

// materials.hpp:125-140: the phase function of a constant-density medium
struct Isotropic : public BRDF {
  Isotropic(const Texture* t) : texture(t) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_ISOTROPIC, {tex_index(texture, g_context), -1}, {0, 0, 0, 0}};
    return record(g_context, d);
  }
  const Texture* texture;
};
// materials.hpp:143-159: 0.5 N + 0.5 of the shading or geometric normal
struct Normal_Shader : public BRDF {
  Normal_Shader(const bool useShadingNormal = false) : useShadingNormal(useShadingNormal) {}
  Material assignTo(Context& g_context) const override {
    pt_material_desc d{PT_MAT_NORMAL, {-1, -1}, {useShadingNormal ? 1.f : 0.f, 0, 0, 0}};
    return record(g_context, d);
  }
  const bool useShadingNormal;
};
