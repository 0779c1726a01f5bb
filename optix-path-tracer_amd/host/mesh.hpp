// mesh.hpp -- triangle meshes of the host scene API
// (host_includes/mesh.hpp:10-345).  Meshes are flattened into unindexed
// triangles (i, i+1, i+2) as mesh.hpp:110-149 does; OBJ/MTL files are read
// by obj_loader.hpp, in-memory meshes come from vertex/index arrays.
#pragma once

#include "hitables.hpp"
#include "obj_loader.hpp"

class Mesh {
 public:
  Mesh(std::string fileName, bool RTX) : RTX_MODE(RTX), givenMaterial(nullptr), fileName(fileName) {}
  Mesh(std::string fileName, std::string assetsFolder, bool RTX)
      : RTX_MODE(RTX), givenMaterial(nullptr), fileName(fileName), assetsFolder(assetsFolder) {}
  Mesh(std::string fileName, BRDF* givenMaterial, bool RTX)
      : RTX_MODE(RTX), givenMaterial(givenMaterial), fileName(fileName) {}
  Mesh(std::string fileName, std::string assetsFolder, BRDF* givenMaterial, bool RTX)
      : RTX_MODE(RTX), givenMaterial(givenMaterial), fileName(fileName), assetsFolder(assetsFolder) {}
  // in-memory mesh: vertices, triangles and optional per-vertex normals / uvs
  Mesh(std::vector<float3> vertices, std::vector<uint3> triangles, BRDF* material,
       std::vector<float3> normals = {}, std::vector<float2> uvs = {})
      : RTX_MODE(true), givenMaterial(material), mem_v(std::move(vertices)), mem_t(std::move(triangles)),
        mem_n(std::move(normals)), mem_uv(std::move(uvs)), in_memory(true) {}

  // mesh.hpp:49-204
  GeometryInstance getGeometryInstance(Context& g_context) {
    ptfacade::MeshData md;
    int mat = -1;
    if (in_memory) {
      if (!givenMaterial) throw "Mesh has no material";
      mat = givenMaterial->assignTo(g_context)->index;
      for (auto& v : mem_v) { md.v.push_back(v.x); md.v.push_back(v.y); md.v.push_back(v.z); }
      for (auto& n : mem_n) { md.n.push_back(n.x); md.n.push_back(n.y); md.n.push_back(n.z); }
      for (auto& t : mem_uv) { md.uv.push_back(t.x); md.uv.push_back(t.y); }
      for (auto& t : mem_t) {
        if (t.x >= mem_v.size() || t.y >= mem_v.size() || t.z >= mem_v.size()) throw "Mesh index out of range";
        md.idx.push_back(t.x); md.idx.push_back(t.y); md.idx.push_back(t.z);
      }
      if (!md.n.empty() && md.n.size() != md.v.size()) throw "Mesh normal count differs from vertex count";
      if (!md.uv.empty() && md.uv.size() / 2 != md.v.size() / 3) throw "Mesh uv count differs from vertex count";
    } else {
      ObjMesh om;
      std::string err;
      if (!obj_load(assetsFolder + fileName, assetsFolder, om, err)) throw std::string("Failed to load/parse .obj: ") + err;
      if (givenMaterial) {
        // mesh.hpp:142-145: with a given material every face's
        // material-buffer entry is 0 (a Vector_Texture picks its first child)
        mat = givenMaterial->assignTo(g_context)->index;
      } else {
        // one Lambertian per MTL material, textured by its diffuse map or
        // coloured by its ambient colour (mesh.hpp:77-102: one Lambertian
        // whose Vector_Texture picks the face's texture, the same image)
        std::vector<int> mats;
        for (auto& m : om.materials) {
          pt_texture_desc t{};
          int ti;
          if (!m.diffuse_texname.empty()) {
            uint32_t w = 0, h = 0;
            auto texels = image_texels(assetsFolder + m.diffuse_texname, w, h);
            t.type = PT_TEX_IMAGE;
            t.width = w;
            t.height = h;
            ti = g_context->rec.add_texture(nullptr, t, texels);
          } else {
            t.type = PT_TEX_CONSTANT;
            t.rgb[0] = m.ambient[0]; t.rgb[1] = m.ambient[1]; t.rgb[2] = m.ambient[2];
            ti = g_context->rec.add_texture(nullptr, t, nullptr);
          }
          pt_material_desc d{PT_MAT_LAMBERTIAN, {ti, -1}, {0, 0, 0, 0}};
          mats.push_back(g_context->rec.add_material(nullptr, d));
        }
        if (mats.empty()) {
          pt_texture_desc t{};
          t.type = PT_TEX_CONSTANT;
          int ti = g_context->rec.add_texture(nullptr, t, nullptr);
          pt_material_desc d{PT_MAT_LAMBERTIAN, {ti, -1}, {0, 0, 0, 0}};
          mats.push_back(g_context->rec.add_material(nullptr, d));
        }
        for (int f : om.face_material) md.fmat.push_back(mats[f < 0 ? 0 : f]);
        mat = mats[0];
      }
      md.v = std::move(om.v);
      md.n = std::move(om.n);
      md.uv = std::move(om.uv);
      md.idx = std::move(om.idx);
    }
    md.material = mat;
    int mi = (int)g_context->rec.meshes.size();
    g_context->rec.meshes.push_back(std::move(md));
    GeometryInstance gi = g_context->createGeometryInstance();
    gi->material = mat;
    ptfacade::Prim p;
    p.kind = PT_GEOM_MESH;
    p.mesh = mi;
    gi->prims.push_back(p);
    return gi;
  }
  void rotate(float angle, AXIS axis) {
    arr.push_back(TransformParameter(Rotate_Transform, angle, axis, make_float3(0.f), make_float3(0.f)));
  }
  void scale(float3 s) { arr.push_back(TransformParameter(Scale_Transform, 0.f, X_AXIS, s, make_float3(0.f))); }
  void translate(float3 pos) {
    arr.push_back(TransformParameter(Translate_Transform, 0.f, X_AXIS, make_float3(0.f), pos));
  }
  // mesh.hpp:237-242
  void addTo(Group& d_world, Context& g_context) {
    std::reverse(arr.begin(), arr.end());
    GeometryInstance gi = getGeometryInstance(g_context);
    addAndTransform(gi, d_world, g_context, arr);
  }

 private:
  bool RTX_MODE;
  BRDF* givenMaterial;
  std::string fileName, assetsFolder;
  std::vector<TransformParameter> arr;
  std::vector<float3> mem_v;
  std::vector<uint3> mem_t;
  std::vector<float3> mem_n;
  std::vector<float2> mem_uv;
  bool in_memory = false;
};

// mesh.hpp:273-345
class Mesh_List {
 public:
  Mesh_List() {}
  int push(Mesh* mesh) {
    int index = (int)list.size();
    list.push_back(mesh);
    return index;
  }
  void rotate(float angle, AXIS axis) {
    arr.push_back(TransformParameter(Rotate_Transform, angle, axis, make_float3(0.f), make_float3(0.f)));
  }
  void scale(float3 s) { arr.push_back(TransformParameter(Scale_Transform, 0.f, X_AXIS, s, make_float3(0.f))); }
  void translate(float3 pos) {
    arr.push_back(TransformParameter(Translate_Transform, 0.f, X_AXIS, make_float3(0.f), pos));
  }
  GeometryGroup getGroup(Context& g_context) {
    GeometryGroup gg = g_context->createGeometryGroup();
    gg->setAcceleration(g_context->createAcceleration("Trbvh"));
    for (size_t i = 0; i < list.size(); i++) gg->addChild(list[i]->getGeometryInstance(g_context));
    return gg;
  }
  void addListTo(Group& d_world, Context& g_context) {
    GeometryGroup gg = getGroup(g_context);
    addAndTransform(gg, d_world, g_context, arr);
  }
  void addElementsTo(Group& d_world, Context& g_context) {
    for (int i = 0; i < (int)list.size(); i++) list[i]->addTo(d_world, g_context);
  }

 private:
  std::vector<Mesh*> list;
  std::vector<TransformParameter> arr;
};
