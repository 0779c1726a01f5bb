// optix_facade.hpp -- the host object model the reference's scene code is
// written against (Context, Group, GeometryGroup, GeometryInstance,
// Transform, Program, Material, Buffer, Variable and the float2/float3
// vector helpers), re-targeted to the C ABI of libpt_hip.so.
//
// The reference gets these types from the OptiX 6 C++ wrapper and the CUDA
// vector headers (host_includes/host_common.hpp:10 -> programs/vec.hpp ->
// programs/common.hpp:6-8).  Here they are plain host C++: the scene graph is
// recorded, flattened into a pt_scene_desc at the first launch and handed to
// pt_set_scene / pt_build_accel; launch(0, W, H) becomes pt_render.
//
// Errors keep the reference's behaviour: invalid use throws const char* or
// std::string, as host_common.hpp:53 and transforms.hpp:15 do.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/pt_api.h"

// ------------------------------------------------------------ vector types
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct int3 { int x, y, z; };
struct uint2 { unsigned int x, y; };
struct uint3 { unsigned int x, y, z; };
struct uchar1 { unsigned char x; };
struct uchar4 { unsigned char x, y, z, w; };

inline float2 make_float2(float x, float y) { return {x, y}; }
inline float2 make_float2(float s) { return {s, s}; }
inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float3 make_float3(float s) { return {s, s, s}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline float4 make_float4(float s) { return {s, s, s, s}; }
inline int3 make_int3(int x, int y, int z) { return {x, y, z}; }
inline uint3 make_uint3(unsigned x, unsigned y, unsigned z) { return {x, y, z}; }
inline uchar4 make_uchar4(unsigned char x, unsigned char y, unsigned char z, unsigned char w) { return {x, y, z, w}; }

inline float3 operator+(float3 a, float3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator-(float3 a, float3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator-(float3 a) { return {-a.x, -a.y, -a.z}; }
inline float3 operator*(float3 a, float3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float3 operator*(float3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float3 operator*(float s, float3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float3 operator/(float3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float3 operator/(float3 a, float3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline float3 operator+(float3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline float3 operator-(float3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
inline float3& operator+=(float3& a, float3 b) { return a = a + b; }
inline float3& operator-=(float3& a, float3 b) { return a = a - b; }
inline float3& operator*=(float3& a, float3 b) { return a = a * b; }
inline float3& operator*=(float3& a, float s) { return a = a * s; }
inline float3& operator/=(float3& a, float s) { return a = a / s; }
inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float3 cross(float3 a, float3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(float3 a) { return std::sqrt(dot(a, a)); }
inline float3 normalize(float3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline float clamp(float x, float lo, float hi) { return std::fmax(lo, std::fmin(x, hi)); }
inline float3 clamp(float3 v, float lo, float hi) { return {clamp(v.x, lo, hi), clamp(v.y, lo, hi), clamp(v.z, lo, hi)}; }
inline float3 sqrt(float3 v) { return {std::sqrt(v.x), std::sqrt(v.y), std::sqrt(v.z)}; }

// programs/vec.hpp helpers the host code uses
inline float ffmin(float a, float b) { return a < b ? a : b; }
inline float ffmax(float a, float b) { return a > b ? a : b; }
inline float3 unit_vector(const float3& v) { return v / length(v); }
inline float squared_length(const float3& a) { return length(a) * length(a); }

// programs/common.hpp:17-29
#ifndef PI_F
#define PI_F 3.141592654f
#endif
#ifndef PI_D
#define PI_D 3.14159265358979323846264338327950288
#endif
typedef enum { X_AXIS, Y_AXIS, Z_AXIS } AXIS;
inline float Radians(float deg) { return (PI_F / 180.f) * deg; }
inline float Degrees(float rad) { return (180.f / PI_F) * rad; }

// OptiX enums the host code names
enum RTbuffertype { RT_BUFFER_INPUT = 1, RT_BUFFER_OUTPUT = 2, RT_BUFFER_INPUT_OUTPUT = 3 };
enum RTformat {
  RT_FORMAT_FLOAT = 1, RT_FORMAT_FLOAT2, RT_FORMAT_FLOAT3, RT_FORMAT_FLOAT4, RT_FORMAT_INT, RT_FORMAT_UNSIGNED_INT3,
  RT_FORMAT_UNSIGNED_BYTE4, RT_FORMAT_PROGRAM_ID
};

namespace ptfacade {

// ------------------------------------------------------------ 4x4 matrices
// Row-major affine transform kept in double; rounded to float at flatten.
struct Mat4 {
  double m[16];
  static Mat4 identity() {
    Mat4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
    return r;
  }
  static Mat4 translate(float3 t) {
    Mat4 r = identity();
    r.m[3] = t.x; r.m[7] = t.y; r.m[11] = t.z;
    return r;
  }
  static Mat4 scale(float3 s) {
    Mat4 r = identity();
    r.m[0] = s.x; r.m[5] = s.y; r.m[10] = s.z;
    return r;
  }
  // right-handed rotation by `radians` about a unit axis (Rodrigues)
  static Mat4 rotate(double radians, float3 axis) {
    double x = axis.x, y = axis.y, z = axis.z;
    double n = std::sqrt(x * x + y * y + z * z);
    x /= n; y /= n; z /= n;
    double c = std::cos(radians), s = std::sin(radians), t = 1.0 - c;
    Mat4 r = identity();
    r.m[0] = c + t * x * x;     r.m[1] = t * x * y - s * z; r.m[2] = t * x * z + s * y;
    r.m[4] = t * x * y + s * z; r.m[5] = c + t * y * y;     r.m[6] = t * y * z - s * x;
    r.m[8] = t * x * z - s * y; r.m[9] = t * y * z + s * x; r.m[10] = c + t * z * z;
    return r;
  }
  Mat4 operator*(const Mat4& b) const {
    Mat4 r{};
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        double acc = 0;
        for (int k = 0; k < 4; k++) acc += m[4 * i + k] * b.m[4 * k + j];
        r.m[4 * i + j] = acc;
      }
    return r;
  }
  bool is_identity() const {
    Mat4 I = identity();
    for (int i = 0; i < 16; i++)
      if (m[i] != I.m[i]) return false;
    return true;
  }
  // general inverse of the affine part (3x3 inverse + translation)
  Mat4 inverse() const {
    const double* a = m;
    double c00 = a[5] * a[10] - a[6] * a[9], c01 = a[6] * a[8] - a[4] * a[10], c02 = a[4] * a[9] - a[5] * a[8];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (det == 0.0) throw "Transform matrix is singular";
    double id = 1.0 / det;
    Mat4 r = identity();
    r.m[0] = c00 * id;
    r.m[1] = (a[2] * a[9] - a[1] * a[10]) * id;
    r.m[2] = (a[1] * a[6] - a[2] * a[5]) * id;
    r.m[4] = c01 * id;
    r.m[5] = (a[0] * a[10] - a[2] * a[8]) * id;
    r.m[6] = (a[2] * a[4] - a[0] * a[6]) * id;
    r.m[8] = c02 * id;
    r.m[9] = (a[1] * a[8] - a[0] * a[9]) * id;
    r.m[10] = (a[0] * a[5] - a[1] * a[4]) * id;
    for (int i = 0; i < 3; i++)
      r.m[4 * i + 3] = -(r.m[4 * i] * a[3] + r.m[4 * i + 1] * a[7] + r.m[4 * i + 2] * a[11]);
    return r;
  }
};

struct Recorder;
struct ContextObj;

// ------------------------------------------------------------------ objects
struct ObjBase {
  virtual ~ObjBase() = default;
};

// A Program is an opaque value.  The recorder reads the payload of the kinds
// it understands (textures, light sample/pdf, miss, raygen).
struct ProgramObj : ObjBase {
  enum Kind { OTHER, TEXTURE, LIGHT_SAMPLE, LIGHT_PDF, RAYGEN, MISS, EXCEPTION } kind = OTHER;
  int texture = -1;
  pt_light_desc light{};
  std::string name;
};
struct MaterialObj : ObjBase {
  int index = -1;
};
struct AccelerationObj : ObjBase {
  std::string builder;
  bool dirty = true;
  void markDirty() { dirty = true; }
  void setProperty(const std::string&, const std::string&) {}
};

template <class T>
struct Handle {
  std::shared_ptr<T> p;
  Handle() = default;
  Handle(std::shared_ptr<T> q) : p(std::move(q)) {}
  T* operator->() const {
    if (!p) throw "Dereferenced a NULL scene object";
    return p.get();
  }
  T* get() const { return p.get(); }
  explicit operator bool() const { return (bool)p; }
  bool operator!() const { return !p; }
};

using Program = Handle<ProgramObj>;
using Material = Handle<MaterialObj>;
using Acceleration = Handle<AccelerationObj>;

// One recorded primitive (sphere, rect, box or mesh reference)
struct Prim {
  int kind = PT_GEOM_SPHERES;
  pt_sphere_desc s{};
  pt_rect_desc r{};
  pt_box_desc b{};
  pt_volume_desc v{};
  pt_cylinder_desc c{};
  pt_moving_sphere_desc m{};
  int mesh = -1;
};

struct GeometryInstanceObj : ObjBase {
  std::vector<Prim> prims;  // one primitive (or one mesh) per instance
  int material = -1;
  void setMaterialCount(int) {}
  void setMaterial(int, Material m) { material = m ? m->index : -1; }
};
using GeometryInstance = Handle<GeometryInstanceObj>;

struct GeometryGroupObj : ObjBase {
  std::vector<GeometryInstance> children;
  Acceleration accel;
  void setAcceleration(Acceleration a) { accel = a; }
  Acceleration getAcceleration() { return accel; }
  void addChild(GeometryInstance gi) { children.push_back(gi); }
  unsigned getChildCount() const { return (unsigned)children.size(); }
};
using GeometryGroup = Handle<GeometryGroupObj>;

struct TransformObj;
using Transform = Handle<TransformObj>;
struct TransformObj : ObjBase {
  Mat4 matrix = Mat4::identity();
  GeometryInstance child_gi;
  GeometryGroup child_gg;
  Transform child_tr;
  void setChild(GeometryGroup g) { child_gg = g; child_gi = {}; child_tr = {}; }
  void setChild(Transform t) { child_tr = t; child_gi = {}; child_gg = {}; }
  // transform.hpp:65 -- row-major 4x4 (transposed when `transpose`)
  void setMatrix(bool transpose, const float* m, const float* /*inverse*/) {
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) matrix.m[4 * i + j] = transpose ? m[4 * j + i] : m[4 * i + j];
  }
  void setMatrix(const Mat4& m) { matrix = m; }
};

struct GroupObj : ObjBase {
  std::vector<GeometryGroup> groups;
  std::vector<Transform> transforms;
  Acceleration accel;
  void setAcceleration(Acceleration a) { accel = a; }
  Acceleration getAcceleration() {
    if (!accel) accel = Acceleration(std::make_shared<AccelerationObj>());
    return accel;
  }
  void addChild(GeometryGroup g) { groups.push_back(g); }
  void addChild(Transform t) { transforms.push_back(t); }
  unsigned getChildCount() const { return (unsigned)(groups.size() + transforms.size()); }
};
using Group = Handle<GroupObj>;

// Buffers: the two output buffers of main.cpp:92-97 map to the library's
// accumulation and display images; input buffers hold host data.
struct BufferObj : ObjBase {
  ContextObj* ctx = nullptr;
  int type = RT_BUFFER_INPUT, format = RT_FORMAT_FLOAT4;
  size_t w = 0, h = 1;
  int role = 0;  // 0 input, 1 accumulation (float4), 2 display (uchar4)
  std::vector<unsigned char> host;
  void setFormat(int f) { format = f; }
  void setSize(size_t W) { w = W; h = 1; }
  void setSize(size_t W, size_t H) { w = W; h = H; }
  void getSize(size_t& W, size_t& H) const { W = w; H = h; }
  void* map();
  void unmap() {}
};
using Buffer = Handle<BufferObj>;

struct VariableObj {
  ContextObj* ctx;
  std::string name;
  void set(Group g);
  void set(Buffer b);
  void set(Program p);
  void setBuffer(Buffer b) { set(b); }
  void setProgramId(Program p) { set(p); }
  void setInt(int v);
  void setUint(unsigned v) { setInt((int)v); }
  void setFloat(float x);
  void setFloat(float x, float y);
  void setFloat(float x, float y, float z);
  void set3fv(const float* v) { setFloat(v[0], v[1], v[2]); }
  int getInt() const;
};
struct Variable {
  std::shared_ptr<VariableObj> p;
  VariableObj* operator->() const { return p.get(); }
};

// --------------------------------------------------------------- recorder
struct MeshData {
  std::vector<float> v, n, uv;
  std::vector<uint32_t> idx;
  std::vector<int32_t> fmat;
  std::vector<int32_t> fidx;  // per-face vector-texture index (the material buffer)
  int32_t material = 0;
};

struct FlatScene {
  std::vector<pt_sphere_desc> spheres;
  std::vector<pt_rect_desc> rects;
  std::vector<pt_box_desc> boxes;
  std::vector<pt_volume_desc> volumes;
  std::vector<pt_cylinder_desc> cylinders;
  std::vector<pt_moving_sphere_desc> moving;
  std::vector<pt_mesh_desc> meshes;
  std::vector<pt_instance_desc> instances;
  std::vector<pt_material_desc> materials;
  std::vector<pt_texture_desc> textures;
  std::vector<pt_light_desc> lights;
  pt_scene_desc desc{};
};

struct Recorder {
  std::vector<pt_texture_desc> textures;
  std::vector<std::shared_ptr<std::vector<float>>> texels;
  std::map<const void*, int> tex_of;
  std::vector<pt_material_desc> materials;
  std::map<const void*, int> mat_of;
  std::vector<MeshData> meshes;
  std::vector<pt_light_desc> lights;
  pt_miss_desc miss{PT_MISS_CONSTANT, {0, 0, 0}, {0, 0, 0}, -1, 1};
  Group world;
  std::map<std::string, std::vector<float>> fvars;
  std::map<std::string, int> ivars;

  int add_texture(const void* key, const pt_texture_desc& t, std::shared_ptr<std::vector<float>> data) {
    auto it = tex_of.find(key);
    if (key && it != tex_of.end()) return it->second;
    int idx = (int)textures.size();
    textures.push_back(t);
    texels.push_back(std::move(data));
    if (key) tex_of[key] = idx;
    return idx;
  }
  int add_material(const void* key, const pt_material_desc& m) {
    auto it = mat_of.find(key);
    if (key && it != mat_of.end()) return it->second;
    int idx = (int)materials.size();
    materials.push_back(m);
    if (key) mat_of[key] = idx;
    return idx;
  }

  // Flatten the world group.  Untransformed spheres, rects, boxes and volumes share
  // one identity instance per kind; every transformed node and every mesh
  // gets its own instance.
  void flatten(FlatScene& out) const {
    if (!world) throw "Scene variable 'world' was not set";
    out = FlatScene{};
    struct Item { Prim p; int mat; bool xf; Mat4 m; int group; };
    std::vector<Item> items;
    int next_group = 0;
    std::function<void(const GeometryInstance&, const Mat4&, bool, int)> visit_gi =
        [&](const GeometryInstance& gi, const Mat4& M, bool xf, int grp) {
          for (const Prim& p : gi->prims) items.push_back({p, gi->material, xf, M, grp});
        };
    std::function<void(const Transform&, const Mat4&)> visit_tr = [&](const Transform& t, const Mat4& M) {
      Mat4 C = M * t->matrix;
      if (t->child_tr) visit_tr(t->child_tr, C);
      else if (t->child_gg) {
        int g = next_group++;
        for (auto& gi : t->child_gg->children) visit_gi(gi, C, true, g);
      } else if (t->child_gi) {
        visit_gi(t->child_gi, C, true, next_group++);
      } else {
        throw "Assigned Transform has no child";
      }
    };
    for (auto& gg : world->groups) {
      int g = next_group++;
      for (auto& gi : gg->children) visit_gi(gi, Mat4::identity(), false, g);
    }
    for (auto& t : world->transforms) visit_tr(t, Mat4::identity());
    for (auto& it : items)
      if (it.xf && it.m.is_identity()) it.xf = false;
    auto to_f12 = [](const Mat4& m, float* o) {
      for (int i = 0; i < 12; i++) o[i] = (float)m.m[i];
    };
    auto add_instance = [&](int kind, uint32_t first, uint32_t count, bool xf, const Mat4& m) {
      pt_instance_desc in{};
      in.kind = kind;
      in.first = first;
      in.count = count;
      in.has_transform = xf ? 1 : 0;
      Mat4 inv = xf ? m.inverse() : Mat4::identity();
      to_f12(xf ? m : Mat4::identity(), in.obj_to_world);
      to_f12(inv, in.world_to_obj);
      out.instances.push_back(in);
    };
    // identity instances of the analytic kinds, then transformed groups
    static const int analytic[6] = {PT_GEOM_SPHERES, PT_GEOM_RECTS,     PT_GEOM_BOXES,
                                    PT_GEOM_VOLUMES, PT_GEOM_CYLINDERS, PT_GEOM_MOVING_SPHERES};
    for (int kind : analytic) {
      uint32_t first = 0, count = 0;
      for (auto& it : items) {
        if (it.p.kind != kind || it.xf) continue;
        uint32_t idx = push_prim(out, it.p, it.mat);
        if (count == 0) first = idx;
        count++;
      }
      if (count) add_instance(kind, first, count, false, Mat4::identity());
    }
    for (int g = 0; g < next_group; g++) {
      for (int kind : analytic) {
        uint32_t first = 0, count = 0;
        Mat4 m = Mat4::identity();
        for (auto& it : items) {
          if (it.group != g || it.p.kind != kind || !it.xf) continue;
          uint32_t idx = push_prim(out, it.p, it.mat);
          if (count == 0) first = idx;
          m = it.m;
          count++;
        }
        if (count) add_instance(kind, first, count, true, m);
      }
    }
    for (auto& it : items) {
      if (it.p.kind != PT_GEOM_MESH) continue;
      add_instance(PT_GEOM_MESH, (uint32_t)it.p.mesh, 0, it.xf, it.m);
    }
    for (auto& m : meshes) {
      pt_mesh_desc d{};
      d.num_vertices = (uint32_t)(m.v.size() / 3);
      d.vertices = m.v.data();
      d.normals = m.n.empty() ? nullptr : m.n.data();
      d.uvs = m.uv.empty() ? nullptr : m.uv.data();
      d.num_triangles = (uint32_t)(m.idx.size() / 3);
      d.indices = m.idx.data();
      d.face_materials = m.fmat.empty() ? nullptr : m.fmat.data();
      d.face_indices = m.fidx.empty() ? nullptr : m.fidx.data();
      d.material = m.material;
      out.meshes.push_back(d);
    }
    out.materials = materials;
    out.textures = textures;
    for (size_t i = 0; i < out.textures.size(); i++)
      out.textures[i].texels = texels[i] ? texels[i]->data() : nullptr;
    out.lights = lights;
    pt_scene_desc& d = out.desc;
    d.num_spheres = (uint32_t)out.spheres.size();
    d.spheres = out.spheres.data();
    d.num_rects = (uint32_t)out.rects.size();
    d.rects = out.rects.data();
    d.num_boxes = (uint32_t)out.boxes.size();
    d.boxes = out.boxes.data();
    d.num_volumes = (uint32_t)out.volumes.size();
    d.volumes = out.volumes.data();
    d.num_cylinders = (uint32_t)out.cylinders.size();
    d.cylinders = out.cylinders.data();
    d.num_moving_spheres = (uint32_t)out.moving.size();
    d.moving_spheres = out.moving.data();
    d.num_meshes = (uint32_t)out.meshes.size();
    d.meshes = out.meshes.data();
    d.num_instances = (uint32_t)out.instances.size();
    d.instances = out.instances.data();
    d.num_materials = (uint32_t)out.materials.size();
    d.materials = out.materials.data();
    d.num_textures = (uint32_t)out.textures.size();
    d.textures = out.textures.data();
    d.num_lights = (uint32_t)out.lights.size();
    d.lights = out.lights.data();
    d.miss = miss;
  }

  static uint32_t push_prim(FlatScene& out, const Prim& p, int mat) {
    if (mat < 0) throw "GeometryInstance has no material";
    if (p.kind == PT_GEOM_SPHERES) {
      pt_sphere_desc s = p.s;
      s.material = mat;
      out.spheres.push_back(s);
      return (uint32_t)out.spheres.size() - 1;
    }
    if (p.kind == PT_GEOM_RECTS) {
      pt_rect_desc r = p.r;
      r.material = mat;
      out.rects.push_back(r);
      return (uint32_t)out.rects.size() - 1;
    }
    if (p.kind == PT_GEOM_CYLINDERS) {
      pt_cylinder_desc c = p.c;
      c.material = mat;
      out.cylinders.push_back(c);
      return (uint32_t)out.cylinders.size() - 1;
    }
    if (p.kind == PT_GEOM_MOVING_SPHERES) {
      pt_moving_sphere_desc m = p.m;
      m.material = mat;
      out.moving.push_back(m);
      return (uint32_t)out.moving.size() - 1;
    }
    if (p.kind == PT_GEOM_VOLUMES) {
      pt_volume_desc v = p.v;
      v.material = mat;
      out.volumes.push_back(v);
      return (uint32_t)out.volumes.size() - 1;
    }
    pt_box_desc b = p.b;
    b.material = mat;
    out.boxes.push_back(b);
    return (uint32_t)out.boxes.size() - 1;
  }

  // camera variables of Camera::set (camera.hpp:40-51)
  pt_camera camera() const {
    pt_camera c{};
    auto get3 = [&](const char* n, float* o) {
      auto it = fvars.find(n);
      if (it == fvars.end() || it->second.size() < 3) throw std::string("Camera variable ") + n + " not set";
      std::copy(it->second.begin(), it->second.begin() + 3, o);
    };
    auto get1 = [&](const char* n) {
      auto it = fvars.find(n);
      return it == fvars.end() || it->second.empty() ? 0.f : it->second[0];
    };
    get3("camera_lower_left_corner", c.lower_left);
    get3("camera_horizontal", c.horizontal);
    get3("camera_vertical", c.vertical);
    get3("camera_origin", c.origin);
    get3("camera_u", c.u);
    get3("camera_v", c.v);
    auto w = fvars.find("camera_w");
    if (w != fvars.end() && w->second.size() >= 3) std::copy(w->second.begin(), w->second.begin() + 3, c.w);
    c.lens_radius = get1("camera_lens_radius");
    c.time0 = get1("time0");
    c.time1 = get1("time1");
    return c;
  }
  int ivar(const char* n, int def) const {
    auto it = ivars.find(n);
    return it == ivars.end() ? def : it->second;
  }
};

struct ContextObj : ObjBase {
  Recorder rec;
  pt_context* pt = nullptr;
  bool uploaded = false;
  FlatScene flat;
  int ray_types = 2, max_trace_depth = 5;
  Buffer acc, disp;
  std::vector<float4> acc_host;
  std::vector<uchar4> disp_host;
  uint32_t W = 0, H = 0;

  ~ContextObj() {
    if (pt) pt_destroy(pt);
  }
  Group createGroup() { return Group(std::make_shared<GroupObj>()); }
  GeometryGroup createGeometryGroup() { return GeometryGroup(std::make_shared<GeometryGroupObj>()); }
  GeometryInstance createGeometryInstance() { return GeometryInstance(std::make_shared<GeometryInstanceObj>()); }
  Transform createTransform() { return Transform(std::make_shared<TransformObj>()); }
  Acceleration createAcceleration(const std::string& builder) {
    auto a = std::make_shared<AccelerationObj>();
    a->builder = builder;
    return Acceleration(a);
  }
  Material createMaterial() { return Material(std::make_shared<MaterialObj>()); }
  Buffer createBuffer(int type) {
    auto b = std::make_shared<BufferObj>();
    b->ctx = this;
    b->type = type;
    return Buffer(b);
  }
  Buffer createBuffer(int type, int format, size_t w) {
    Buffer b = createBuffer(type);
    b->setFormat(format);
    b->setSize(w);
    return b;
  }
  Buffer createBuffer(int type, int format, size_t w, size_t h) {
    Buffer b = createBuffer(type);
    b->setFormat(format);
    b->setSize(w, h);
    return b;
  }
  void setRayTypeCount(int n) { ray_types = n; }
  void setMaxTraceDepth(int n) { max_trace_depth = n; }
  void setEntryPointCount(int) {}
  void setRayGenerationProgram(int, Program) {}
  void setMissProgram(int, Program) {}
  void setExceptionProgram(int, Program) {}
  void setPrintEnabled(bool) {}
  void validate() {
    if (!rec.world) throw "Scene variable 'world' was not set";
  }
  void invalidate() { uploaded = false; }

  void ensure_context() {
    if (pt) return;
    pt_status s = pt_create(0, &pt);
    if (s != PT_OK) throw std::string("pt_create failed: ") + pt_last_error(nullptr);
  }
  void check(pt_status s) {
    if (s != PT_OK) throw std::string("path tracer error: ") + pt_last_error(pt);
  }
  // Context::launch (main.cpp:26-39): the first launch flattens and uploads
  // the scene and builds the acceleration structures; a non-empty launch
  // renders frame ctx["frame"] and accumulates it.
  void launch(int /*entry*/, int w, int h) {
    validate();
    ensure_context();
    if (!uploaded) {
      rec.flatten(flat);
      check(pt_set_scene(pt, &flat.desc));
      check(pt_build_accel(pt, nullptr));
      uploaded = true;
    }
    if (w <= 0 || h <= 0) return;
    pt_camera cam = rec.camera();
    check(pt_set_camera(pt, &cam));
    pt_render_params rp{};
    rp.width = (uint32_t)w;
    rp.height = (uint32_t)h;
    rp.max_depth = (uint32_t)rec.ivar("maxDepth", 50);
    rp.russian = rec.ivar("russian", 1);
    rp.rr_min_depth = (uint32_t)rec.ivar("rr_min_depth", 11);
    rp.flags = PT_FLAG_MATERIAL_SORT;
    int frame = rec.ivar("frame", 0);
    check(pt_render(pt, &rp, (uint32_t)frame, 1, nullptr));
    W = (uint32_t)w;
    H = (uint32_t)h;
  }
  void launch(int entry, int w) { launch(entry, w, 1); }
  // accBuffer->map(): float4 (sum rgb, frames) rows top to bottom
  void* map_accum() {
    check(pt_sync(pt));
    std::vector<double> a((size_t)W * H * 3);
    uint32_t frames = 0;
    check(pt_read_accum(pt, a.data(), &frames));
    acc_host.resize((size_t)W * H);
    for (size_t i = 0; i < acc_host.size(); i++)
      acc_host[i] = {(float)a[3 * i], (float)a[3 * i + 1], (float)a[3 * i + 2], (float)frames};
    return acc_host.data();
  }
  // pt_save_image of the accumulation (Save_PNG / Save_HDR)
  bool save_image(const char* path, int32_t format) {
    if (!pt) return false;
    return pt_sync(pt) == PT_OK && pt_save_image(pt, path, format) == PT_OK;
  }
  void* map_display() {
    disp_host.resize((size_t)W * H);
    check(pt_read_display(pt, reinterpret_cast<uint8_t*>(disp_host.data())));
    return disp_host.data();
  }
};

inline void* BufferObj::map() {
  if (role == 1) return ctx->map_accum();
  if (role == 2) return ctx->map_display();
  return host.data();
}

struct Context : Handle<ContextObj> {
  Context() = default;
  Context(std::shared_ptr<ContextObj> q) : Handle<ContextObj>(std::move(q)) {}
  static Context create() { return Context(std::make_shared<ContextObj>()); }
  Variable operator[](const std::string& name) {
    Variable v;
    v.p = std::make_shared<VariableObj>(VariableObj{get(), name});
    return v;
  }
};

inline void VariableObj::set(Group g) {
  if (name == "world") ctx->rec.world = g;
  ctx->invalidate();
}
inline void VariableObj::set(Buffer b) {
  if (name == "acc_buffer") { b->role = 1; ctx->acc = b; }
  else if (name == "display_buffer") { b->role = 2; ctx->disp = b; }
}
inline void VariableObj::set(Program) {}
inline void VariableObj::setInt(int v) { ctx->rec.ivars[name] = v; }
inline void VariableObj::setFloat(float x) { ctx->rec.fvars[name] = {x}; }
inline void VariableObj::setFloat(float x, float y) { ctx->rec.fvars[name] = {x, y}; }
inline void VariableObj::setFloat(float x, float y, float z) { ctx->rec.fvars[name] = {x, y, z}; }
inline int VariableObj::getInt() const { return ctx->rec.ivar(name.c_str(), 0); }

}  // namespace ptfacade

using ptfacade::Acceleration;
using ptfacade::Buffer;
using ptfacade::Context;
using ptfacade::GeometryGroup;
using ptfacade::GeometryInstance;
using ptfacade::Group;
using ptfacade::Material;
using ptfacade::Program;
using ptfacade::Transform;
using ptfacade::Variable;
