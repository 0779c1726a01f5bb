// hitables.hpp -- geometry builders of the host scene API
// (host_includes/hitables.hpp:20-528): Sphere, Moving_Sphere, AARect, Box,
// the volumetric sphere and box, Cylinder and Hitable_List with the reference's transform-order semantics
// (Hitable::addTo reverses its list, Hitable_List::addElementsTo and
// addListTo do not).
#pragma once

#include <algorithm>

#include "host_common.hpp"
#include "materials.hpp"
#include "programs.hpp"
#include "transforms.hpp"

// hitables.hpp:21-86
class Hitable {
 public:
  std::vector<TransformParameter> transforms;
  Hitable(BRDF* material) : material(material) {}
  virtual ~Hitable() = default;
  virtual GeometryInstance getGeometryInstance(Context& g_context) = 0;
  virtual void rotate(float angle, AXIS axis) {
    transforms.push_back(TransformParameter(Rotate_Transform, angle, axis, make_float3(0.f), make_float3(0.f)));
  }
  virtual void scale(float3 s) {
    transforms.push_back(TransformParameter(Scale_Transform, 0.f, X_AXIS, s, make_float3(0.f)));
  }
  virtual void translate(float3 pos) {
    transforms.push_back(TransformParameter(Translate_Transform, 0.f, X_AXIS, make_float3(0.f), pos));
  }
  virtual void addTo(Group& d_world, Context& g_context) {
    std::reverse(transforms.begin(), transforms.end());
    GeometryInstance gi = getGeometryInstance(g_context);
    addAndTransform(gi, d_world, g_context, transforms);
  }

 protected:
  BRDF* material;
  GeometryInstance make_instance(Context& g_context, const ptfacade::Prim& p) {
    if (!material) throw "Hitable has no material";
    GeometryInstance gi = g_context->createGeometryInstance();
    gi->setMaterialCount(1);
    gi->setMaterial(0, material->assignTo(g_context));
    gi->prims.push_back(p);
    return gi;
  }
};

// hitables.hpp:89-128
class Sphere : public Hitable {
 public:
  Sphere(const float3& c, const float r, BRDF* material) : Hitable(material), center(c), radius(r) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_SPHERES;
    p.s = pt_sphere_desc{{center.x, center.y, center.z}, radius, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float3 center;
  const float radius;
};

// hitables.hpp:211-272: a0..a1 and b0..b1 span the two axes other than
// `ax` (Y,Z for X; X,Z for Y; X,Y for Z), k lies on `ax`.
class AARect : public Hitable {
 public:
  AARect(const float a0, const float a1, const float b0, const float b1, const float k, const bool flip,
         const AXIS ax, BRDF* material)
      : Hitable(material), a0(a0), a1(a1), b0(b0), b1(b1), k(k), axis(ax), flip(flip) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_RECTS;
    p.r = pt_rect_desc{(int32_t)axis, flip ? 1 : 0, a0, a1, b0, b1, k, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float a0, a1, b0, b1, k;
  const AXIS axis;
  const bool flip;
};

// hitables.hpp:275-313
class Box : public Hitable {
 public:
  Box(const float3& p0, const float3& p1, BRDF* material) : Hitable(material), p0(p0), p1(p1) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_BOXES;
    p.b = pt_box_desc{{p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float3 p0, p1;
};

// hitables.hpp:175-209: constant-density medium inside a sphere
class Volumetric_Sphere : public Hitable {
 public:
  Volumetric_Sphere(const float3& c, const float r, const float d, BRDF* material)
      : Hitable(material), center(c), radius(r), density(d) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_VOLUMES;
    p.v = pt_volume_desc{PT_VOLUME_SPHERE, {center.x, center.y, center.z}, {radius, 0.f, 0.f}, density, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float3 center;
  const float radius;
  const float density;
};
// hitables.hpp:316-349: constant-density medium inside a box
class Volumetric_Box : public Hitable {
 public:
  Volumetric_Box(const float3& p0, const float3& p1, const float d, BRDF* material)
      : Hitable(material), p0(p0), p1(p1), density(d) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_VOLUMES;
    p.v = pt_volume_desc{PT_VOLUME_BOX, {p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, density, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float3 p0, p1;
  const float density;
};

// hitables.hpp:131-172: a sphere moving from c0 at t0 to c1 at t1 (marked
// broken in the reference at :130; built here as Shirley's moving sphere,
// with each path's time drawn in the camera's shutter interval)
class Moving_Sphere : public Hitable {
 public:
  Moving_Sphere(const float3& c0, const float3& c1, const float r, const float t0, const float t1, BRDF* material)
      : Hitable(material), center0(c0), center1(c1), radius(r), time0(t0), time1(t1) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_MOVING_SPHERES;
    p.m = pt_moving_sphere_desc{{center0.x, center0.y, center0.z}, {center1.x, center1.y, center1.z}, radius, time0,
                                time1, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float3 center0, center1;
  const float radius;
  const float time0, time1;
};
// hitables.hpp:411-454: the open lateral surface of a +Y cylinder
// (cylinder.cu's intersection fixed to its intent)
class Cylinder : public Hitable {
 public:
  Cylinder(const float3& oo, const float& ll, const float r, BRDF* material)
      : Hitable(material), O(oo), length(ll), radius(r) {}
  GeometryInstance getGeometryInstance(Context& g_context) override {
    ptfacade::Prim p;
    p.kind = PT_GEOM_CYLINDERS;
    p.c = pt_cylinder_desc{{O.x, O.y, O.z}, length, radius, -1};
    return make_instance(g_context, p);
  }

 protected:
  const float3 O;
  const float length;
  const float radius;
};

// hitables.hpp:457-528
class Hitable_List {
 public:
  Hitable_List() {}
  int push(Hitable* hit) {
    int index = (int)hitList.size();
    hitList.push_back(hit);
    return index;
  }
  void rotate(float angle, AXIS axis) {
    transforms.push_back(TransformParameter(Rotate_Transform, angle, axis, make_float3(0.f), make_float3(0.f)));
  }
  void scale(float3 s) { transforms.push_back(TransformParameter(Scale_Transform, 0.f, X_AXIS, s, make_float3(0.f))); }
  void translate(float3 pos) {
    transforms.push_back(TransformParameter(Translate_Transform, 0.f, X_AXIS, make_float3(0.f), pos));
  }
  GeometryGroup getGroup(Context& g_context) {
    GeometryGroup gg = g_context->createGeometryGroup();
    gg->setAcceleration(g_context->createAcceleration("Trbvh"));
    for (size_t i = 0; i < hitList.size(); i++) gg->addChild(hitList[i]->getGeometryInstance(g_context));
    return gg;
  }
  void addListTo(Group& d_world, Context& g_context) {
    GeometryGroup gg = getGroup(g_context);
    addAndTransform(gg, d_world, g_context, transforms);
  }
  void addElementsTo(Group& d_world, Context& g_context) {
    for (int i = 0; i < (int)hitList.size(); i++) {
      GeometryInstance gi = hitList[i]->getGeometryInstance(g_context);
      addAndTransform(gi, d_world, g_context, hitList[i]->transforms);
    }
  }

 protected:
  std::vector<Hitable*> hitList;
  std::vector<TransformParameter> transforms;
};
