// transforms.hpp -- instance transforms of the host scene API
// (host_includes/transforms.hpp:9-419).  Same node structure: each
// elementary transform wraps its child (GeometryInstance, GeometryGroup or
// Transform); applyTransform takes the LAST parameter as the innermost
// transform, so the composed object-to-world matrix is
// M(params[0]) * M(params[1]) * ... * M(params[last]).
#pragma once

#include "host_common.hpp"
#include "materials.hpp"

inline void check_if_null(GeometryInstance gi) {
  if (!gi) throw "Assigned GeometryInstance is NULL";
}
inline void check_if_null(GeometryGroup gg) {
  if (!gg) throw "Assigned GeometryGroup is NULL";
}
inline void check_if_null(Transform t) {
  if (!t) throw "Assigned Transform is NULL";
}

// transforms.hpp:29-44
typedef enum { Rotate_Transform, Translate_Transform, Scale_Transform } Transform_Type;
struct TransformParameter {
  TransformParameter(Transform_Type t, float a, AXIS ax, float3 s, float3 p)
      : type(t), angle(a), axis(ax), scale(s), pos(p) {}
  Transform_Type type;
  float angle;
  AXIS axis;
  float3 scale;
  float3 pos;
};

inline float3 axis_vector(AXIS ax) {
  switch (ax) {
    case X_AXIS: return make_float3(1.f, 0.f, 0.f);
    case Y_AXIS: return make_float3(0.f, 1.f, 0.f);
    case Z_AXIS: return make_float3(0.f, 0.f, 1.f);
  }
  throw "Invalid rotation axis";
}

inline Transform wrap(const ptfacade::Mat4& m, Context& g_context) {
  Transform t = g_context->createTransform();
  t->setMatrix(m);
  return t;
}
inline Transform wrap_gi(GeometryInstance gi, const ptfacade::Mat4& m, Context& g_context) {
  check_if_null(gi);
  GeometryGroup group = g_context->createGeometryGroup();
  group->setAcceleration(g_context->createAcceleration("Trbvh"));
  group->addChild(gi);
  Transform t = wrap(m, g_context);
  t->setChild(group);
  return t;
}

// transforms.hpp:51-98
inline Transform translate(GeometryInstance gi, float3& xyz, Context& c) { return wrap_gi(gi, ptfacade::Mat4::translate(xyz), c); }
inline Transform translate(GeometryGroup gg, float3& xyz, Context& c) {
  check_if_null(gg);
  Transform t = wrap(ptfacade::Mat4::translate(xyz), c);
  t->setChild(gg);
  return t;
}
inline Transform translate(Transform tr, float3& xyz, Context& c) {
  check_if_null(tr);
  Transform t = wrap(ptfacade::Mat4::translate(xyz), c);
  t->setChild(tr);
  return t;
}
// transforms.hpp:105-201 (angle in degrees)
inline ptfacade::Mat4 rotation(float angle, AXIS ax) {
  return ptfacade::Mat4::rotate((double)angle * PI_D / 180.0, axis_vector(ax));
}
inline Transform rotate(GeometryInstance gi, float angle, AXIS ax, Context& c) { return wrap_gi(gi, rotation(angle, ax), c); }
inline Transform rotate(GeometryGroup gg, float angle, AXIS ax, Context& c) {
  check_if_null(gg);
  Transform t = wrap(rotation(angle, ax), c);
  t->setChild(gg);
  return t;
}
inline Transform rotate(Transform tr, float angle, AXIS ax, Context& c) {
  check_if_null(tr);
  Transform t = wrap(rotation(angle, ax), c);
  t->setChild(tr);
  return t;
}
// transforms.hpp:208-255
inline Transform scale(GeometryInstance gi, float3& xyz, Context& c) { return wrap_gi(gi, ptfacade::Mat4::scale(xyz), c); }
inline Transform scale(GeometryGroup gg, float3& xyz, Context& c) {
  check_if_null(gg);
  Transform t = wrap(ptfacade::Mat4::scale(xyz), c);
  t->setChild(gg);
  return t;
}
inline Transform scale(Transform tr, float3& xyz, Context& c) {
  check_if_null(tr);
  Transform t = wrap(ptfacade::Mat4::scale(xyz), c);
  t->setChild(tr);
  return t;
}

// transforms.hpp:266-362
template <class Node>
inline Transform applyTransform(Node node, std::vector<TransformParameter>& params, Context& c) {
  TransformParameter p = params[params.size() - 1];
  params.pop_back();
  Transform t;
  switch (p.type) {
    case Rotate_Transform: t = rotate(node, p.angle, p.axis, c); break;
    case Translate_Transform: t = translate(node, p.pos, c); break;
    case Scale_Transform: t = scale(node, p.scale, c); break;
    default: throw "Invalid Transform operation";
  }
  if (params.size() > 0) return applyTransform(t, params, c);
  return t;
}

// transforms.hpp:365-419
inline void addAndTransform(Transform tr, Group& d_world, Context& c, std::vector<TransformParameter> params) {
  check_if_null(tr);
  Transform t = params.size() > 0 ? applyTransform(tr, params, c) : tr;
  d_world->addChild(t);
  d_world->getAcceleration()->markDirty();
  c->invalidate();
}
inline void addAndTransform(GeometryGroup gg, Group& d_world, Context& c, std::vector<TransformParameter> params) {
  check_if_null(gg);
  if (params.size() == 0) {
    d_world->addChild(gg);
    d_world->getAcceleration()->markDirty();
    c->invalidate();
  } else {
    Transform t = applyTransform(gg, params, c);
    addAndTransform(t, d_world, c, params);
  }
}
inline void addAndTransform(GeometryInstance gi, Group& d_world, Context& c, std::vector<TransformParameter> params) {
  check_if_null(gi);
  if (params.size() == 0) {
    GeometryGroup group = c->createGeometryGroup();
    group->setAcceleration(c->createAcceleration("Trbvh"));
    group->addChild(gi);
    d_world->addChild(group);
    d_world->getAcceleration()->markDirty();
    c->invalidate();
  } else {
    Transform t = applyTransform(gi, params, c);
    addAndTransform(t, d_world, c, params);
  }
}
