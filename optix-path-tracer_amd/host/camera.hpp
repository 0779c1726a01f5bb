// camera.hpp -- thin-lens camera of the host scene API
// (host_includes/camera.hpp:13-62).
#pragma once

#include "host_common.hpp"

class Camera {
 public:
  Camera() {}
  // camera.hpp:17-38
  Camera(const float3& lookfrom, const float3& lookat, const float3& vup, float vfov, float aspect, float aperture,
         float focus_dist, float t0, float t1) {
    lens_radius = aperture / 2.f;
    time0 = t0;
    time1 = t1;
    float theta = vfov * PI_F / 180.f;
    float half_height = std::tan(theta / 2.f);
    float half_width = aspect * half_height;
    origin = lookfrom;
    w = unit_vector(lookfrom - lookat);
    u = unit_vector(cross(vup, w));
    v = cross(w, u);
    lower_left_corner = origin - half_width * focus_dist * u - half_height * focus_dist * v - focus_dist * w;
    horizontal = 2.f * half_width * focus_dist * u;
    vertical = 2.f * half_height * focus_dist * v;
  }
  // camera.hpp:40-51
  void set(Context& g_context) {
    g_context["camera_lower_left_corner"]->set3fv(&lower_left_corner.x);
    g_context["camera_horizontal"]->set3fv(&horizontal.x);
    g_context["camera_vertical"]->set3fv(&vertical.x);
    g_context["camera_origin"]->set3fv(&origin.x);
    g_context["camera_u"]->set3fv(&u.x);
    g_context["camera_v"]->set3fv(&v.x);
    g_context["camera_w"]->set3fv(&w.x);
    g_context["time0"]->setFloat(time0);
    g_context["time1"]->setFloat(time1);
    g_context["camera_lens_radius"]->setFloat(lens_radius);
  }

 private:
  float3 origin;
  float3 lower_left_corner;
  float3 horizontal;
  float3 vertical;
  float3 u, v, w;
  float time0, time1;
  float lens_radius;
};
