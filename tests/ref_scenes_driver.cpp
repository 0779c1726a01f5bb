// Runs one scene function of a host scene file written against the
// reference's host API (host_includes/scenes.hpp, compiled unchanged against
// this package's host/ headers) and prints the flattened scene as JSON: what
// the facade would hand to pt_set_scene and pt_set_camera at the first
// Context::launch.  No GPU is touched.
//
//   ref_scenes_driver <InOneWeekend|Cornell|Final_Next_Week> [W H]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "scenes.hpp"

static void floats(const char* name, const float* v, int n, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (int i = 0; i < n; i++) std::printf("%s%.9g", i ? ", " : "", (double)v[i]);
  std::printf("]%s", comma ? ", " : "");
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string which = argv[1];
  App_State app;
  if (argc >= 4) {
    app.W = std::atoi(argv[2]);
    app.H = std::atoi(argv[3]);
  }
  try {
    if (which == "InOneWeekend") InOneWeekend(app);
    else if (which == "Cornell") Cornell(app);
    else if (which == "Final_Next_Week") Final_Next_Week(app);
    else return 2;
  } catch (const char* e) {
    std::fprintf(stderr, "scene error: %s\n", e);
    return 1;
  } catch (const std::string& e) {
    std::fprintf(stderr, "scene error: %s\n", e.c_str());
    return 1;
  }
  ptfacade::FlatScene flat;
  app.context->rec.flatten(flat);
  const pt_scene_desc& d = flat.desc;
  const pt_camera cam = app.context->rec.camera();
  std::printf("{\"spheres\": [");
  for (uint32_t i = 0; i < d.num_spheres; i++) {
    std::printf("%s{", i ? ", " : "");
    floats("center", d.spheres[i].center, 3);
    std::printf("\"radius\": %.9g, \"material\": %d}", (double)d.spheres[i].radius, d.spheres[i].material);
  }
  std::printf("], \"rects\": [");
  for (uint32_t i = 0; i < d.num_rects; i++) {
    const pt_rect_desc& r = d.rects[i];
    std::printf("%s{\"axis\": %d, \"flip\": %d, ", i ? ", " : "", r.axis, r.flip);
    const float v[5] = {r.a0, r.a1, r.b0, r.b1, r.k};
    floats("range", v, 5);
    std::printf("\"material\": %d}", r.material);
  }
  std::printf("], \"boxes\": [");
  for (uint32_t i = 0; i < d.num_boxes; i++) {
    std::printf("%s{", i ? ", " : "");
    floats("pmin", d.boxes[i].pmin, 3);
    floats("pmax", d.boxes[i].pmax, 3);
    std::printf("\"material\": %d}", d.boxes[i].material);
  }
  std::printf("], \"instances\": [");
  for (uint32_t i = 0; i < d.num_instances; i++) {
    const pt_instance_desc& s = d.instances[i];
    std::printf("%s{\"kind\": %d, \"first\": %u, \"count\": %u, \"has_transform\": %d, ", i ? ", " : "", s.kind,
                s.first, s.count, s.has_transform);
    floats("obj_to_world", s.obj_to_world, 12, false);
    std::printf("}");
  }
  std::printf("], \"materials\": [");
  for (uint32_t i = 0; i < d.num_materials; i++) {
    const pt_material_desc& m = d.materials[i];
    std::printf("%s{\"type\": %d, \"tex\": [%d, %d], ", i ? ", " : "", m.type, m.tex[0], m.tex[1]);
    floats("param", m.param, 4, false);
    std::printf("}");
  }
  std::printf("], \"textures\": [");
  for (uint32_t i = 0; i < d.num_textures; i++) {
    const pt_texture_desc& t = d.textures[i];
    std::printf("%s{\"type\": %d, ", i ? ", " : "", t.type);
    floats("rgb", t.rgb, 3, false);
    std::printf("}");
  }
  std::printf("], \"lights\": [");
  for (uint32_t i = 0; i < d.num_lights; i++) {
    const pt_light_desc& l = d.lights[i];
    std::printf("%s{\"type\": %d, \"axis\": %d, ", i ? ", " : "", l.type, l.axis);
    const float v[5] = {l.a0, l.a1, l.b0, l.b1, l.k};
    floats("range", v, 5);
    floats("center", l.center, 3);
    std::printf("\"radius\": %.9g, ", (double)l.radius);
    floats("emission", l.emission, 3, false);
    std::printf("}");
  }
  std::printf("], \"miss\": {\"type\": %d, ", d.miss.type);
  floats("color1", d.miss.color1, 3);
  floats("color2", d.miss.color2, 3, false);
  std::printf("}, \"volumes\": %u, \"camera\": {", d.num_volumes);
  floats("origin", cam.origin, 3);
  floats("lower_left", cam.lower_left, 3);
  floats("horizontal", cam.horizontal, 3);
  floats("vertical", cam.vertical, 3);
  std::printf("\"lens_radius\": %.9g}, \"max_depth\": %d, \"russian\": %d}\n", (double)cam.lens_radius,
              app.context->rec.ivar("maxDepth", -1), app.context->rec.ivar("russian", -1));
  return 0;
}
