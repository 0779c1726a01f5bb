// pt_scene_export.cpp -- C entry points (include/pt_scene.h) that build the
// synthetic scenes through the host scene API and hand out the flattened
// pt_scene_desc and camera, for callers that are not C++ (tests, bench).
#include <exception>
#include <string>

#include "../../include/pt_scene.h"
#include "config_scenes.hpp"

struct pt_scene_handle {
  App_State app;
  SceneArena arena;
  ptfacade::FlatScene flat;
  pt_camera camera{};
  pt_scene_info info{};
};

static std::string g_scene_err;

extern "C" {

const char* pt_scene_last_error(void) { return g_scene_err.c_str(); }

pt_status pt_scene_build(int32_t config, uint32_t width, uint32_t height, uint32_t variant, pt_scene_handle** out) {
  if (!out) return PT_ERR_INVALID_ARG;
  *out = nullptr;
  pt_scene_handle* h = new pt_scene_handle();
  try {
    ConfigDefaults d{};
    // first pass for the defaults, then with the requested size (the camera
    // depends on the aspect ratio)
    switch (config) {
      case PT_SCENE_RANDOM_SPHERES: d = ConfigDefaults{400, 267, 16, 50, true, 11}; break;
      case PT_SCENE_CORNELL: d = ConfigDefaults{512, 512, 64, 16, true, 3}; break;
      case PT_SCENE_MOVING_SPHERES: d = ConfigDefaults{400, 267, 16, 50, true, 11}; break;
      case PT_SCENE_NEXT_WEEK: d = ConfigDefaults{500, 500, 64, 50, true, 11}; break;
      case PT_SCENE_MESHES:
        d = ConfigDefaults{1920, 1080, 256, 12, true, 11};
        if (variant > 6) {
          delete h;
          g_scene_err = "mesh scene variant must be 0..6";
          return PT_ERR_INVALID_ARG;
        }
        break;
      case PT_SCENE_BATCH: {
        if (variant >= BATCH_SCENES) {
          delete h;
          g_scene_err = "batch scene index must be 0..63";
          return PT_ERR_INVALID_ARG;
        }
        const BatchSceneSpec bs = batch_scene_spec(variant);
        d = ConfigDefaults{(int)bs.side, (int)bs.side, (int)bs.spp, 16, true, 11};
        break;
      }
      default:
        delete h;
        g_scene_err = "unknown scene config";
        return PT_ERR_INVALID_ARG;
    }
    h->app.W = width ? (int)width : d.W;
    h->app.H = height ? (int)height : d.H;
    if (config == PT_SCENE_RANDOM_SPHERES) d = RandomSpheres(h->app, h->arena, variant);
    else if (config == PT_SCENE_CORNELL) d = CornellBox(h->app, h->arena);
    else if (config == PT_SCENE_BATCH) d = BatchScene(h->app, h->arena, variant);
    else if (config == PT_SCENE_MOVING_SPHERES) d = MovingSpheres(h->app, h->arena, variant);
    else if (config == PT_SCENE_NEXT_WEEK) d = NextWeek(h->app, h->arena, variant);
    else if (variant == 0) d = MeshScene(h->app, h->arena);
    else d = MeshScene(h->app, h->arena, (int)variant, 256, 128);
    h->app.context->rec.flatten(h->flat);
    h->camera = h->app.context->rec.camera();
    h->info.width = (uint32_t)h->app.W;
    h->info.height = (uint32_t)h->app.H;
    h->info.spp = (uint32_t)d.spp;
    h->info.max_depth = (uint32_t)d.depth;
    h->info.russian = d.russian ? 1 : 0;
    h->info.rr_min_depth = (uint32_t)d.rr_min_depth;
    h->info.flags = d.flags;
  } catch (const char* e) {
    g_scene_err = e;
    delete h;
    return PT_ERR_INVALID_ARG;
  } catch (const std::string& e) {
    g_scene_err = e;
    delete h;
    return PT_ERR_INVALID_ARG;
  } catch (const std::exception& e) {
    g_scene_err = e.what();
    delete h;
    return PT_ERR_INVALID_ARG;
  }
  *out = h;
  return PT_OK;
}

// An OBJ/MTL model (mesh.hpp:49-204, as scenes.hpp:377-526 loads its
// models) under a Radiance .hdr environment (importance-sampled, rendered
// with PT_FLAG_ENV_MIS) or, without one, the gradient sky; the camera looks
// at the model's bounding box from +Z.
pt_status pt_scene_build_obj(const char* obj_file, const char* assets_folder, const char* hdr_file, uint32_t width,
                             uint32_t height, pt_scene_handle** out) {
  if (!out || !obj_file) return PT_ERR_INVALID_ARG;
  *out = nullptr;
  pt_scene_handle* h = new pt_scene_handle();
  try {
    const std::string folder = assets_folder ? assets_folder : "";
    h->app.W = width ? (int)width : 512;
    h->app.H = height ? (int)height : 512;
    ObjMesh om;
    std::string err;
    if (!obj_load(folder + obj_file, folder, om, err)) throw std::string("Failed to load/parse .obj: ") + err;
    if (om.v.empty()) throw std::string("the .obj file has no faces");
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (size_t i = 0; i < om.v.size(); i++) {
      lo[i % 3] = std::fmin(lo[i % 3], om.v[i]);
      hi[i % 3] = std::fmax(hi[i % 3], om.v[i]);
    }
    App_State& app = h->app;
    SceneArena& A = h->arena;
    Light_Sampler lights;
    ConfigDefaults d{app.W, app.H, 16, 8, true, 3};
    if (hdr_file) {
      Environment_PDF env;
      lights.pdf.push_back(env.createPDF(app.context));
      lights.sample.push_back(env.createSample(app.context));
      lights.emissions.push_back(make_float3(1.f));
      setRayGenerationProgram(app.context, lights);
      setMissProgram(app.context, HDR, std::string(hdr_file), true);
      d.flags = PT_FLAG_MATERIAL_SORT | PT_FLAG_ENV_MIS;
    } else {
      setRayGenerationProgram(app.context, lights);
      setMissProgram(app.context, GRADIENT, make_float3(1.f), make_float3(0.5f, 0.7f, 1.f));
      d.flags = PT_FLAG_MATERIAL_SORT;
    }
    setExceptionProgram(app.context);
    Group group = app.context->createGroup();
    group->setAcceleration(app.context->createAcceleration("Trbvh"));
    A.meshes.emplace_back(new Mesh(std::string(obj_file), folder, true));
    Mesh_List meshes;
    meshes.push(A.meshes.back().get());
    meshes.addElementsTo(group, app.context);
    app.context["world"]->set(group);
    const float3 c = make_float3(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]));
    const float r = 0.5f * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                     (hi[2] - lo[2]) * (hi[2] - lo[2]));
    Camera camera(c + make_float3(0.f, 0.3f * r, 2.8f * r), c, make_float3(0.f, 1.f, 0.f), 40.f,
                  float(app.W) / float(app.H), 0.f, 10.f, 0.f, 1.f);
    camera.set(app.context);
    app.context->rec.flatten(h->flat);
    h->camera = app.context->rec.camera();
    h->info.width = (uint32_t)app.W;
    h->info.height = (uint32_t)app.H;
    h->info.spp = (uint32_t)d.spp;
    h->info.max_depth = (uint32_t)d.depth;
    h->info.russian = d.russian ? 1 : 0;
    h->info.rr_min_depth = (uint32_t)d.rr_min_depth;
    h->info.flags = d.flags;
  } catch (const char* e) {
    g_scene_err = e;
    delete h;
    return PT_ERR_INVALID_ARG;
  } catch (const std::string& e) {
    g_scene_err = e;
    delete h;
    return PT_ERR_INVALID_ARG;
  } catch (const std::exception& e) {
    g_scene_err = e.what();
    delete h;
    return PT_ERR_INVALID_ARG;
  }
  *out = h;
  return PT_OK;
}

const pt_scene_desc* pt_scene_get_desc(const pt_scene_handle* h) { return h ? &h->flat.desc : nullptr; }
const pt_camera* pt_scene_get_camera(const pt_scene_handle* h) { return h ? &h->camera : nullptr; }
const pt_scene_info* pt_scene_get_info(const pt_scene_handle* h) { return h ? &h->info : nullptr; }
void pt_scene_free(pt_scene_handle* h) { delete h; }

}  // extern "C"
