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

const pt_scene_desc* pt_scene_get_desc(const pt_scene_handle* h) { return h ? &h->flat.desc : nullptr; }
const pt_camera* pt_scene_get_camera(const pt_scene_handle* h) { return h ? &h->camera : nullptr; }
const pt_scene_info* pt_scene_get_info(const pt_scene_handle* h) { return h ? &h->info : nullptr; }
void pt_scene_free(pt_scene_handle* h) { delete h; }

}  // extern "C"
