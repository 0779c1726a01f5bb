/*
 * pt_scene.h -- synthetic benchmark scenes built through the host scene API
 * (the C++ facade over pt_api.h that mirrors the reference's host_includes/
 * builders), exported as flattened pt_scene_desc + pt_camera for callers
 * that are not C++.  Part of libpt_hip.so; no GPU is touched.
 *
 * No counterpart in the reference: its scenes are C++ functions
 * (host_includes/scenes.hpp) run inside the GUI application.
 */
#ifndef PT_SCENE_H
#define PT_SCENE_H

#include "pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PT_SCENE_RANDOM_SPHERES 0 /* random-spheres scene  */
#define PT_SCENE_CORNELL 1        /* Cornell box with MIS  */
#define PT_SCENE_MESHES 2         /* displaced meshes under an importance-sampled HDRI */
#define PT_SCENE_BATCH 3          /* scene `variant` (0..63) of the 64-scene batch workload */
#define PT_SCENE_MOVING_SPHERES 4 /* checkered ground, noise and gradient spheres, rect light */
#define PT_SCENE_NEXT_WEEK 5      /* ground boxes, media, Perlin sphere, 1000-sphere group */

typedef struct pt_scene_handle pt_scene_handle;

typedef struct pt_scene_info {
  uint32_t width, height;  /* image size the camera was built for */
  uint32_t spp;            /* samples per pixel of the workload   */
  uint32_t max_depth;
  int32_t russian;
  uint32_t rr_min_depth;
  uint32_t flags;          /* PT_FLAG_* the workload renders with */
  uint32_t reserved;
} pt_scene_info;

/* Build scene `config` for a width x height image (0 = the workload's size).
 * `variant`: the scene RNG seed of PT_SCENE_RANDOM_SPHERES,
 * PT_SCENE_MOVING_SPHERES and PT_SCENE_NEXT_WEEK; for
 * PT_SCENE_MESHES 0 builds the full workload and 1..6 the same layout with
 * octaspheres subdivided `variant` times and a 256x128 sky (small copies for
 * tests); for PT_SCENE_BATCH the scene index: 0..31 sphere scenes of
 * 50..5000 spheres, 32..63 mesh scenes of 10k..2M triangles under a sky
 * light, each with its own square resolution (side 256..1024) and 4, 16 or
 * 64 spp in its info. */
pt_status pt_scene_build(int32_t config, uint32_t width, uint32_t height, uint32_t variant,
                         pt_scene_handle** out);
/* A scene around one Wavefront OBJ model (mesh.hpp:49-204: the file's
 * faces, normals, uvs and per-face MTL materials, each a Lambertian with
 * its map_Kd image, PNG or binary PPM/PGM, or its Ka colour), lit by the
 * Radiance .hdr environment `hdr_file` (importance-sampled; info.flags asks
 * for PT_FLAG_ENV_MIS) or, when hdr_file is NULL, by the gradient sky.
 * Paths inside the files are relative to assets_folder (may be NULL); the
 * camera frames the model's bounding box from +Z; 0 for width / height
 * means 512. */
pt_status pt_scene_build_obj(const char* obj_file, const char* assets_folder, const char* hdr_file,
                             uint32_t width, uint32_t height, pt_scene_handle** out);
/* Views owned by the handle, valid until pt_scene_free. */
const pt_scene_desc* pt_scene_get_desc(const pt_scene_handle* h);
const pt_camera* pt_scene_get_camera(const pt_scene_handle* h);
const pt_scene_info* pt_scene_get_info(const pt_scene_handle* h);
void pt_scene_free(pt_scene_handle* h);
const char* pt_scene_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PT_SCENE_H */
