// ptref_capi64.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// The oracle built over double (Real = double, the 32-bit RNG unchanged):
// converged reference images for relMSE.  Its names carry the ptref64
// prefix and live in namespace ptref64, so the float oracle of
// ptref_capi.cpp links into the same library.
#define PTREF_DOUBLE 1
#define ptref ptref64

#include <exception>
#include <string>

#include "ptref_pixels.hpp"

using namespace ptref64;

namespace {
thread_local std::string g_err64;
}

extern "C" {

const char* ptref64_last_error(void) { return g_err64.c_str(); }

void* ptref64_scene_create(const pt_scene_desc* desc) {
  try {
    return new Scene(*desc);
  } catch (const std::exception& e) {
    g_err64 = e.what();
    return nullptr;
  }
}

void ptref64_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

int32_t ptref64_render_pixels(const void* sp, const pt_camera* cam, const pt_render_params* rp,
                              uint32_t frame_begin, uint32_t frame_count, const uint32_t* pix, uint64_t npix,
                              int32_t nthreads, double* sum, uint64_t* stats) {
  try {
    render_pixel_list(*static_cast<const Scene*>(sp), *cam, *rp, frame_begin, frame_count, pix, npix, nthreads, sum,
                      stats);
    return 0;
  } catch (const std::exception& e) {
    g_err64 = e.what();
    return -1;
  }
}

}  // extern "C"
