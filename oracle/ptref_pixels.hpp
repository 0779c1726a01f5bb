// ptref_pixels.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Render a list of pixels (launch index x + W*y) over a frame range and sum
// de_nan(color()) per pixel in double (raygen.cu:126-168 restricted to the
// listed pixels).  Included by the float oracle (ptref_capi.cpp) and by its
// double build (ptref_capi64.cpp), which converges reference images.
#pragma once

#include <atomic>
#include <thread>
#include <vector>

#include "ptref_render.hpp"

namespace ptref {

// sum: npix * 3 doubles, added to; stats: [0] paths, [1] extension rays,
// [2] shadow rays, [3] paths whose color has a NaN channel
inline void render_pixel_list(const Scene& s, const pt_camera& cam, const pt_render_params& rp,
                              uint32_t frame_begin, uint32_t frame_count, const uint32_t* pix, uint64_t npix,
                              int nthreads, double* sum, uint64_t* stats) {
  Params p{rp.width, rp.height, rp.max_depth, rp.russian != 0, rp.rr_min_depth, (rp.flags & PT_FLAG_ENV_MIS) != 0};
  if (nthreads < 1) nthreads = 1;
  std::atomic<uint64_t> ext{0}, shd{0}, paths{0}, nans{0};
  // work items: (pixel chunk, all frames), so each pixel's frames are summed
  // in frame order by one thread (the result does not depend on threads)
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    PathCounters pc;
    uint64_t np = 0, nn = 0;
    while (true) {
      uint64_t b = next.fetch_add(16);
      if (b >= npix) break;
      uint64_t e = std::min<uint64_t>(npix, b + 16);
      for (uint64_t k = b; k < e; k++) {
        uint32_t x = pix[k] % p.W, y = pix[k] / p.W;
        double a0 = 0, a1 = 0, a2 = 0;
        for (uint32_t f = frame_begin; f < frame_begin + frame_count; f++) {
          PathResult r = trace_path(s, cam, p, x, y, f, pc);
          nn += (r.L.x == r.L.x && r.L.y == r.L.y && r.L.z == r.L.z) ? 0 : 1;
          f3 c = de_nan(r.L);
          a0 += (double)c.x;
          a1 += (double)c.y;
          a2 += (double)c.z;
          np++;
        }
        sum[3 * k] += a0;
        sum[3 * k + 1] += a1;
        sum[3 * k + 2] += a2;
      }
    }
    ext += pc.ext;
    shd += pc.shadow;
    paths += np;
    nans += nn;
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; t++) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  if (stats) {
    stats[0] += paths;
    stats[1] += ext;
    stats[2] += shd;
    stats[3] += nans;
  }
}

}  // namespace ptref
