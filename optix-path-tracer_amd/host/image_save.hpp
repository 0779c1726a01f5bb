// image_save.hpp -- output of the host scene API (host_includes/image_save.hpp):
// Save_PNG averages, applies sqrt gamma, clamps and scales by 255.99
// (:7-37); Save_HDR averages and applies Reinhard c / (1 + c) (:40-73), with
// the row loop over H rather than W (fixes :46).
#pragma once

#include <cstdio>
#include <vector>

#include "host_common.hpp"
#include "png_write.hpp"
#include "rgbe.hpp"

// image_save.hpp:7-37
inline int Save_PNG(App_State& app, Buffer& buffer) {
  std::vector<unsigned char> arr((size_t)app.W * app.H * 3);
  const float4* cols = (const float4*)buffer->map();
  for (int j = app.H - 1; j >= 0; j--)
    for (int i = 0; i < app.W; i++) {
      int index = app.W * j + i;
      int pixel_index = 3 * (app.W * j + i);
      float3 col = make_float3(cols[index].x, cols[index].y, cols[index].z);
      col = sqrt(col / float(app.samples));
      col = 255.99f * clamp(col, 0.f, 1.f);
      arr[pixel_index + 0] = (int)col.x;
      arr[pixel_index + 1] = (int)col.y;
      arr[pixel_index + 2] = (int)col.z;
    }
  buffer->unmap();
  app.fileName += ".png";
  return write_png_rgb(app.fileName.c_str(), app.W, app.H, arr.data());
}

// image_save.hpp:40-73
inline int Save_HDR(App_State& app, Buffer& buffer) {
  std::vector<float> arr((size_t)app.W * app.H * 3);
  const float4* cols = (const float4*)buffer->map();
  for (int j = app.H - 1; j >= 0; j--)
    for (int i = 0; i < app.W; i++) {
      int index = app.W * j + i;
      int pixel_index = 3 * (app.W * j + i);
      float3 col = make_float3(cols[index].x, cols[index].y, cols[index].z);
      col = col / float(app.samples);
      col = col / (make_float3(1.f) + col);
      arr[pixel_index + 0] = col.x;
      arr[pixel_index + 1] = col.y;
      arr[pixel_index + 2] = col.z;
    }
  buffer->unmap();
  app.fileName += ".hdr";
  return rgbe_save(app.fileName, (uint32_t)app.W, (uint32_t)app.H, arr.data()) ? 1 : 0;
}
