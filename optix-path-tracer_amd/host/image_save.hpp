// image_save.hpp -- output of the host scene API (host_includes/image_save.hpp):
// Save_PNG averages, applies sqrt gamma, clamps and scales by 255.99
// (:7-37); Save_HDR averages and applies Reinhard c / (1 + c) (:40-73), with
// the row loop over H rather than W (fixes :46).
#pragma once

#include <cstdio>
#include <vector>

#include "host_common.hpp"
#include "rgbe.hpp"

namespace png_detail {
inline uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t v = i;
      for (int k = 0; k < 8; k++) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
      table[i] = v;
    }
    init = true;
  }
  for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c;
}
inline void put32(std::vector<unsigned char>& o, uint32_t v) {
  o.push_back(v >> 24); o.push_back(v >> 16); o.push_back(v >> 8); o.push_back(v);
}
inline void chunk(FILE* f, const char* type, const std::vector<unsigned char>& data) {
  std::vector<unsigned char> buf;
  put32(buf, (uint32_t)data.size());
  buf.insert(buf.end(), type, type + 4);
  buf.insert(buf.end(), data.begin(), data.end());
  uint32_t c = crc32(buf.data() + 4, buf.size() - 4) ^ 0xFFFFFFFFu;
  put32(buf, c);
  std::fwrite(buf.data(), 1, buf.size(), f);
}
}  // namespace png_detail

// 8-bit RGB PNG, zlib stream of stored (uncompressed) deflate blocks
inline int write_png_rgb(const char* name, int w, int h, const unsigned char* rgb) {
  FILE* f = std::fopen(name, "wb");
  if (!f) return 0;
  const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  std::fwrite(sig, 1, 8, f);
  std::vector<unsigned char> ihdr;
  png_detail::put32(ihdr, (uint32_t)w);
  png_detail::put32(ihdr, (uint32_t)h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
  png_detail::chunk(f, "IHDR", ihdr);
  std::vector<unsigned char> raw;
  for (int y = 0; y < h; y++) {
    raw.push_back(0);
    raw.insert(raw.end(), rgb + (size_t)3 * w * y, rgb + (size_t)3 * w * (y + 1));
  }
  std::vector<unsigned char> z = {0x78, 0x01};
  for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
    size_t len = std::min<size_t>(65535, raw.size() - off);
    bool last = off + len >= raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(len & 0xFF); z.push_back(len >> 8);
    z.push_back(~len & 0xFF); z.push_back((~len >> 8) & 0xFF);
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
    if (last) break;
  }
  uint32_t a = 1, b = 0;
  for (unsigned char c : raw) { a = (a + c) % 65521; b = (b + a) % 65521; }
  png_detail::put32(z, (b << 16) | a);
  png_detail::chunk(f, "IDAT", z);
  png_detail::chunk(f, "IEND", {});
  return std::fclose(f) == 0 ? 1 : 0;
}

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
