// png_write.hpp -- 8-bit RGB PNG writer (stored deflate blocks, so no
// compressor is needed) used by Save_PNG and pt_save_image.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace png_detail {
struct CrcTable {
  uint32_t t[256];
  CrcTable() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t v = i;
      for (int k = 0; k < 8; k++) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
      t[i] = v;
    }
  }
};
// CRC-32 of the PNG specification (reflected 0xEDB88320)
inline uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
  static const CrcTable table;  // initialised once, thread-safe
  for (size_t i = 0; i < n; i++) c = table.t[(c ^ p[i]) & 0xFF] ^ (c >> 8);
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

