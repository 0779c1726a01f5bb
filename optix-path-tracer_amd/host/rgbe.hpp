// rgbe.hpp -- Radiance RGBE (.hdr) reader and writer for the host scene API.
//
// Reader: the format lib/HDRloader.cpp handles (header up to an empty line,
// resolution line "-Y h +X w", new-style per-channel run-length scanlines
// with an old-style/flat fallback), converting each pixel to float RGB as
// (mantissa / 256) * 2^(exponent - 128); rows are kept in file order (top
// first).  Writer: flat RGBE scanlines, as image_save.hpp:40-73 saves via
// stbi_write_hdr.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct RGBEImage {
  uint32_t width = 0, height = 0;
  std::vector<float> rgb;  // width * height * 3, row 0 = top
};

namespace rgbe_detail {
inline bool read_old(FILE* f, unsigned char (*scan)[4], int len, bool has_prev) {
  int shift = 0;
  while (len > 0) {
    unsigned char p[4];
    if (std::fread(p, 1, 4, f) != 4) return false;
    if (p[0] == 1 && p[1] == 1 && p[2] == 1) {
      if (!has_prev) return false;
      int count = (int)p[3] << shift;
      if (count > len) return false;
      for (int i = 0; i < count; i++) {
        std::memcpy(scan[0], scan[-1], 4);
        scan++;
        len--;
      }
      shift += 8;
    } else {
      std::memcpy(scan[0], p, 4);
      scan++;
      len--;
      shift = 0;
      has_prev = true;
    }
  }
  return true;
}
inline bool read_scanline(FILE* f, unsigned char (*scan)[4], int len) {
  if (len < 8 || len > 0x7fff) return read_old(f, scan, len, false);
  int c0 = std::fgetc(f);
  if (c0 == EOF) return false;
  if (c0 != 2) {
    std::ungetc(c0, f);
    return read_old(f, scan, len, false);
  }
  int c1 = std::fgetc(f), c2 = std::fgetc(f), c3 = std::fgetc(f);
  if (c1 == EOF || c2 == EOF || c3 == EOF) return false;
  if (c1 != 2 || (c2 & 128)) {
    scan[0][0] = 2;
    scan[0][1] = (unsigned char)c1;
    scan[0][2] = (unsigned char)c2;
    scan[0][3] = (unsigned char)c3;
    return read_old(f, scan + 1, len - 1, true);
  }
  if (((c2 << 8) | c3) != len) return false;
  for (int ch = 0; ch < 4; ch++) {
    for (int j = 0; j < len;) {
      int code = std::fgetc(f);
      if (code == EOF) return false;
      if (code > 128) {
        int run = code & 127;
        int val = std::fgetc(f);
        if (val == EOF || run == 0 || j + run > len) return false;
        while (run--) scan[j++][ch] = (unsigned char)val;
      } else {
        if (code == 0 || j + code > len) return false;
        while (code--) {
          int v = std::fgetc(f);
          if (v == EOF) return false;
          scan[j++][ch] = (unsigned char)v;
        }
      }
    }
  }
  return true;
}
}  // namespace rgbe_detail

inline bool rgbe_load(const std::string& path, RGBEImage& img, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  char line[512];
  if (!std::fgets(line, sizeof(line), f) || (std::strncmp(line, "#?RADIANCE", 10) && std::strncmp(line, "#?RGBE", 6))) {
    std::fclose(f);
    err = "not a Radiance file";
    return false;
  }
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '\n' || line[0] == '\r') break;
    if (!std::strncmp(line, "FORMAT=", 7) && std::strncmp(line + 7, "32-bit_rle_rgbe", 15)) {
      std::fclose(f);
      err = "unsupported pixel format";
      return false;
    }
  }
  long h = 0, w = 0;
  if (!std::fgets(line, sizeof(line), f) || std::sscanf(line, "-Y %ld +X %ld", &h, &w) != 2 || w <= 0 || h <= 0 ||
      w > (1 << 20) || h > (1 << 20)) {
    std::fclose(f);
    err = "unsupported resolution line";
    return false;
  }
  img.width = (uint32_t)w;
  img.height = (uint32_t)h;
  img.rgb.assign((size_t)w * h * 3, 0.f);
  std::vector<unsigned char> scan((size_t)w * 4);
  for (long y = 0; y < h; y++) {
    auto* s = reinterpret_cast<unsigned char(*)[4]>(scan.data());
    if (!rgbe_detail::read_scanline(f, s, (int)w)) {
      std::fclose(f);
      err = "truncated or corrupt scanline";
      return false;
    }
    for (long x = 0; x < w; x++) {
      int e = (int)s[x][3] - 128;
      float scale = std::ldexp(1.0f, e);
      float* o = &img.rgb[3 * ((size_t)y * w + x)];
      o[0] = ((float)s[x][0] / 256.0f) * scale;
      o[1] = ((float)s[x][1] / 256.0f) * scale;
      o[2] = ((float)s[x][2] / 256.0f) * scale;
    }
  }
  std::fclose(f);
  return true;
}

// float RGB -> flat RGBE scanlines
inline bool rgbe_save(const std::string& path, uint32_t w, uint32_t h, const float* rgb) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
  for (size_t i = 0; i < (size_t)w * h; i++) {
    float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
    float m = std::fmax(r, std::fmax(g, b));
    unsigned char p[4] = {0, 0, 0, 0};
    if (m > 1e-32f) {
      int e;
      float fr = std::frexp(m, &e);  // m = fr * 2^e, fr in [0.5, 1)
      float s = fr * 256.0f / m;
      p[0] = (unsigned char)(r * s);
      p[1] = (unsigned char)(g * s);
      p[2] = (unsigned char)(b * s);
      p[3] = (unsigned char)(e + 128);
    }
    std::fwrite(p, 1, 4, f);
  }
  return std::fclose(f) == 0;
}
