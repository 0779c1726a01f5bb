// image_load.hpp -- 8-bit image readers behind Image_Texture
// (host_includes/textures.hpp:114-175 loads with stb_image).  PNG (ISO/IEC
// 15948: zlib stream inflated by the system zlib; bit depths 1, 2, 4, 8 and
// 16; grey, grey + alpha, RGB, RGBA and palette images; the five scanline
// filters; Adam7-interlaced files are refused) and binary PPM / PGM (P6 /
// P5, maxval up to 65535).  Pixels come out as RGBA8, row 0 = the file's top
// row; 16-bit samples keep their high byte, grey is replicated to RGB,
// alpha is 255 where the file has none (palette alpha from tRNS).
#pragma once

#include <zlib.h>

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct LdrImage {
  uint32_t width = 0, height = 0;
  std::vector<uint8_t> rgba;  // width * height * 4, row 0 = top
};

namespace image_detail {
inline bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)n);
  bool ok = n == 0 || std::fread(out.data(), 1, (size_t)n, f) == (size_t)n;
  std::fclose(f);
  return ok;
}
inline uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
inline uint8_t paeth(int a, int b, int c) {
  int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  return (uint8_t)(pb <= pc ? b : c);
}
}  // namespace image_detail

inline bool load_png(const std::vector<uint8_t>& f, LdrImage& img, std::string& err) {
  using namespace image_detail;
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return err = "not a PNG file", false;
  uint32_t W = 0, H = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  size_t p = 8;
  bool seen_end = false;
  while (p + 12 <= f.size()) {
    const uint32_t len = be32(&f[p]);
    if (len > f.size() - p - 12) return err = "truncated PNG chunk", false;
    const char* type = (const char*)&f[p + 4];
    const uint8_t* d = &f[p + 8];
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len < 13) return err = "bad IHDR", false;
      W = be32(d);
      H = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      if (d[10] != 0 || d[11] != 0) return err = "unknown PNG compression or filter method", false;
      interlace = d[12];
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(d, d + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(d, d + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      seen_end = true;
      break;
    }
    p += 12 + (size_t)len;
  }
  if (!seen_end || W == 0 || H == 0) return err = "incomplete PNG", false;
  if (interlace) return err = "interlaced PNG images are not supported", false;
  int ch;
  switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: return err = "unknown PNG colour type", false;
  }
  const bool ok_depth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                        (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                        ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
  if (!ok_depth) return err = "invalid PNG bit depth", false;
  if (ctype == 3 && plte.empty()) return err = "palette image without PLTE", false;
  if ((uint64_t)W * H > (1ull << 28)) return err = "PNG image too large", false;
  const size_t bits = (size_t)ch * depth;
  const size_t rowbytes = ((size_t)W * bits + 7) / 8;
  const size_t bpp = bits >= 8 ? bits / 8 : 1;  // filter byte distance
  std::vector<uint8_t> raw((rowbytes + 1) * H);
  uLongf got = (uLongf)raw.size();
  if (uncompress(raw.data(), &got, idat.data(), (uLong)idat.size()) != Z_OK || got != raw.size())
    return err = "corrupt PNG image data", false;
  // undo the scanline filters in place (rows of rowbytes after a filter byte)
  std::vector<uint8_t> prev(rowbytes, 0), cur(rowbytes);
  std::vector<uint8_t> plain(rowbytes * H);
  for (uint32_t y = 0; y < H; y++) {
    const uint8_t* r = &raw[(size_t)y * (rowbytes + 1)];
    const int ft = r[0];
    for (size_t i = 0; i < rowbytes; i++) {
      const int x = r[1 + i], a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
      int v;
      switch (ft) {
        case 0: v = x; break;
        case 1: v = x + a; break;
        case 2: v = x + b; break;
        case 3: v = x + ((a + b) >> 1); break;
        case 4: v = x + paeth(a, b, c); break;
        default: return err = "unknown PNG filter type", false;
      }
      cur[i] = (uint8_t)v;
    }
    std::memcpy(&plain[(size_t)y * rowbytes], cur.data(), rowbytes);
    prev.swap(cur);
  }
  img.width = W;
  img.height = H;
  img.rgba.assign((size_t)W * H * 4, 255);
  for (uint32_t y = 0; y < H; y++) {
    const uint8_t* row = &plain[(size_t)y * rowbytes];
    for (uint32_t x = 0; x < W; x++) {
      uint8_t* o = &img.rgba[4 * ((size_t)y * W + x)];
      // sample s of pixel x (high byte of 16-bit samples; sub-byte samples
      // scaled to 0..255 for grey, used as indices for palettes)
      auto sample = [&](int s) -> int {
        if (depth == 8) return row[(size_t)x * ch + s];
        if (depth == 16) return row[2 * ((size_t)x * ch + s)];
        const size_t bit = (size_t)x * depth;
        return (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
      };
      if (ctype == 3) {
        const int k = sample(0);
        if ((size_t)3 * k + 2 >= plte.size()) return err = "PNG palette index out of range", false;
        o[0] = plte[3 * k];
        o[1] = plte[3 * k + 1];
        o[2] = plte[3 * k + 2];
        if ((size_t)k < trns.size()) o[3] = trns[k];
      } else if (ch <= 2) {
        int g = sample(0);
        if (depth < 8) g = g * 255 / ((1 << depth) - 1);
        o[0] = o[1] = o[2] = (uint8_t)g;
        if (ch == 2) o[3] = (uint8_t)sample(1);
      } else {
        o[0] = (uint8_t)sample(0);
        o[1] = (uint8_t)sample(1);
        o[2] = (uint8_t)sample(2);
        if (ch == 4) o[3] = (uint8_t)sample(3);
      }
    }
  }
  return true;
}

// binary PPM (P6) / PGM (P5)
inline bool load_pnm(const std::vector<uint8_t>& f, LdrImage& img, std::string& err) {
  size_t p = 2;
  auto token = [&](long& v) {
    for (;;) {
      while (p < f.size() && std::isspace(f[p])) p++;
      if (p < f.size() && f[p] == '#') {
        while (p < f.size() && f[p] != '\n') p++;
        continue;
      }
      break;
    }
    if (p >= f.size() || !std::isdigit(f[p])) return false;
    v = 0;
    while (p < f.size() && std::isdigit(f[p])) v = v * 10 + (f[p++] - '0');
    return true;
  };
  if (f.size() < 3 || f[0] != 'P' || (f[1] != '6' && f[1] != '5')) return err = "not a binary PPM/PGM file", false;
  const int ch = f[1] == '6' ? 3 : 1;
  long w, h, maxv;
  if (!token(w) || !token(h) || !token(maxv) || w <= 0 || h <= 0 || maxv <= 0 || maxv > 65535)
    return err = "bad PPM/PGM header", false;
  if (w > 65536 || h > 65536 || (uint64_t)w * h > (1ull << 28)) return err = "PPM/PGM image too large", false;
  p++;  // the single whitespace after maxval
  const size_t bytes = maxv > 255 ? 2 : 1;
  if (f.size() < p + (size_t)w * h * ch * bytes) return err = "truncated PPM/PGM data", false;
  img.width = (uint32_t)w;
  img.height = (uint32_t)h;
  img.rgba.assign((size_t)w * h * 4, 255);
  for (size_t i = 0; i < (size_t)w * h; i++)
    for (int c = 0; c < 3; c++) {
      const size_t s = (i * ch + (ch == 3 ? c : 0)) * bytes;
      const long v = bytes == 2 ? ((long)f[p + s] << 8 | f[p + s + 1]) : f[p + s];
      img.rgba[4 * i + c] = (uint8_t)((v * 255 + maxv / 2) / maxv);
    }
  return true;
}

inline bool load_image(const std::string& path, LdrImage& img, std::string& err) {
  std::vector<uint8_t> f;
  if (!image_detail::read_file(path, f)) return err = "cannot read " + path, false;
  if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') return load_png(f, img, err);
  if (f.size() >= 2 && f[0] == 'P' && (f[1] == '6' || f[1] == '5')) return load_pnm(f, img, err);
  return err = path + ": only PNG and binary PPM/PGM images are read", false;
}
