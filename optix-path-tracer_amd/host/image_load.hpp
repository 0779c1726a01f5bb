// image_load.hpp -- 8-bit image readers behind Image_Texture
// (host_includes/textures.hpp:114-175 loads with stb_image).  PNG (ISO/IEC
// 15948: zlib stream inflated by the system zlib; bit depths 1, 2, 4, 8 and
// 16; grey, grey + alpha, RGB, RGBA and palette images; the five scanline
// filters; Adam7-interlaced files are refused), baseline / extended
// sequential JPEG (ITU-T T.81, Huffman coded), TGA (raw and RLE), BMP
// (uncompressed 24/32-bit) and binary PPM / PGM (P6 / P5, maxval up to
// 65535).  Pixels come out as RGBA8, row 0 = the file's top row; 16-bit
// samples keep their high byte, grey is replicated to RGB, alpha is 255
// where the file has none (palette alpha from tRNS).  stb_image's JPEG
// decoder upsamples chroma by interpolation; this one replicates.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
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

// ---------------------------------------------------------------- JPEG
// Baseline and extended sequential (8-bit) JPEG, ITU-T T.81: Huffman
// entropy coding, 1 or 3 components with any sampling factors up to 4,
// interleaved and non-interleaved scans, restart intervals.  Inverse DCT in
// double precision from the definition (A.3.3), chroma upsampled by
// replication, JFIF YCbCr -> RGB.  Progressive and arithmetic-coded files
// are refused.
namespace jpeg_detail {
static const uint8_t zigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                   12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                   35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                   58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
struct Huff {
  bool defined = false;
  uint8_t vals[256] = {};
  int mincode[17] = {}, maxcode[18] = {}, valptr[17] = {};
};
struct Comp {
  int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
  int bw = 0, bh = 0;  // blocks across and down, padded to whole MCUs
  int dc = 0;
  std::vector<uint8_t> pix;  // bw*8 x bh*8 samples
};
struct Bits {
  const std::vector<uint8_t>& f;
  size_t pos;
  uint32_t acc = 0;
  int n = 0;
  bool at_marker = false;
  Bits(const std::vector<uint8_t>& f, size_t p) : f(f), pos(p) {}
  bool fill() {
    while (n <= 24) {
      uint32_t b = 0;
      if (!at_marker && pos < f.size()) {
        b = f[pos];
        if (b == 0xFF) {
          uint8_t nx = pos + 1 < f.size() ? f[pos + 1] : 0;
          if (nx == 0x00) pos += 2;
          else { at_marker = true; b = 0; }  // a marker: feed zeros
        } else {
          pos++;
        }
      }
      acc |= b << (24 - n);
      n += 8;
    }
    return true;
  }
  int bit() {
    if (n == 0) fill();
    int b = (int)(acc >> 31);
    acc <<= 1;
    n--;
    return b;
  }
  int receive(int s) {
    int v = 0;
    for (int i = 0; i < s; i++) v = (v << 1) | bit();
    return v;
  }
  void reset() { acc = 0; n = 0; }
};
inline int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }
inline int decode(Bits& b, const Huff& h) {
  int code = 0;
  for (int l = 1; l <= 16; l++) {
    code = (code << 1) | b.bit();
    if (h.maxcode[l] >= 0 && code <= h.maxcode[l] && code >= h.mincode[l]) return h.vals[h.valptr[l] + code - h.mincode[l]];
  }
  return -1;
}
}  // namespace jpeg_detail

inline bool load_jpeg(const std::vector<uint8_t>& f, LdrImage& img, std::string& err) {
  using namespace jpeg_detail;
  if (f.size() < 4 || f[0] != 0xFF || f[1] != 0xD8) return err = "not a JPEG file", false;
  uint16_t qt[4][64] = {};
  bool qdef[4] = {};
  Huff hdc[4], hac[4];
  std::vector<Comp> comps;
  int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0, restart = 0;
  bool frame = false, done = false;
  double cosv[8][8];
  for (int x = 0; x < 8; x++)
    for (int u = 0; u < 8; u++) cosv[x][u] = std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0);
  size_t p = 2;
  auto be16 = [&](size_t q) { return (int)f[q] << 8 | f[q + 1]; };
  while (!done) {
    while (p < f.size() && f[p] != 0xFF) p++;
    while (p < f.size() && f[p] == 0xFF) p++;
    if (p >= f.size()) return err = "truncated JPEG", false;
    const int m = f[p++];
    if (m == 0xD9) break;                       // EOI
    if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
    if (p + 2 > f.size()) return err = "truncated JPEG", false;
    const int len = be16(p);
    if (len < 2 || p + len > f.size()) return err = "truncated JPEG segment", false;
    const size_t q0 = p + 2, q1 = p + len;
    if (m == 0xDB) {  // DQT
      for (size_t q = q0; q < q1;) {
        const int pq = f[q] >> 4, tq = f[q] & 15;
        if (tq > 3) return err = "bad JPEG quantization table", false;
        q++;
        for (int k = 0; k < 64; k++) {
          if (q + (pq ? 2 : 1) > q1) return err = "bad JPEG quantization table", false;
          qt[tq][k] = pq ? (uint16_t)be16(q) : f[q];
          q += pq ? 2 : 1;
        }
        qdef[tq] = true;
      }
    } else if (m == 0xC4) {  // DHT
      for (size_t q = q0; q < q1;) {
        if (q + 17 > q1) return err = "bad JPEG Huffman table", false;
        const int tc = f[q] >> 4, th = f[q] & 15;
        if (tc > 1 || th > 3) return err = "bad JPEG Huffman table", false;
        Huff& h = tc ? hac[th] : hdc[th];
        int total = 0, code = 0;
        for (int l = 1; l <= 16; l++) {
          const int cnt = f[q + l];
          h.valptr[l] = total;
          h.mincode[l] = code;
          code += cnt;
          total += cnt;
          h.maxcode[l] = cnt ? code - 1 : -1;
          code <<= 1;
        }
        if (total > 256 || q + 17 + total > q1) return err = "bad JPEG Huffman table", false;
        for (int k = 0; k < total; k++) h.vals[k] = f[q + 17 + k];
        h.defined = true;
        q += 17 + total;
      }
    } else if (m == 0xDD) {  // DRI
      if (len < 4) return err = "bad JPEG restart interval", false;
      restart = be16(q0);
    } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1
      if (len < 8 || q0 + 6 + 3 * (size_t)f[q0 + 5] > q1) return err = "bad JPEG frame header", false;
      if (f[q0] != 8) return err = "only 8-bit JPEG images are read", false;
      H = be16(q0 + 1);
      W = be16(q0 + 3);
      const int nc = f[q0 + 5];
      if (W <= 0 || H <= 0 || (nc != 1 && nc != 3)) return err = "unsupported JPEG frame (size or components)", false;
      if ((uint64_t)W * H > (1ull << 28)) return err = "JPEG image too large", false;
      comps.resize(nc);
      for (int c = 0; c < nc; c++) {
        const size_t q = q0 + 6 + 3 * c;
        comps[c].id = f[q];
        comps[c].h = f[q + 1] >> 4;
        comps[c].v = f[q + 1] & 15;
        comps[c].tq = f[q + 2] & 3;
        if (comps[c].h < 1 || comps[c].h > 4 || comps[c].v < 1 || comps[c].v > 4)
          return err = "bad JPEG sampling factors", false;
        hmax = std::max(hmax, comps[c].h);
        vmax = std::max(vmax, comps[c].v);
      }
      mcux = (W + 8 * hmax - 1) / (8 * hmax);
      mcuy = (H + 8 * vmax - 1) / (8 * vmax);
      for (auto& c : comps) {
        c.bw = mcux * c.h;
        c.bh = mcuy * c.v;
        c.pix.assign((size_t)c.bw * 8 * c.bh * 8, 0);
      }
      frame = true;
    } else if (m == 0xC2 || m == 0xC6 || m == 0xCA || m == 0xCE) {
      return err = "progressive JPEG images are not supported", false;
    } else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
      return err = "lossless, hierarchical and arithmetic-coded JPEG images are not supported", false;
    } else if (m == 0xDA) {  // SOS
      if (!frame) return err = "JPEG scan before the frame header", false;
      if (len < 3 || q0 + 1 + 2 * (size_t)f[q0] + 3 > q1 || f[q0] == 0 || f[q0] > 4)
        return err = "bad JPEG scan header", false;
      const int ns = f[q0];
      std::vector<int> sc(ns);
      for (int k = 0; k < ns; k++) {
        const int cid = f[q0 + 1 + 2 * k];
        int ci = -1;
        for (size_t c = 0; c < comps.size(); c++)
          if (comps[c].id == cid) ci = (int)c;
        if (ci < 0) return err = "JPEG scan names an unknown component", false;
        comps[ci].td = f[q0 + 2 + 2 * k] >> 4;
        comps[ci].ta = f[q0 + 2 + 2 * k] & 15;
        if (comps[ci].td > 3 || comps[ci].ta > 3 || !hdc[comps[ci].td].defined || !hac[comps[ci].ta].defined ||
            !qdef[comps[ci].tq])
          return err = "JPEG scan uses an undefined table", false;
        sc[k] = ci;
      }
      for (int ci : sc) comps[ci].dc = 0;
      Bits b(f, q1);
      // one block: Huffman decode, dequantize, inverse DCT, level shift
      auto block = [&](Comp& c, int bx, int by) -> bool {
        double F[64] = {};
        int s = decode(b, hdc[c.td]);
        if (s < 0 || s > 11) return false;
        c.dc += extend(b.receive(s), s);
        F[0] = (double)c.dc * qt[c.tq][0];
        for (int k = 1; k < 64;) {
          const int rs = decode(b, hac[c.ta]);
          if (rs < 0) return false;
          const int r = rs >> 4, ss = rs & 15;
          if (ss == 0) {
            if (r != 15) break;
            k += 16;
            continue;
          }
          k += r;
          if (k > 63) return false;
          F[zigzag[k]] = (double)extend(b.receive(ss), ss) * qt[c.tq][k];
          k++;
        }
        double tmp[64];
        for (int y = 0; y < 8; y++)
          for (int u = 0; u < 8; u++) {
            double a = 0;
            for (int v = 0; v < 8; v++) a += (v ? 1.0 : std::sqrt(0.5)) * F[v * 8 + u] * cosv[y][v];
            tmp[y * 8 + u] = a;
          }
        const size_t stride = (size_t)c.bw * 8;
        for (int y = 0; y < 8; y++)
          for (int x = 0; x < 8; x++) {
            double a = 0;
            for (int u = 0; u < 8; u++) a += (u ? 1.0 : std::sqrt(0.5)) * tmp[y * 8 + u] * cosv[x][u];
            const double val = std::floor(a / 4.0 + 128.0 + 0.5);
            c.pix[((size_t)by * 8 + y) * stride + (size_t)bx * 8 + x] =
                (uint8_t)(val < 0 ? 0 : (val > 255 ? 255 : val));
          }
        return true;
      };
      int units = 0;
      auto restart_check = [&]() -> bool {
        if (!restart || ++units % restart) return true;
        // byte-align and step over the RSTn marker
        b.reset();
        size_t q = b.pos;
        if (!(q + 1 < f.size() && f[q] == 0xFF && f[q + 1] >= 0xD0 && f[q + 1] <= 0xD7)) return false;
        b.pos = q + 2;
        b.at_marker = false;
        for (int ci : sc) comps[ci].dc = 0;
        return true;
      };
      if (ns == 1) {
        Comp& c = comps[sc[0]];
        const int cw = ((W * c.h + hmax - 1) / hmax + 7) / 8, ch = ((H * c.v + vmax - 1) / vmax + 7) / 8;
        const long total = (long)cw * ch;
        for (long n = 0; n < total; n++) {
          if (!block(c, (int)(n % cw), (int)(n / cw))) return err = "corrupt JPEG entropy data", false;
          if (n + 1 < total && !restart_check()) return err = "missing JPEG restart marker", false;
        }
      } else {
        const long total = (long)mcux * mcuy;
        for (long n = 0; n < total; n++) {
          const int mx = (int)(n % mcux), my = (int)(n / mcux);
          for (int ci : sc) {
            Comp& c = comps[ci];
            for (int v = 0; v < c.v; v++)
              for (int h = 0; h < c.h; h++)
                if (!block(c, mx * c.h + h, my * c.v + v)) return err = "corrupt JPEG entropy data", false;
          }
          if (n + 1 < total && !restart_check()) return err = "missing JPEG restart marker", false;
        }
      }
      p = b.pos;
      continue;
    }
    p = q1;
  }
  if (!frame) return err = "JPEG without a frame", false;
  img.width = (uint32_t)W;
  img.height = (uint32_t)H;
  img.rgba.assign((size_t)W * H * 4, 255);
  auto sample = [&](const Comp& c, int x, int y) {
    return (double)c.pix[(size_t)(y * c.v / vmax) * c.bw * 8 + (size_t)(x * c.h / hmax)];
  };
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      uint8_t* o = &img.rgba[4 * ((size_t)y * W + x)];
      if (comps.size() == 1) {
        o[0] = o[1] = o[2] = (uint8_t)sample(comps[0], x, y);
        continue;
      }
      const double Y = sample(comps[0], x, y), cb = sample(comps[1], x, y) - 128.0,
                   cr = sample(comps[2], x, y) - 128.0;
      const double rgb[3] = {Y + 1.402 * cr, Y - 0.344136 * cb - 0.714136 * cr, Y + 1.772 * cb};
      for (int k = 0; k < 3; k++) {
        const double v = std::floor(rgb[k] + 0.5);
        o[k] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
      }
    }
  return true;
}

// ----------------------------------------------------------------- TGA
// Truevision TGA: types 2 / 10 (true colour, raw / RLE, 24 or 32 bits,
// BGR(A)) and 3 / 11 (grey, 8 bits); rows bottom-up unless descriptor bit
// 5 says top-down.
inline bool load_tga(const std::vector<uint8_t>& f, LdrImage& img, std::string& err) {
  if (f.size() < 18) return err = "not a TGA file", false;
  const int idlen = f[0], cmap = f[1], type = f[2];
  const int W = f[12] | f[13] << 8, H = f[14] | f[15] << 8, bpp = f[16], desc = f[17];
  if (cmap != 0) return err = "colour-mapped TGA images are not supported", false;
  const bool grey = type == 3 || type == 11, rle = type == 10 || type == 11;
  if (!(type == 2 || type == 3 || type == 10 || type == 11)) return err = "unsupported TGA image type", false;
  if ((grey && bpp != 8) || (!grey && bpp != 24 && bpp != 32)) return err = "unsupported TGA pixel depth", false;
  if (W <= 0 || H <= 0) return err = "bad TGA size", false;
  const int bytes = bpp / 8;
  std::vector<uint8_t> px((size_t)W * H * bytes);
  size_t p = 18 + (size_t)idlen, o = 0;
  while (o < px.size()) {
    if (!rle) {
      if (p + px.size() > f.size()) return err = "truncated TGA data", false;
      std::memcpy(px.data(), &f[p], px.size());
      break;
    }
    if (p >= f.size()) return err = "truncated TGA data", false;
    const int hdr = f[p++], n = (hdr & 0x7F) + 1;
    if (o + (size_t)n * bytes > px.size()) return err = "corrupt TGA run", false;
    if (hdr & 0x80) {
      if (p + bytes > f.size()) return err = "truncated TGA data", false;
      for (int k = 0; k < n; k++) std::memcpy(&px[o + (size_t)k * bytes], &f[p], bytes);
      p += bytes;
    } else {
      if (p + (size_t)n * bytes > f.size()) return err = "truncated TGA data", false;
      std::memcpy(&px[o], &f[p], (size_t)n * bytes);
      p += (size_t)n * bytes;
    }
    o += (size_t)n * bytes;
  }
  img.width = (uint32_t)W;
  img.height = (uint32_t)H;
  img.rgba.assign((size_t)W * H * 4, 255);
  const bool top = (desc & 0x20) != 0;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const uint8_t* s = &px[((size_t)(top ? y : H - 1 - y) * W + x) * bytes];
      uint8_t* d = &img.rgba[4 * ((size_t)y * W + x)];
      if (grey) {
        d[0] = d[1] = d[2] = s[0];
      } else {
        d[0] = s[2];
        d[1] = s[1];
        d[2] = s[0];
        if (bytes == 4) d[3] = s[3];
      }
    }
  return true;
}

// ----------------------------------------------------------------- BMP
// Windows bitmap, uncompressed 24 or 32 bits (BGR(A)), rows padded to 4
// bytes, bottom-up for a positive height, top-down for a negative one.
inline bool load_bmp(const std::vector<uint8_t>& f, LdrImage& img, std::string& err) {
  if (f.size() < 54 || f[0] != 'B' || f[1] != 'M') return err = "not a BMP file", false;
  auto le32 = [&](size_t q) { return (uint32_t)f[q] | (uint32_t)f[q + 1] << 8 | (uint32_t)f[q + 2] << 16 | (uint32_t)f[q + 3] << 24; };
  const uint32_t off = le32(10);
  const int32_t W = (int32_t)le32(18), Hs = (int32_t)le32(22);
  const int bpp = f[28] | f[29] << 8;
  const uint32_t comp = le32(30);
  if (comp != 0 || (bpp != 24 && bpp != 32)) return err = "only uncompressed 24/32-bit BMP images are read", false;
  if (W <= 0 || Hs == 0 || Hs == INT32_MIN) return err = "bad BMP size", false;
  const int H = Hs < 0 ? -Hs : Hs, bytes = bpp / 8;
  const size_t row = ((size_t)W * bytes + 3) / 4 * 4;
  if ((uint64_t)W * H > (1ull << 28) || off + row * H > f.size()) return err = "truncated BMP data", false;
  img.width = (uint32_t)W;
  img.height = (uint32_t)H;
  img.rgba.assign((size_t)W * H * 4, 255);
  for (int y = 0; y < H; y++) {
    const uint8_t* r = &f[off + row * (size_t)(Hs > 0 ? H - 1 - y : y)];
    for (int x = 0; x < W; x++) {
      uint8_t* d = &img.rgba[4 * ((size_t)y * W + x)];
      d[0] = r[x * bytes + 2];
      d[1] = r[x * bytes + 1];
      d[2] = r[x * bytes];
      if (bytes == 4) d[3] = r[x * bytes + 3];
    }
  }
  return true;
}

inline bool load_image(const std::string& path, LdrImage& img, std::string& err) {
  std::vector<uint8_t> f;
  if (!image_detail::read_file(path, f)) return err = "cannot read " + path, false;
  if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') return load_png(f, img, err);
  if (f.size() >= 2 && f[0] == 'P' && (f[1] == '6' || f[1] == '5')) return load_pnm(f, img, err);
  if (f.size() >= 3 && f[0] == 0xFF && f[1] == 0xD8 && f[2] == 0xFF) return load_jpeg(f, img, err);
  if (f.size() >= 2 && f[0] == 'B' && f[1] == 'M') return load_bmp(f, img, err);
  std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : "";
  for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
  if (ext == ".tga") return load_tga(f, img, err);
  return err = path + ": PNG, JPEG, TGA, BMP and binary PPM/PGM images are read", false;
}
