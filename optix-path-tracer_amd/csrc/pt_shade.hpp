// pt_shade.hpp -- hit records, textures, miss programs and light sampling of
// the gfx950 path tracer.
#pragma once

#include "pt_trace.hpp"

namespace pt {

struct HitRec {
  v3 P, N, Wo;
  float u, v;
  int mat;
  int fidx;  // the face index a vector texture selects by (0 off meshes)
};

// Get_HitRecord of sphere.cu:65-93, aarect.cu:71-115, box.cu:24-33 + 56-81,
// triangle.cu:85-149, rebuilt in object space from the object-space ray and
// then transformed to world space.  u, v are only produced when `need_uv`.
// KINDS / EXT as in trace_closest: only the kinds that can occur are
// compiled in.
template <int KINDS = 15, bool EXT = true>
PTD HitRec hit_record(const DevScene& sc, int inst_pos, int prim, float t, v3 o, v3 d, bool need_uv_hint,
                      float time = 0.f) {
  const uint4* ir = sc.insts + (size_t)inst_pos * 8;
  float4 w0 = as_f4(ir[0]), w1 = as_f4(ir[1]), w2 = as_f4(ir[2]);
  uint4 info = ir[3];
  uint32_t kind = info.z & 0xFFu;
  bool xf = (info.z >> 8) != 0;
  v3 ro = o, rd = d;
  if (xf) {
    ro = xf_point(w0, w1, w2, o);
    rd = xf_vector(w0, w1, w2, d);
  }
  HitRec r;
  r.Wo = normalize(-d);
  r.u = 0.f;
  r.v = 0.f;
  r.fidx = 0;
  v3 hp = ro, n = mk3(0.f, 1.f, 0.f);  // every kind that occurs sets both
  r.mat = 0;
  if ((KINDS & 1) && kind == PT_GEOM_SPHERES) {
    float4 s = sc.sph_o[prim];
    r.mat = sc.sph_mat[prim];
    hp = ro + t * rd;
    n = (hp - mk3(s)) / s.w;
    if (need_uv_hint) {
      float phi = patan2(n.z, n.x);
      float theta = pasin(n.y);
      r.u = 1.f - (phi + PI_F) / (2.f * PI_F);
      r.v = (theta + PI_F / 2.f) / PI_F;
    }
  } else if ((KINDS & 2) && kind == PT_GEOM_RECTS) {
    float4 A = sc.rect_o[2 * prim], B = sc.rect_o[2 * prim + 1];
    r.mat = sc.rect_mat[prim];
    int axis = (int)__float_as_uint(B.y);
    hp = ro + t * rd;
    // aarect.cu:88-100; u, v only where a texture reads them
    if (axis == 0) {
      n = mk3(1.f, 0.f, 0.f);
      if (need_uv_hint) { r.u = (hp.y - A.x) / (A.y - A.x); r.v = (hp.z - A.z) / (A.w - A.z); }
    } else if (axis == 1) {
      n = mk3(0.f, 1.f, 0.f);
      if (need_uv_hint) { r.u = (hp.x - A.x) / (A.y - A.x); r.v = (hp.z - A.z) / (A.w - A.z); }
    } else {
      n = mk3(0.f, 0.f, 1.f);
      if (need_uv_hint) { r.u = (hp.x - A.x) / (A.y - A.x); r.v = (hp.y - A.z) / (A.w - A.z); }
    }
    if (__float_as_uint(B.z)) n = -n;
  } else if ((KINDS & 4) && kind == PT_GEOM_BOXES) {
    float4 lo = sc.box_o[2 * prim], hi = sc.box_o[2 * prim + 1];
    r.mat = sc.box_mat[prim];
    hp = ro + t * rd;
    // the slab distances exactly as the intersection formed them
    v3 rinv = mk3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    v3 t0 = mulv(mk3(lo) - ro, rinv);
    v3 t1 = mulv(mk3(hi) - ro, rinv);
    v3 neg = mk3(t == t0.x ? 1.f : 0.f, t == t0.y ? 1.f : 0.f, t == t0.z ? 1.f : 0.f);
    v3 pos = mk3(t == t1.x ? 1.f : 0.f, t == t1.y ? 1.f : 0.f, t == t1.z ? 1.f : 0.f);
    n = pos - neg;
  } else if (EXT && kind == PT_GEOM_CYLINDERS) {
    // cylinder.cu:70-97: radial normal, no uv
    float4 A = sc.cyl_o[2 * prim];
    r.mat = sc.cyl_mat[prim];
    hp = ro + t * rd;
    n = normalize(hp - mk3(A.x, hp.y, A.z));
  } else if (EXT && kind == PT_GEOM_MOVING_SPHERES) {
    // moving_sphere.cu:78-91 with the center at the path's time, in object space
    float4 A = sc.mov_o[3 * prim], B = sc.mov_o[3 * prim + 1], C = sc.mov_o[3 * prim + 2];
    r.mat = sc.mov_mat[prim];
    v3 c = moving_center(A, B, C, time);
    hp = ro + t * rd;
    n = (hp - c) / A.w;
    if (need_uv_hint) {
      float phi = patan2(n.z, n.x);
      float theta = pasin(n.y);
      r.u = 1.f - (phi + PI_F) / (2.f * PI_F);
      r.v = (theta + PI_F / 2.f) / PI_F;
    }
  } else if (EXT && kind == PT_GEOM_VOLUMES) {
    // volume_sphere.cu:88-100: the scatter point, normal +X, no uv
    r.mat = sc.vol_mat[prim];
    hp = ro + t * rd;
    n = mk3(1.f, 0.f, 0.f);
  } else if (KINDS & 8) {
    const DevMesh& m = sc.meshes[ir[7].z];
    uint4 ix = sc.idx[m.ibase + prim];
    v3 a = mk3(sc.vtx[m.vbase + ix.x]), b = mk3(sc.vtx[m.vbase + ix.y]), c = mk3(sc.vtx[m.vbase + ix.z]);
    TriRay tr = tri_setup(rd);
    float tt = 0.f, b1 = 0.f, b2 = 0.f;
    tri_test(tr, ro, a, b, c, tt, b1, b2);
    float b0 = 1.f - b1 - b2;
    hp = a * b0 + b * b1 + c * b2;
    n = cross(b - a, c - a);
    if (m.has_n) {
      n = mk3(sc.nrm[m.vbase + ix.x]) * b0 + mk3(sc.nrm[m.vbase + ix.y]) * b1 + mk3(sc.nrm[m.vbase + ix.z]) * b2;
    }
    if (m.has_uv) {
      float2 ua = sc.uv[m.vbase + ix.x], ub = sc.uv[m.vbase + ix.y], uc = sc.uv[m.vbase + ix.z];
      r.u = ua.x * b0 + ub.x * b1 + uc.x * b2;
      r.v = ua.y * b0 + ub.y * b1 + uc.y * b2;
    }
    r.mat = m.fmat_base >= 0 ? sc.fmat[m.fmat_base + prim] : m.mat;
    if (m.fidx_base >= 0) r.fidx = sc.fidx[m.fidx_base + prim];
  }
  if (xf) {
    float4 m0 = as_f4(ir[4]), m1 = as_f4(ir[5]), m2 = as_f4(ir[6]);
    r.P = xf_point(m0, m1, m2, hp);
    r.N = normalize(xf_normal(w0, w1, w2, n));
  } else {
    r.P = hp;
    r.N = normalize(n);
  }
  return r;
}

// image_texture.cu:5-7 with the sampler of textures.hpp:177-230: bilinear,
// repeat wrap, texel centres at (i + 0.5) / w, row 0 at v = 0.
PTD v3 tex_image(const DevScene& sc, const DevTexture& t, float u, float v) {
  float x = u * (float)t.w - 0.5f, y = v * (float)t.h - 0.5f;
  float x0 = floorf(x), y0 = floorf(y);
  float fx = x - x0, fy = y - y0;
  long long W = t.w, H = t.h;
  long long xi = (long long)x0, yi = (long long)y0;
  long long i0 = xi % W; if (i0 < 0) i0 += W;
  long long i1 = (xi + 1) % W; if (i1 < 0) i1 += W;
  long long j0 = yi % H; if (j0 < 0) j0 += H;
  long long j1 = (yi + 1) % H; if (j1 < 0) j1 += H;
  const float4* p = sc.texels + t.offset;
  v3 c00 = mk3(p[j0 * W + i0]), c10 = mk3(p[j0 * W + i1]);
  v3 c01 = mk3(p[j1 * W + i0]), c11 = mk3(p[j1 * W + i1]);
  v3 top = c00 * (1.f - fx) + c10 * fx;
  v3 bot = c01 * (1.f - fx) + c11 * fx;
  return top * (1.f - fy) + bot * fy;
}
// noise_texture.cu:13-54 perlin_interp / noise over the texel-pool tables
// (256 gradient vectors, then perm_x, perm_y, perm_z as floats)
PTD float perlin_noise(const float4* tab, v3 p) {
  float u = p.x - floorf(p.x), v = p.y - floorf(p.y), w = p.z - floorf(p.z);
  int i = (int)floorf(p.x), j = (int)floorf(p.y), k = (int)floorf(p.z);
  const float* perm = reinterpret_cast<const float*>(tab + 256);
  float uu = u * u * (3 - 2 * u), vv = v * v * (3 - 2 * v), ww = w * w * (3 - 2 * w);
  float accum = 0;
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++)
      for (int dk = 0; dk < 2; dk++) {
        int h = (int)perm[(i + di) & 255] ^ (int)perm[256 + ((j + dj) & 255)] ^ (int)perm[512 + ((k + dk) & 255)];
        float4 c = tab[h];
        v3 wv = mk3(u - (float)di, v - (float)dj, w - (float)dk);
        accum += ((float)di * uu + (float)(1 - di) * (1 - uu)) * ((float)dj * vv + (float)(1 - dj) * (1 - vv)) *
                 ((float)dk * ww + (float)(1 - dk) * (1 - ww)) * dot(mk3(c), wv);
      }
  return accum;
}
// noise_texture.cu:56-68 turb (7 octaves)
PTD float perlin_turb(const float4* tab, v3 p) {
  float accum = 0, weight = 1.0f;
  v3 tp = p;
  for (int i = 0; i < 7; i++) {
    accum += weight * perlin_noise(tab, tp);
    weight *= 0.5f;
    tp = tp * 2.f;
  }
  return fabsf(accum);
}
// every texture kind but the checker, at (u, v) and world point P
PTD v3 tex_leaf(const DevScene& sc, const DevTexture& t, float u, float v, v3 P) {
  if (t.type == PT_TEX_CONSTANT) return mk3(t.r, t.g, t.b);
  if (t.type == PT_TEX_NOISE) {
    // noise_texture.cu:70-84: the axis switch falls through to the Z case
    const float4* tab = sc.texels + t.offset;
    float sv = psin(t.r * P.z + 5 * perlin_turb(tab, t.r * P));
    return mk3(1.f) * 0.5f * (1 + sv);
  }
  if (t.type == PT_TEX_GRADIENT) {
    // gradient_texture.cu:7-14
    const float4* c = sc.texels + t.offset;
    v3 ud = normalize(P);
    float x = fabsf(ud.x), y = fabsf(ud.y), z = fabsf(ud.z);
    return x * mk3(c[0]) + y * mk3(c[1]) + z * mk3(c[2]);
  }
  return tex_image(sc, t, u, v);
}
// vector.cu:7-12: child `fidx` of a vector texture (children see index 0),
// black when it is out of range
PTD v3 tex_eval(const DevScene& sc, int ti, float u, float v, v3 P = mk3(0.f), int fidx = 0) {
  if (sc.texs[ti].type == PT_TEX_VECTOR) {
    const DevTexture& tv = sc.texs[ti];
    if (fidx < 0 || (uint32_t)fidx >= tv.w) return mk3(0.f);
    ti = (int)reinterpret_cast<const float*>(sc.texels + tv.offset)[fidx];
  }
  const DevTexture& t = sc.texs[ti];
  if (t.type == PT_TEX_CHECKER) {
    // checkered_texture.cu:8-15 (odd in w, even in h)
    float sines = psin(10 * P.x) * psin(10 * P.y) * psin(10 * P.z);
    return tex_leaf(sc, sc.texs[sines < 0 ? t.w : t.h], u, v, P);
  }
  return tex_leaf(sc, t, u, v, P);
}
// Geometric normal of a hit for normal_shader.cu:27-31: differs from the
// hit record's (shading) normal only on meshes with vertex normals.
PTD v3 geometric_normal(const DevScene& sc, int inst_pos, int prim, v3 n_shading) {
  const uint4* ir = sc.insts + (size_t)inst_pos * 8;
  if ((ir[3].z & 0xFFu) != PT_GEOM_MESH) return n_shading;
  const DevMesh& m = sc.meshes[ir[7].z];
  if (!m.has_n) return n_shading;
  uint4 ix = sc.idx[m.ibase + prim];
  v3 a = mk3(sc.vtx[m.vbase + ix.x]), b = mk3(sc.vtx[m.vbase + ix.y]), c = mk3(sc.vtx[m.vbase + ix.z]);
  v3 n = cross(b - a, c - a);
  if ((ir[3].z >> 8) != 0) n = xf_normal(as_f4(ir[0]), as_f4(ir[1]), as_f4(ir[2]), n);
  return normalize(n);
}

PTD bool tex_needs_uv(const DevScene& sc, int ti) {
  if (ti < 0) return false;
  const DevTexture& t = sc.texs[ti];
  if (t.type == PT_TEX_CHECKER) return sc.texs[t.w].type == PT_TEX_IMAGE || sc.texs[t.h].type == PT_TEX_IMAGE;
  return t.type == PT_TEX_IMAGE || t.type == PT_TEX_VECTOR;  // vectors: conservatively
}

PTD Surface make_surface(const DevScene& sc, const HitRec& r) {
  const DevMaterial& m = sc.mats[r.mat];
  Surface f;
  f.type = m.type;
  f.c0 = m.tex0 >= 0 ? tex_eval(sc, m.tex0, r.u, r.v, r.P, r.fidx) : mk3(0.f);  // the normal shader has none
  f.c1 = (m.type == PT_MAT_ASHIKHMIN_SHIRLEY) ? tex_eval(sc, m.tex1, r.u, r.v, r.P, r.fidx) : mk3(0.f);
  f.p0 = m.p0;
  f.p1 = m.p1;
  return f;
}

// ------------------------------------------------------- environment
// Spherical HDR mapping of miss.cu:64-95: u = atan2(z, x) / 2pi,
// v = acos(y / |d|) / pi (u wraps through the texture's repeat mode).
PTD void env_uv(v3 d, float& u, float& v) {
  float r = length(d);
  float lon = patan2(d.z, d.x);
  float lat = pacos(d.y / r);
  u = lon * (1.f / (PI_F * 2.f));
  v = lat * (1.f / PI_F);
}
// radiance the HDR miss program returns toward d (x2, miss.cu:93)
PTD v3 env_radiance(const DevScene& sc, v3 d) {
  float u, v;
  env_uv(d, u, v);
  return 2.f * tex_eval(sc, sc.miss.texture, u, v);
}
// floor(total * r / 2^24) for a 24-bit r, exactly
PTD unsigned long long scale24(unsigned long long total, uint32_t r) {
  return (total >> 24) * r + (((total & 0xFFFFFFull) * r) >> 24);
}
// first index in [0, n) whose inclusive prefix exceeds target
PTD uint32_t upper_index(const unsigned long long* a, uint32_t n, unsigned long long target) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > target) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}
// Importance sampling of the environment (no reference counterpart):
// density per unit solid angle p(u, v) W H / (2 pi^2 sin(theta)), p the
// texel's share of the fixed-point weight total.
PTD float env_pdf(const DevScene& sc, v3 d) {
  if (sc.env_total == 0) return 0.f;
  float u, v;
  env_uv(d, u, v);
  float uw = u - floorf(u);
  long long i = (long long)(uw * (float)sc.env_w), j = (long long)(v * (float)sc.env_h);
  if (i >= (long long)sc.env_w) i = sc.env_w - 1;
  if (j >= (long long)sc.env_h) j = sc.env_h - 1;
  if (i < 0) i = 0;
  if (j < 0) j = 0;
  const unsigned long long* row = sc.env_cond + (size_t)j * sc.env_w;
  unsigned long long w = row[i] - (i ? row[i - 1] : 0ull);
  float st = psin(v * PI_F);
  if (w == 0 || !(st > 0.f)) return 0.f;
  const float p = ((float)w / (float)sc.env_total) * ((float)sc.env_w * (float)sc.env_h);
  return p / ((2.0f * PI_F * PI_F) * st);
}
// two draws: row from the marginal table, column from the row's table,
// continuous offsets inside the texel
// first index of a (n inclusive prefix sums, total > x) whose prefix
// exceeds x, searched between the guide entries around x's bucket (the
// bucket is estimated in float; an estimate off by one still brackets the
// answer, and the exact integer comparisons decide it): the same index as
// upper_index over the whole array.
PTD uint32_t guided_index(const unsigned long long* a, uint32_t n, unsigned long long total, unsigned long long x,
                          const uint32_t* guide) {
  const float est = (float)x * ((float)ENV_GUIDE / (float)total);
  uint32_t b = est >= (float)(ENV_GUIDE - 1) ? ENV_GUIDE - 1 : (uint32_t)est;
  uint32_t lo = guide[b > 0 ? b - 1 : 0], hi = guide[b + 2 < ENV_GUIDE ? b + 2 : ENV_GUIDE];
  if (hi > n - 1) hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}
PTD v3 env_sample(const DevScene& sc, uint32_t& seed) {
  uint32_t r1 = lcg24(seed);
  uint32_t r2 = lcg24(seed);
  if (sc.env_total == 0) return mk3(0.f, 1.f, 0.f);
  unsigned long long tr = scale24(sc.env_total, r1);
  const uint32_t* gm = sc.env_guide + (size_t)sc.env_h * (ENV_GUIDE + 1);
  uint32_t j = guided_index(sc.env_marg, sc.env_h, sc.env_total, tr, gm);
  unsigned long long rprev = j ? sc.env_marg[j - 1] : 0ull;
  unsigned long long rw = sc.env_marg[j] - rprev;
  const unsigned long long* row = sc.env_cond + (size_t)j * sc.env_w;
  unsigned long long tc = scale24(rw, r2);
  uint32_t i = guided_index(row, sc.env_w, rw, tc, sc.env_guide + (size_t)j * (ENV_GUIDE + 1));
  unsigned long long cprev = i ? row[i - 1] : 0ull;
  // offsets inside the texel, below 1 (a ratio of two roundings may reach it)
  const float fu = fminf((float)(tc - cprev) / (float)(row[i] - cprev), 0x1.fffffep-1f);
  const float fv = fminf((float)(tr - rprev) / (float)rw, 0x1.fffffep-1f);
  float u = ((float)i + fu) / (float)sc.env_w;
  float v = ((float)j + fv) / (float)sc.env_h;
  float lon = 2.f * PI_F * u, lat = PI_F * v;
  float sl = psin(lat);
  return mk3(sl * pcos(lon), pcos(lat), sl * psin(lon));
}

// miss.cu:30-95
PTD v3 miss_color(const DevScene& sc, v3 d) {
  const pt_miss_desc& m = sc.miss;
  if (m.type == PT_MISS_GRADIENT) {
    v3 ud = normalize(d);
    float t = 0.5f * (ud.y + 1.f);
    v3 c = (1.f - t) * mk3(m.color1[0], m.color1[1], m.color1[2]);
    c = c + t * mk3(m.color2[0], m.color2[1], m.color2[2]);
    return c;
  }
  if (m.type == PT_MISS_CONSTANT) return mk3(m.color1[0], m.color1[1], m.color1[2]);
  if (m.type == PT_MISS_IMAGE) {
    const float PI = 3.14159265358979323846f, INV_PI = 0.318309886183790671538f;
    float theta = patan2(d.x, d.z);
    float phi = PI * 0.5f - pacos(d.y);
    float u = (theta + PI) * (0.5f * INV_PI);
    float v = 0.5f * (1.f + psin(phi));
    return tex_eval(sc, m.texture, u, v);
  }
  float u, v;
  if (m.spherical) {
    env_uv(d, u, v);
  } else {
    float theta = patan2(d.x, d.z);
    theta = theta < 0.f ? theta + (2.f * PI_F) : theta;
    float phi = pacos(d.y);
    u = 1.f - (theta / (2.f * PI_F));
    v = phi / PI_F;
  }
  return 2.f * tex_eval(sc, m.texture, u, v);
}

// rect_pdf.cu:115-146 and sphere_pdf.cu:53-74
PTD v3 light_sample(const DevLight& l, v3 P, uint32_t& seed) {
  if (l.type == PT_LIGHT_RECT) {
    float ra = rnd(seed);
    float ca = l.a0 + ra * (l.a1 - l.a0);
    float rb = rnd(seed);
    float cb = l.b0 + rb * (l.b1 - l.b0);
    v3 pt = l.axis == 0 ? mk3(l.k, ca, cb) : (l.axis == 1 ? mk3(ca, l.k, cb) : mk3(ca, cb, l.k));
    return pt - P;
  }
  float r1 = rnd(seed);
  float r2 = rnd(seed);
  v3 c = mk3(l.cx, l.cy, l.cz);
  float dsq = squared_length(c - P);
  float z = 1.f + r2 * (sqrtf(1.f - l.r * l.r / dsq) - 1.f);
  float phi = 2.f * PI_F * r1;
  float x = pcos(phi) * sqrtf(1.f - z * z);
  float y = psin(phi) * sqrtf(1.f - z * z);
  v3 Wi = mk3(x, y, z);
  return onb_to_world(normalize(Wi), Wi);
}
// rect_pdf.cu:64-113 and sphere_pdf.cu:7-50
PTD float light_pdf(const DevLight& l, v3 P, v3 Wi) {
  const float FLTMAX = 3.402823466e+38f;
  if (l.type == PT_LIGHT_RECT) {
    float t, a, b;
    v3 rn;
    if (l.axis == 0) { t = (l.k - P.x) / Wi.x; a = P.y + t * Wi.y; b = P.z + t * Wi.z; rn = mk3(1.f, 0.f, 0.f); }
    else if (l.axis == 1) { t = (l.k - P.y) / Wi.y; a = P.x + t * Wi.x; b = P.z + t * Wi.z; rn = mk3(0.f, 1.f, 0.f); }
    else { t = (l.k - P.z) / Wi.z; a = P.x + t * Wi.x; b = P.y + t * Wi.y; rn = mk3(0.f, 0.f, 1.f); }
    if (a < l.a0 || a > l.a1 || b < l.b0 || b > l.b1) return 0.f;
    if (!(t < FLTMAX && t > 0.001f)) return 0.f;
    float dsq = t * t * squared_length(Wi);
    float cosine = fabsf(dot(Wi, rn) / length(Wi));
    float area = (l.a1 - l.a0) * (l.b1 - l.b0);
    return dsq / (cosine * area);
  }
  v3 c = mk3(l.cx, l.cy, l.cz);
  v3 oc = P - c;
  float a = dot(Wi, Wi), b = dot(oc, Wi), cc = dot(oc, oc) - l.r * l.r;
  float disc = b * b - a * cc;
  if (disc < 0.f) return 0.f;
  float tt = (-b - sqrtf(disc)) / a;
  bool hit = (tt < FLTMAX && tt > 0.001f);
  if (!hit) {
    tt = (-b + sqrtf(disc)) / a;
    hit = (tt < FLTMAX && tt > 0.001f);
  }
  if (!hit) return 0.f;
  float dsq = squared_length(c - P);
  float ctm = sqrtf(1.f - l.r * l.r / dsq);
  float sa = 2.f * PI_F * (1.f - ctm);
  return 1.f / sa;
}

PTD v3 light_emission(const DevLight& l) { return mk3(l.ex, l.ey, l.ez); }

// material of a hit (for the shadow-ray rule and the material sort)
PTD int hit_material(const DevScene& sc, int inst_pos, int prim) {
  const uint4* ir = sc.insts + (size_t)inst_pos * 8;
  uint32_t kind = ir[3].z & 0xFFu;
  if (kind == PT_GEOM_SPHERES) return sc.sph_mat[prim];
  if (kind == PT_GEOM_RECTS) return sc.rect_mat[prim];
  if (kind == PT_GEOM_BOXES) return sc.box_mat[prim];
  if (kind == PT_GEOM_VOLUMES) return sc.vol_mat[prim];
  if (kind == PT_GEOM_CYLINDERS) return sc.cyl_mat[prim];
  if (kind == PT_GEOM_MOVING_SPHERES) return sc.mov_mat[prim];
  const DevMesh& m = sc.meshes[ir[7].z];
  return m.fmat_base >= 0 ? sc.fmat[m.fmat_base + prim] : m.mat;
}

}  // namespace pt
