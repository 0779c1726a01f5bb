// ptref_capi.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// C entry points of the CPU oracle (liboracle_ptref.so) for the tests, the
// smoke check and bench.py's cpu_baseline leg.  Nothing in the product links
// or loads this library.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "ptref_pixels.hpp"

using namespace ptref;

namespace {
thread_local std::string g_err;
}

extern "C" {

const char* ptref_last_error(void) { return g_err.c_str(); }

uint32_t ptref_tea(uint32_t v0, uint32_t v1, int32_t rounds) { return tea(v0, v1, rounds); }

// n successive rnd() draws starting from `seed`; returns the final state
uint32_t ptref_rnd_sequence(uint32_t seed, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; i++) out[i] = rnd(seed);
  return seed;
}

void* ptref_scene_create(const pt_scene_desc* desc) {
  try {
    return new Scene(*desc);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void ptref_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

// Sizes of the acceleration structures: out[0] = total nodes (all BLAS then
// TLAS), out[1] = total primitives (all BLAS then TLAS instances),
// out[2] = number of BLAS.
void ptref_accel_sizes(const void* sp, uint64_t* out) {
  const Scene& s = *static_cast<const Scene*>(sp);
  uint64_t nodes = 0, prims = 0;
  for (const auto& b : s.blas) { nodes += b.bvh.num_nodes(); prims += b.bvh.order.size(); }
  nodes += s.tlas.num_nodes();
  prims += s.tlas.order.size();
  out[0] = nodes;
  out[1] = prims;
  out[2] = s.blas.size();
}

// Node words, sorted orders and Morton codes concatenated in the same order
// (every BLAS in first-use order, then the TLAS); blas_info gets 6 words per
// BLAS: node_base, node_count, prim_base, prim_count, kind, first.
void ptref_accel_read(const void* sp, uint32_t* nodes, uint32_t* order, uint32_t* morton, uint32_t* blas_info) {
  const Scene& s = *static_cast<const Scene*>(sp);
  uint64_t nb = 0, pb = 0;
  auto put = [&](const Bvh4& b) {
    std::memcpy(nodes + nb * 32, b.words.data(), b.words.size() * 4);
    std::memcpy(order + pb, b.order.data(), b.order.size() * 4);
    std::memcpy(morton + pb, b.morton.data(), b.morton.size() * 4);
    nb += b.num_nodes();
    pb += b.order.size();
  };
  for (size_t i = 0; i < s.blas.size(); i++) {
    blas_info[6 * i + 0] = (uint32_t)nb;
    blas_info[6 * i + 1] = s.blas[i].bvh.num_nodes();
    blas_info[6 * i + 2] = (uint32_t)pb;
    blas_info[6 * i + 3] = (uint32_t)s.blas[i].bvh.order.size();
    blas_info[6 * i + 4] = (uint32_t)s.blas[i].kind;
    blas_info[6 * i + 5] = s.blas[i].first;
    put(s.blas[i].bvh);
  }
  put(s.tlas);
}

// Closest hit through the BVH (mode 0) or by brute force (mode 1).
// counts (optional, mode 0): 3 uint32 per ray (nodes, prims, instances).
void ptref_trace(const void* sp, const pt_ray* rays, uint64_t n, pt_hit* hits, uint32_t* counts, int32_t mode,
                 int32_t nthreads) {
  const Scene& s = *static_cast<const Scene*>(sp);
  if (nthreads < 1) nthreads = 1;
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    while (true) {
      uint64_t b = next.fetch_add(4096);
      if (b >= n) break;
      uint64_t e = std::min<uint64_t>(n, b + 4096);
      for (uint64_t i = b; i < e; i++) {
        const pt_ray& r = rays[i];
        f3 o = mk(r.org[0], r.org[1], r.org[2]), d = mk(r.dir[0], r.dir[1], r.dir[2]);
        Counts c;
        // volumes draw their free flight from the ray's index
        HitState h = mode == 1 ? s.brute(o, d, r.tmin, r.tmax, query_vseed(i))
                               : s.closest(o, d, r.tmin, r.tmax, &c, query_vseed(i));
        hits[i].t = h.found() ? h.t : INF;
        hits[i].inst = h.inst;
        hits[i].prim = h.prim;
        hits[i].reserved = 0;
        if (counts) {
          counts[3 * i + 0] = (uint32_t)c.nodes;
          counts[3 * i + 1] = (uint32_t)c.prims;
          counts[3 * i + 2] = (uint32_t)c.insts;
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; t++) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
}

// Shadow-ray rule: occluded[i] = 1 unless the closest hit is a diffuse light
// or there is no hit.
void ptref_occluded(const void* sp, const pt_ray* rays, uint64_t n, uint8_t* occ) {
  const Scene& s = *static_cast<const Scene*>(sp);
  for (uint64_t i = 0; i < n; i++) {
    const pt_ray& r = rays[i];
    HitState h = s.closest(mk(r.org[0], r.org[1], r.org[2]), mk(r.dir[0], r.dir[1], r.dir[2]), r.tmin, r.tmax,
                           nullptr, query_vseed(i));
    occ[i] = h.found() ? (s.mats[s.material_of(h.inst, h.prim)].type == PT_MAT_DIFFUSE_LIGHT ? 0 : 1) : 0;
  }
}

// Render frames [frame_begin, frame_begin + frame_count) of pixels
// [pix_begin, pix_end) (launch index x + W*y) into accum (W*H*3 doubles,
// row 0 = top).  Optional last-frame outputs are indexed by launch index.
// stats: [0] paths, [1] extension rays, [2] shadow rays, [3] paths whose
// color has a NaN channel.
int32_t ptref_render(const void* sp, const pt_camera* cam, const pt_render_params* rp, uint32_t frame_begin,
                     uint32_t frame_count, uint64_t pix_begin, uint64_t pix_end, int32_t nthreads, double* accum,
                     float* last_rgb, uint32_t* last_sig, uint32_t* last_bounces, uint64_t* stats) {
  try {
    const Scene& s = *static_cast<const Scene*>(sp);
    Params p{rp->width, rp->height, rp->max_depth, rp->russian != 0, rp->rr_min_depth,
             (rp->flags & PT_FLAG_ENV_MIS) != 0};
    uint64_t npix = (uint64_t)p.W * p.H;
    if (pix_end > npix) pix_end = npix;
    if (nthreads < 1) nthreads = 1;
    std::atomic<uint64_t> ext{0}, shd{0}, paths{0}, nans{0};
    for (uint32_t f = frame_begin; f < frame_begin + frame_count; f++) {
      std::atomic<uint64_t> next{pix_begin};
      bool last = (f + 1 == frame_begin + frame_count);
      auto work = [&]() {
        PathCounters pc;
        uint64_t np = 0, nn = 0;
        while (true) {
          uint64_t b = next.fetch_add(256);
          if (b >= pix_end) break;
          uint64_t e = std::min<uint64_t>(pix_end, b + 256);
          for (uint64_t i = b; i < e; i++) {
            uint32_t x = (uint32_t)(i % p.W), y = (uint32_t)(i / p.W);
            PathResult r = trace_path(s, *cam, p, x, y, f, pc);
            np++;
            nn += (r.L.x == r.L.x && r.L.y == r.L.y && r.L.z == r.L.z) ? 0 : 1;
            f3 c = de_nan(r.L);
            size_t row = (size_t)(p.H - 1 - y);
            double* a = &accum[3 * (row * p.W + x)];
            a[0] += (double)c.x;
            a[1] += (double)c.y;
            a[2] += (double)c.z;
            if (last) {
              if (last_rgb) { last_rgb[3 * i] = r.L.x; last_rgb[3 * i + 1] = r.L.y; last_rgb[3 * i + 2] = r.L.z; }
              if (last_sig) last_sig[i] = r.sig;
              if (last_bounces) last_bounces[i] = r.bounces;
            }
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
    }
    if (stats) {
      stats[0] = paths;
      stats[1] = ext;
      stats[2] = shd;
      stats[3] = nans;
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// The listed pixels (launch index) over frames [frame_begin, +frame_count):
// sum[3k..3k+2] += sum of de_nan(color()) of pixel pix[k].
int32_t ptref_render_pixels(const void* sp, const pt_camera* cam, const pt_render_params* rp, uint32_t frame_begin,
                            uint32_t frame_count, const uint32_t* pix, uint64_t npix, int32_t nthreads, double* sum,
                            uint64_t* stats) {
  try {
    render_pixel_list(*static_cast<const Scene*>(sp), *cam, *rp, frame_begin, frame_count, pix, npix, nthreads, sum,
                      stats);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Display mapping of an accumulation (raygen.cu:136-146): rgba W*H*4.
void ptref_display(const double* accum, uint64_t npix, uint32_t frames, uint8_t* rgba) {
  for (uint64_t i = 0; i < npix; i++) {
    for (int c = 0; c < 3; c++) rgba[4 * i + c] = display_channel(accum[3 * i + c], frames);
    rgba[4 * i + 3] = 255;
  }
}

// Single-sample helpers for unit tests of the shading functions.
// brdf: surface (type, c0[3], c1[3], p0, p1), Wo, N, seed -> Wi, f, pdf.
void ptref_brdf_sample_eval(int32_t type, const float* c0, const float* c1, float p0, float p1, const float* wo,
                            const float* n, uint32_t* seed, float* wi, float* f, float* pdf) {
  Surface s{type, mk(c0[0], c0[1], c0[2]), mk(c1[0], c1[1], c1[2]), p0, p1};
  f3 Wo = mk(wo[0], wo[1], wo[2]), N = mk(n[0], n[1], n[2]);
  f3 Wi = brdf_sample(s, Wo, N, *seed);
  float pd = 1.f;
  f3 fv = brdf_eval(s, Wo, Wi, N, pd);
  wi[0] = Wi.x; wi[1] = Wi.y; wi[2] = Wi.z;
  f[0] = fv.x; f[1] = fv.y; f[2] = fv.z;
  *pdf = pd;
}
void ptref_brdf_eval(int32_t type, const float* c0, const float* c1, float p0, float p1, const float* wo,
                     const float* wi, const float* n, float* f, float* pdf) {
  Surface s{type, mk(c0[0], c0[1], c0[2]), mk(c1[0], c1[1], c1[2]), p0, p1};
  float pd = 1.f;
  f3 fv = brdf_eval(s, mk(wo[0], wo[1], wo[2]), mk(wi[0], wi[1], wi[2]), mk(n[0], n[1], n[2]), pd);
  f[0] = fv.x; f[1] = fv.y; f[2] = fv.z;
  *pdf = pd;
}
// `count` samples of brdf_sample + brdf_eval, sample k seeded with
// tea(k, seed0, 8); u0[k] = the first rnd() that sample draws
void ptref_brdf_sample_many(int32_t type, const float* c0, const float* c1, float p0, float p1, const float* wo,
                            const float* n, uint32_t seed0, uint64_t count, float* wi, float* f, float* pdf,
                            float* u0) {
  Surface s{type, mk(c0[0], c0[1], c0[2]), mk(c1[0], c1[1], c1[2]), p0, p1};
  f3 Wo = mk(wo[0], wo[1], wo[2]), N = mk(n[0], n[1], n[2]);
  for (uint64_t k = 0; k < count; k++) {
    uint32_t seed = tea((uint32_t)k, seed0, 8);
    uint32_t peek = seed;
    u0[k] = rnd(peek);
    f3 Wi = brdf_sample(s, Wo, N, seed);
    float pd = 1.f;
    f3 fv = brdf_eval(s, Wo, Wi, N, pd);
    wi[3 * k] = Wi.x; wi[3 * k + 1] = Wi.y; wi[3 * k + 2] = Wi.z;
    f[3 * k] = fv.x; f[3 * k + 1] = fv.y; f[3 * k + 2] = fv.z;
    pdf[k] = pd;
  }
}
// `count` metal scatter directions (seeds as above)
void ptref_metal_scatter_many(const float* wo, const float* n, float fuzz, uint32_t seed0, uint64_t count,
                              float* dir) {
  f3 Wo = mk(wo[0], wo[1], wo[2]), N = mk(n[0], n[1], n[2]);
  for (uint64_t k = 0; k < count; k++) {
    uint32_t seed = tea((uint32_t)k, seed0, 8);
    f3 d = metal_scatter(Wo, N, fuzz, seed);
    dir[3 * k] = d.x; dir[3 * k + 1] = d.y; dir[3 * k + 2] = d.z;
  }
}
// cosine hemisphere sample about n: out = world direction
void ptref_cosine_sample(float u1, float u2, const float* n, float* out) {
  f3 w = Onb(mk(n[0], n[1], n[2])).to_world(cosine_sample_hemisphere(u1, u2));
  out[0] = w.x; out[1] = w.y; out[2] = w.z;
}
// light sample + pdf for light index `li` of the scene (environment lights
// through the importance-sampling tables)
void ptref_light_sample(const void* sp, int32_t li, const float* P, uint32_t* seed, float* wi, float* pdf) {
  const Scene& s = *static_cast<const Scene*>(sp);
  f3 p = mk(P[0], P[1], P[2]);
  const Light& L = s.lights[li];
  f3 w = L.type == PT_LIGHT_ENV ? env_sample(s, *seed) : light_sample(L, p, *seed);
  wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
  *pdf = L.type == PT_LIGHT_ENV ? env_pdf(s, w) : light_pdf(L, p, w);
}
float ptref_light_pdf(const void* sp, int32_t li, const float* P, const float* wi) {
  const Scene& s = *static_cast<const Scene*>(sp);
  const Light& L = s.lights[li];
  f3 w = mk(wi[0], wi[1], wi[2]);
  return L.type == PT_LIGHT_ENV ? env_pdf(s, w) : light_pdf(L, mk(P[0], P[1], P[2]), w);
}
// environment tables: out[0] = W, out[1] = H, out[2] = total; cond (H*W)
// and marg (H) copied when non-null
void ptref_env_tables(const void* sp, uint64_t* out, uint64_t* cond, uint64_t* marg) {
  const Scene& s = *static_cast<const Scene*>(sp);
  out[0] = s.env.W;
  out[1] = s.env.H;
  out[2] = s.env.total;
  if (cond && !s.env.cond.empty()) std::memcpy(cond, s.env.cond.data(), s.env.cond.size() * 8);
  if (marg && !s.env.marg.empty()) std::memcpy(marg, s.env.marg.data(), s.env.marg.size() * 8);
}
// radiance of the environment (the HDR miss program) toward d
void ptref_env_radiance(const void* sp, const float* d, float* out) {
  const Scene& s = *static_cast<const Scene*>(sp);
  f3 c = env_radiance(s, mk(d[0], d[1], d[2]));
  out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
void ptref_miss(const void* sp, const float* d, float* out) {
  const Scene& s = *static_cast<const Scene*>(sp);
  f3 c = miss_color(s, mk(d[0], d[1], d[2]));
  out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
// texture `tex` at (u, v) and world point P
void ptref_texture(const void* sp, int32_t tex, float u, float v, const float* P, int32_t fidx, float* out) {
  const Scene& s = *static_cast<const Scene*>(sp);
  f3 c = tex_eval(s, tex, u, v, mk(P[0], P[1], P[2]), fidx);
  out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
// hit record of a given hit: out = P[3], N[3], Wo[3], u, v, material
void ptref_hit_record(const void* sp, int32_t inst, int32_t prim, float t, const float* o, const float* d,
                      float* out, int32_t* mat) {
  const Scene& s = *static_cast<const Scene*>(sp);
  HitRec r = hit_record(s, inst, prim, t, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]));
  float v[11] = {r.P.x, r.P.y, r.P.z, r.N.x, r.N.y, r.N.z, r.Wo.x, r.Wo.y, r.Wo.z, r.u, r.v};
  std::memcpy(out, v, sizeof(v));
  *mat = r.mat;
}

// elementary functions of ptref_libm.hpp on arrays: fn 0 sin, 1 cos, 2 tan,
// 3 atan, 4 atan2(x, y), 5 asin, 6 acos, 7 pow(x, y), 8 pow5
void ptref_libm(int32_t fn, const float* x, const float* y, float* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    float a = x[i], b = y ? y[i] : 0.f;
    switch (fn) {
      case 0: out[i] = psin(a); break;
      case 1: out[i] = pcos(a); break;
      case 2: out[i] = ptan(a); break;
      case 3: out[i] = patan(a); break;
      case 4: out[i] = patan2(a, b); break;
      case 5: out[i] = pasin(a); break;
      case 6: out[i] = pacos(a); break;
      case 7: out[i] = ppow(a, b); break;
      default: out[i] = ppow5(a); break;
    }
  }
}

// material.cuh:30-39 schlick
float ptref_schlick(float cosine, float ior) { return schlick(cosine, ior); }

// material.cuh:50-63 Refract; returns 1 and writes the refracted direction
// when the discriminant is positive, else 0
int32_t ptref_refract(const float* v, const float* n, float ni_over_nt, float* out) {
  f3 r = mk(0.f);
  bool ok = Refract(mk(v[0], v[1], v[2]), mk(n[0], n[1], n[2]), ni_over_nt, r);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  return ok ? 1 : 0;
}

// microfacets.cuh:85-99 GGX_D (half vector against world +Y)
float ptref_ggx_d(const float* h, float nu, float nv) { return GGX_D(mk(h[0], h[1], h[2]), nu, nv); }

// the tangent frame around a unit normal: rows t, b, n
void ptref_onb(const float* n, float* out) {
  Onb o(mk(n[0], n[1], n[2]));
  float v[9] = {o.t.x, o.t.y, o.t.z, o.b.x, o.b.y, o.b.z, o.n.x, o.n.y, o.n.z};
  std::memcpy(out, v, sizeof(v));
}

// light_sample.cuh:27-33 PowerHeuristic(1, f, 1, g)
float ptref_power_heuristic(float f, float g) { return PowerHeuristic(1, f, 1, g); }

// random.cuh-based samplers: 0 unit disk, 1 in unit sphere, 2 on unit sphere
uint32_t ptref_random_points(int32_t kind, uint32_t seed, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; i++) {
    f3 p = kind == 0 ? random_in_unit_disk(seed) : (kind == 1 ? random_in_unit_sphere(seed) : random_on_unit_sphere(seed));
    out[3 * i] = p.x;
    out[3 * i + 1] = p.y;
    out[3 * i + 2] = p.z;
  }
  return seed;
}

}  // extern "C"
