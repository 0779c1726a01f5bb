// pt_render.hip -- gfx950 kernels of the per-sample render path: query
// kernels behind pt_trace_closest / pt_trace_occluded and the wavefront
// integrator that replaces the OptiX megakernel (programs/raygen.cu:148-168
// renderPixel and :66-123 color, closest-hit programs of
// programs/materials/*.cu, miss programs, shadow any-hit).
//
// Wavefront, one bounce at a time over SoA path state in HBM:
//   raygen -> [extend -> (material sort) -> shade -> shadow+connect]*
//   -> accumulate
// Queues are slot lists compacted with wave64 ballots: one atomic per wave.
// All kernels are grid-stride loops that read their item count from device
// memory, so a whole batch is enqueued without host synchronisation.
#include "pt_kernels.hpp"
#include "pt_shade.hpp"

namespace pt {

// -------------------------------------------------------------- helpers
PTD uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// Wave-aggregated append: the lanes with `pred` get consecutive indices from
// one atomic on `counter` (usable in divergent code: only active lanes vote).
PTD uint32_t append_index(bool pred, uint32_t* counter) {
  uint64_t m = __ballot(pred);
  if (!pred) return 0u;
  int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + lane_rank(m);
}

// 24-bit hash of (launch index, frame) choosing radiance-cache training
// paths; independent of the path's own random sequence
PTD uint32_t train_hash(uint32_t pix, uint32_t frame) {
  uint32_t h = pix * 0x9E3779B1u ^ (frame + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return h >> 8;
}

// Radiance-cache input features, 96 BF16 (83 used, the rest zero):
//   [0, 72)   sin, cos (2^f pi p), f = 0..11, per axis of the position p
//             normalised to the scene box (angle doubling from pi p)
//   [72, 74)  view direction Wo as (theta / pi, phi / 2 pi + 1/2), theta from +Y
//   [74, 76)  shading normal, same mapping
//   76        roughness, [77, 80) diffuse albedo, [80, 83) specular albedo
PTD uint32_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}
struct FeatWriter {
  uint4* dst;
  float buf[8];
  int n = 0, k = 0;
  PTD void put(float v) {
    buf[n++] = v;
    if (n == 8) {
      dst[k++] = make_uint4(bf16_bits(buf[0]) | (bf16_bits(buf[1]) << 16), bf16_bits(buf[2]) | (bf16_bits(buf[3]) << 16),
                            bf16_bits(buf[4]) | (bf16_bits(buf[5]) << 16), bf16_bits(buf[6]) | (bf16_bits(buf[7]) << 16));
      n = 0;
    }
  }
};
PTD void sph_dir(v3 w, float& a, float& b) {
  a = pacos(fminf(fmaxf(w.y, -1.f), 1.f)) * (1.f / PI_F);
  b = patan2(w.z, w.x) * (0.5f / PI_F) + 0.5f;
}
__device__ __attribute__((noinline)) void write_features(uint4* dst, const CacheConsts& cc, v3 P, v3 Wo, v3 N,
                                                         const Surface& f) {
  FeatWriter fw{dst};
  const float pp[3] = {(P.x - cc.lo[0]) * cc.inv_ext[0], (P.y - cc.lo[1]) * cc.inv_ext[1],
                       (P.z - cc.lo[2]) * cc.inv_ext[2]};
  for (int a = 0; a < 3; a++) {
    float sn = psin(PI_F * pp[a]), cs = pcos(PI_F * pp[a]);
    for (int q = 0; q < 12; q++) {
      fw.put(sn);
      fw.put(cs);
      float s2 = 2.f * sn * cs, c2 = cs * cs - sn * sn;
      sn = s2;
      cs = c2;
    }
  }
  float a, b;
  sph_dir(normalize(Wo), a, b);
  fw.put(a);
  fw.put(b);
  sph_dir(normalize(N), a, b);
  fw.put(a);
  fw.put(b);
  float rough = 1.f;
  v3 rd = mk3(0.f), rs = mk3(0.f);
  switch (f.type) {
    case PT_MAT_LAMBERTIAN:
    case PT_MAT_OREN_NAYAR: rd = f.c0; break;
    case PT_MAT_METAL: rough = f.p0; rs = f.c0; break;
    case PT_MAT_DIELECTRIC: rough = 0.f; rs = mk3(1.f); break;
    case PT_MAT_ASHIKHMIN_SHIRLEY: rough = 1.f / sqrtf(1.f + 0.5f * (f.p0 + f.p1)); rd = f.c0; rs = f.c1; break;
    case PT_MAT_TORRANCE_SPARROW: rough = sqrtf(f.p0); rs = f.c0; break;
    default: break;
  }
  fw.put(rough);
  fw.put(rd.x);
  fw.put(rd.y);
  fw.put(rd.z);
  fw.put(rs.x);
  fw.put(rs.y);
  fw.put(rs.z);
  while (fw.k < 12) fw.put(0.f);
}

PTD Stack make_stack(uint32_t* s_ref, float* s_tn, uint2* spill, uint32_t* overflow) {
  Stack st;
  st.ref = s_ref + threadIdx.x;
  st.tn = s_tn + threadIdx.x;
  st.spill = spill + (size_t)(blockIdx.x * BLOCK + threadIdx.x) * STK_SPILL;
  st.sp = 0;
  st.overflow = overflow;
  return st;
}

// ------------------------------------------------------- query kernels
template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace_closest(DevScene sc, const pt_ray* __restrict__ rays, uint64_t n,
                                                         pt_hit* __restrict__ hits, uint32_t* __restrict__ counts,
                                                         uint2* spill, uint32_t* overflow) {
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  Stack st = make_stack(s_ref, s_tn, spill, overflow);
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    const float4* r4 = reinterpret_cast<const float4*>(rays + i);
    float4 a = r4[0], b = r4[1];
    TraceCounts c{0, 0, 0};
    HitOut h = trace_closest<COUNT>(sc, mk3(a), mk3(b), a.w, b.w, st, c);
    pt_hit out;
    out.t = h.inst >= 0 ? h.t : FINF;
    out.inst = h.inst;
    out.prim = h.prim;
    out.reserved = 0;
    hits[i] = out;
    if (COUNT) {
      counts[3 * i + 0] = c.nodes;
      counts[3 * i + 1] = c.prims;
      counts[3 * i + 2] = c.insts;
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_trace_occluded(DevScene sc, const pt_ray* __restrict__ rays, uint64_t n,
                                                          uint8_t* __restrict__ occ, uint2* spill,
                                                          uint32_t* overflow) {
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  Stack st = make_stack(s_ref, s_tn, spill, overflow);
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    const float4* r4 = reinterpret_cast<const float4*>(rays + i);
    float4 a = r4[0], b = r4[1];
    TraceCounts c{0, 0, 0};
    HitOut h = trace_closest<false>(sc, mk3(a), mk3(b), a.w, b.w, st, c);
    uint8_t o = 0;
    if (h.inst >= 0) o = sc.mats[hit_material(sc, h.pos, h.prim)].type == PT_MAT_DIFFUSE_LIGHT ? 0 : 1;
    occ[i] = o;
  }
}

// ------------------------------------------------------ wavefront kernels
// raygen.cu:148-168 renderPixel up to the first rtTrace, raygen.cu:50-64
// generateRay, raygen.cu:66-74 color() set-up.
__global__ __launch_bounds__(BLOCK) void k_raygen(DevPaths P, RenderConsts rc, uint32_t nslots) {
  PathState S = P.st[0];
  if (rc.frame_dev) rc.frame0 = *rc.frame_dev;  // replayed graphs take the batch's frame from memory
  for (uint32_t s = blockIdx.x * BLOCK + threadIdx.x; s < nslots; s += gridDim.x * BLOCK) {
    uint32_t f = s / rc.WH, pix = s - f * rc.WH;
    uint32_t x = pix % rc.W, y = pix / rc.W;
    uint32_t seed = tea64(rc.W * y + x, rc.frame0 + f);
    float u = ((float)x + rnd(seed)) / (float)rc.W;
    float v = ((float)y + rnd(seed)) / (float)rc.H;
    const DevCamera& c = rc.cam;
    v3 cu = mk3(c.u[0], c.u[1], c.u[2]), cv = mk3(c.v[0], c.v[1], c.v[2]);
    v3 rd = c.lens_radius * random_in_unit_disk(seed);
    v3 lens = cu * rd.x + cv * rd.y;
    v3 origin = mk3(c.org[0], c.org[1], c.org[2]) + lens;
    v3 dir = mk3(c.llc[0], c.llc[1], c.llc[2]) + u * mk3(c.hor[0], c.hor[1], c.hor[2]) +
             v * mk3(c.ver[0], c.ver[1], c.ver[2]) - origin;
    (void)rnd(seed);  // prd.time (raygen.cu:69), consumed but unused
    S.o[s] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(seed));
    uint32_t dw = s;
    if ((rc.flags & PT_FLAG_RADIANCE_CACHE) && train_hash(pix, rc.frame0 + f) < rc.cache.train_threshold) dw |= TRAIN_BIT;
    S.d[s] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(dw));
    S.T[s] = make_float4(1.f, 1.f, 1.f, 1.f);
    S.L[s] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0x811C9DC5u));
    P.result[s] = make_float4(0.f, 0.f, 0.f, __uint_as_float(rc.max_depth));
  }
  if (blockIdx.x == 0 && threadIdx.x < CTR_STACK_OVERFLOW) P.counters[threadIdx.x] = threadIdx.x == 0 ? nslots : 0u;
  if (blockIdx.x == 0 && (threadIdx.x == CTR_QUERY || threadIdx.x == CTR_TRAIN)) P.counters[threadIdx.x] = 0u;
}

// material bucket of a hit: 0 = miss, 1 + material type otherwise
PTD uint32_t material_key(const DevScene& sc, float4 h) {
  int pos = __float_as_int(h.y);
  return pos < 0 ? 0u : 1u + (uint32_t)sc.mats[hit_material(sc, pos, __float_as_int(h.z))].type;
}

// Closest-hit query of every live path (raygen.cu:77).  With SORT the
// material bucket of each hit is counted (per wave by ballot, per block in
// LDS, one global atomic per bucket and block).
template <bool COUNT, bool SORT>
__global__ __launch_bounds__(BLOCK) void k_extend(DevScene sc, DevPaths P, int in, float tmin, uint2* spill,
                                                  uint64_t* stats) {
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  __shared__ uint32_t bcnt[8];
  Stack st = make_stack(s_ref, s_tn, spill, P.counters + CTR_STACK_OVERFLOW);
  const uint32_t n = P.counters[CTR_ACTIVE + in];
  const PathState S = P.st[in];
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&stats[STAT_EXT], (unsigned long long)n);
  if (SORT) {
    if (threadIdx.x < 8) bcnt[threadIdx.x] = 0;
    __syncthreads();
  }
  TraceCounts c{0, 0, 0};
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    uint32_t key = 0xFFu;
    if (i < n) {
      float4 o = S.o[i], d = S.d[i];
      HitOut h = trace_closest<COUNT>(sc, mk3(o), mk3(d), tmin, FINF, st, c);
      float4 hv = make_float4(h.t, __int_as_float(h.pos), __int_as_float(h.prim), __int_as_float(h.inst));
      P.hit[i] = hv;
      if (SORT) {
        key = material_key(sc, hv);
        P.keys[i] = key;
      }
    }
    if (SORT) {
#pragma unroll
      for (uint32_t b = 0; b < 8; b++) {
        uint64_t m = __ballot(key == b);
        if (m && key == b && lane_rank(m) == 0) atomicAdd(&bcnt[b], (uint32_t)__popcll(m));
      }
    }
  }
  if (SORT) {
    __syncthreads();
    if (threadIdx.x < 8 && bcnt[threadIdx.x]) atomicAdd(&P.counters[CTR_BUCKETS + threadIdx.x], bcnt[threadIdx.x]);
  }
  if (COUNT) {
    uint32_t a = c.nodes, b = c.prims, d = c.insts;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      d += __shfl_down(d, off);
    }
    if (__lane_id() == 0) {
      atomicAdd((unsigned long long*)&stats[STAT_NODES], (unsigned long long)a);
      atomicAdd((unsigned long long*)&stats[STAT_PRIMS], (unsigned long long)b);
      atomicAdd((unsigned long long*)&stats[STAT_INSTS], (unsigned long long)d);
    }
  }
}

__global__ void k_sort_offsets(DevPaths P) {
  uint32_t acc = 0;
  for (int b = 0; b < 8; b++) {
    P.counters[CTR_BUCKET_CURSOR + b] = acc;
    acc += P.counters[CTR_BUCKETS + b];
  }
}
// Move path state and hit into material-bucket order (state set 2); cursors
// are claimed per wave (ballot) and per block (LDS).
__global__ __launch_bounds__(BLOCK) void k_sort_scatter(DevPaths P, int in) {
  __shared__ uint32_t cnt[8], base[8];
  const uint32_t n = P.counters[CTR_ACTIVE + in];
  const PathState S = P.st[in], D = P.st[2];
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    uint32_t key = i < n ? P.keys[i] : 0xFFu;
    if (threadIdx.x < 8) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t wave_off = 0;
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t b = 0; b < 8; b++) {
      uint64_t m = __ballot(key == b);
      uint32_t o = 0;
      if (m && key == b && lane_rank(m) == 0) o = atomicAdd(&cnt[b], (uint32_t)__popcll(m));
      if (key == b) mine = m;
      o = __shfl(o, m ? (int)(__ffsll((unsigned long long)m) - 1) : 0);
      if (key == b) wave_off = o;
    }
    __syncthreads();
    if (threadIdx.x < 8)
      base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&P.counters[CTR_BUCKET_CURSOR + threadIdx.x], cnt[threadIdx.x]) : 0u;
    __syncthreads();
    if (key < 8) {
      uint32_t j = base[key] + wave_off + lane_rank(mine);
      D.o[j] = S.o[i];
      D.d[j] = S.d[i];
      D.T[j] = S.T[i];
      D.L[j] = S.L[i];
      P.hit_sorted[j] = P.hit[i];
    }
    __syncthreads();
  }
}

// Scatter of the materials that do not sample lights plus the continuation
// sample of Lambertian / Oren-Nayar, then Russian roulette (raygen.cu:96-118).
// Returns false when the path ended (result written).
template <int MS>
PTD bool scatter_and_roulette(const Surface& f, v3 Wo, v3 N, uint32_t& seed, uint32_t& sig, v3& T, v3 L, v3& dir,
                              uint32_t& flags, float& prev_pdf, uint32_t depth, const RenderConsts& rc,
                              float4& result) {
  bool spec = false;
  if ((has_mat<MS>(PT_MAT_LAMBERTIAN) || has_mat<MS>(PT_MAT_OREN_NAYAR)) &&
      (f.type == PT_MAT_LAMBERTIAN || f.type == PT_MAT_OREN_NAYAR)) {
    // lambertian.cu:62-72, oren_nayar.cu:59-69
    v3 Wi = brdf_sample<MS>(f, Wo, N, seed);
    float pdf = 1.f;
    v3 att = brdf_eval<MS>(f, Wo, Wi, N, pdf);
    dir = Wi;
    T = T * clamp3(att / pdf, 0.f, 1.f);
    prev_pdf = pdf;
  } else if (has_mat<MS>(PT_MAT_METAL) && f.type == PT_MAT_METAL) {  // metal.cu:38-56
    v3 refl = reflect(-Wo, N);
    dir = refl + f.p0 * random_in_unit_sphere(seed);
    T = T * f.c0;
    spec = true;
  } else if (has_mat<MS>(PT_MAT_DIELECTRIC) && f.type == PT_MAT_DIELECTRIC) {  // dielectric.cu:46-96
    float ior = f.p0;
    v3 W = -Wo;
    v3 Nd = N;
    float ni_over_nt;
    float cosine = dot(W, Nd);
    if (cosine > 0.f) {
      Nd = -Nd;
      ni_over_nt = ior;
      cosine = ior * cosine / length(W);
    } else {
      ni_over_nt = 1.f / ior;
      cosine = -cosine / length(W);
    }
    v3 refracted = mk3(0.f);
    float reflect_prob;
    if (Refract(W, Nd, ni_over_nt, refracted)) reflect_prob = schlick(cosine, ior);
    else reflect_prob = 1.f;
    bool refl = rnd(seed) < reflect_prob;
    sig = sig_mix(sig, refl ? 0x11111111u : 0x22222222u);
    dir = refl ? reflect(W, Nd) : normalize(refracted);
    T = T * f.c0;
    spec = true;
  } else if (has_mat<MS>(PT_MAT_ASHIKHMIN_SHIRLEY) && f.type == PT_MAT_ASHIKHMIN_SHIRLEY) {
    // ashikhmin_shirley.cu:48-68
    v3 Wi = brdf_sample<MS>(f, Wo, N, seed);
    float pdf = 1.f;
    v3 att = brdf_eval<MS>(f, Wo, Wi, N, pdf);
    dir = Wi;
    T = T * clamp3(att, 0.f, 1.f);
    spec = true;
  } else if (has_mat<MS>(PT_MAT_TORRANCE_SPARROW)) {  // Torrance-Sparrow, torrance_sparrow.cu:39-59
    v3 Wi = brdf_sample<MS>(f, Wo, N, seed);
    float pdf = 1.f;
    v3 att = brdf_eval<MS>(f, Wo, Wi, N, pdf);
    dir = Wi;
    T = T * clamp3(att / pdf, 0.f, 1.f);
    spec = true;
  }
  flags = spec ? 1u : 0u;
  if (rc.russian) {
    float prob = max_component(T);
    if (depth >= rc.rr_min_depth) {
      bool term = rnd(seed) >= prob;
      sig = sig_mix(sig, term ? 0x44444444u : 0x55555555u);
      if (term) {
        v3 r = L + T;
        result = make_float4(r.x, r.y, r.z, __uint_as_float(depth + 1));
        return false;
      }
      T = T * (1.f / prob);
    }
  }
  return true;
}

// Closest-hit / miss programs of every live path for one bounce: hit record,
// Direct_Light with its shadow ray traced here (light_sample.cuh:35-103;
// lit iff the closest hit in (1e-3, inf) is a diffuse light with
// dot(N, Wi) > 0, or nothing is hit), the emitter rules, the continuation
// sample and Russian roulette.  Surviving paths are appended, wave by wave,
// to state set `out`; finished paths write their slot's result.
// FEAT: SH_COUNT counts the shadow rays' traversal, SH_CACHE records
// radiance-cache training vertices and queries, SH_ENV samples environment
// lights; each is compiled only into the variants that use it.
constexpr int SH_COUNT = 1, SH_CACHE = 2, SH_ENV = 4;
template <int MS, int FEAT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_shade(DevScene sc, DevPaths P, int in, int src, const float4* __restrict__ hits,
                                                 int out, uint32_t depth, RenderConsts rc, uint2* spill,
                                                 uint64_t* stats) {
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  Stack st = make_stack(s_ref, s_tn, spill, P.counters + CTR_STACK_OVERFLOW);
  const uint32_t n = P.counters[CTR_ACTIVE + in];
  const PathState S = P.st[src], D = P.st[out];
  uint32_t* out_count = P.counters + CTR_ACTIVE + out;
  // light table: environment lights take part only with PT_FLAG_ENV_MIS
  constexpr bool COUNT = (FEAT & SH_COUNT) != 0;
  const bool env_mis = (FEAT & SH_ENV) && (rc.flags & PT_FLAG_ENV_MIS) != 0;
  const DevLight* lights = env_mis ? sc.lights : sc.lights_plain;
  const int nl = env_mis ? sc.num_lights : sc.num_lights_plain;
  const bool env_light = env_mis && sc.env_total != 0 && sc.num_lights != sc.num_lights_plain;
  const bool cache_on = (FEAT & SH_CACHE) && (rc.flags & PT_FLAG_RADIANCE_CACHE) != 0;
  TraceCounts tc{0, 0, 0};
  uint32_t shadow_rays = 0;
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    bool alive = false;
    float4 no = make_float4(0.f, 0.f, 0.f, 0.f), nd = no, nT = no, nL = no;
    if (i < n) {
      float4 hit = hits[i];
      float4 o4 = S.o[i], d4 = S.d[i], t4 = S.T[i], l4 = S.L[i];
      v3 d = mk3(d4), T = mk3(t4), L = mk3(l4);
      uint32_t seed = __float_as_uint(o4.w), dw = __float_as_uint(d4.w);
      // dw keeps slot | TRAIN_BIT | SPEC_BIT; fields are taken where used
      uint32_t flags = dw >> 31, sig = __float_as_uint(l4.w);
      float prev_pdf = t4.w;
      int pos = __float_as_int(hit.y), prim = __float_as_int(hit.z), inst = __float_as_int(hit.w);
      sig = sig_mix(sig, pos >= 0 ? (((uint32_t)inst << 20) ^ (uint32_t)prim) : 0xFFFFFFFFu);
      float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
      if (pos < 0) {
        if (env_mis && sc.miss.type == PT_MISS_HDR) {
          // unclamped environment radiance, MIS-weighted against the light
          // sample of the previous (diffuse) bounce
          v3 Le = env_radiance(sc, d);
          float w = 1.f;
          if (env_light && depth > 0 && !(flags & 1u)) w = power_heuristic(prev_pdf, env_pdf(sc, d));
          T = T * (Le * w);
          v3 r = L + T;
          res = make_float4(r.x, r.y, r.z, __uint_as_float(depth + 1));
        } else {
          // miss programs (miss.cu), then color() returns L + clamp(T) (raygen.cu:81-82)
          T = T * miss_color(sc, d);
          v3 r = L + clamp3(T, 0.f, 1.f);
          res = make_float4(r.x, r.y, r.z, __uint_as_float(depth + 1));
        }
      } else {
        int mat = hit_material(sc, pos, prim);
        const DevMaterial& M = sc.mats[mat];
        bool need_uv = tex_needs_uv(sc, M.tex0) || (M.type == PT_MAT_ASHIKHMIN_SHIRLEY && tex_needs_uv(sc, M.tex1));
        HitRec r = hit_record(sc, pos, prim, hit.x, mk3(o4), d, need_uv);
        Surface f = make_surface(sc, r);
        bool is_light = has_mat<MS>(PT_MAT_DIFFUSE_LIGHT) && f.type == PT_MAT_DIFFUSE_LIGHT;
        bool nee = ((has_mat<MS>(PT_MAT_LAMBERTIAN) || has_mat<MS>(PT_MAT_OREN_NAYAR)) &&
                    (f.type == PT_MAT_LAMBERTIAN || f.type == PT_MAT_OREN_NAYAR)) || is_light;
        bool queried = false;
        if (cache_on) {
          // training paths record every vertex (radiance so far, throughput);
          // other paths end at the query depth with a cache lookup
          const bool rec = (dw & TRAIN_BIT) != 0;
          const bool ask = !rec && depth == rc.cache.query_depth;
          uint32_t ti = append_index(rec, P.counters + CTR_TRAIN);
          uint32_t qi = append_index(ask, P.counters + CTR_QUERY);
          if (rec && ti < P.t_cap) {
            write_features(P.t_feat + (size_t)ti * 12, rc.cache, r.P, r.Wo, r.N, f);
            P.t_L[ti] = make_float4(L.x, L.y, L.z, __uint_as_float(dw & SLOT_MASK));
            P.t_T[ti] = make_float4(T.x, T.y, T.z, 0.f);
          }
          if (ask) {
            queried = true;
            if (qi < P.q_cap) {
              write_features(P.q_feat + (size_t)qi * 12, rc.cache, r.P, r.Wo, r.N, f);
              P.q_meta[qi] = make_float4(T.x, T.y, T.z, __uint_as_float(dw & SLOT_MASK));
            }
            res = make_float4(L.x, L.y, L.z, __uint_as_float(depth + 1));
            nee = false;
          }
        }
        if (nee) {
          v3 direct = mk3(0.f);
          if (nl > 0) {
            int index = ((int)(rnd(seed) * nl)) % nl;
            if (!(is_light && nl == 1)) {
              const DevLight& Lt = lights[index];
              const bool env = (FEAT & SH_ENV) && Lt.type == PT_LIGHT_ENV;
              v3 Wi = env ? env_sample(sc, seed) : light_sample(Lt, r.P, seed);
              float lightPDF = env ? env_pdf(sc, Wi) : light_pdf(Lt, r.P, Wi);
              v3 emission = env ? env_radiance(sc, Wi) : light_emission(Lt);
              if (dot(Wi, r.N) < 0.f) {
                sig = sig_mix(sig, 0x33333333u);
              } else {
                shadow_rays++;
                HitOut sh = trace_closest<COUNT>(sc, r.P, Wi, 1e-3f, FINF, st, tc);
                bool occluded = false;
                if (sh.inst >= 0)
                  occluded = env || !(sc.mats[hit_material(sc, sh.pos, sh.prim)].type == PT_MAT_DIFFUSE_LIGHT &&
                                      dot(r.N, Wi) > 0.f);
                sig = sig_mix(sig, occluded ? 0xA5A5A5A5u : 0x5A5A5A5Au);
                if (!occluded) {
                  if (lightPDF != 0.f && !isNull(emission)) {
                    float matPDF = 1.f;
                    v3 mv = brdf_eval<MS>(f, r.Wo, Wi, r.N, matPDF);
                    if (matPDF != 0.f && !isNull(mv)) {
                      float w = power_heuristic(lightPDF, matPDF);
                      direct = direct + mv * emission * w / lightPDF;
                    }
                  }
                  if (!env) {
                    v3 Wb = brdf_sample<MS>(f, r.Wo, r.N, seed);
                    float matPDF = 1.f;
                    v3 mv = brdf_eval<MS>(f, r.Wo, Wb, r.N, matPDF);
                    if (matPDF != 0.f && !isNull(mv)) {
                      float lp = light_pdf(Lt, r.P, Wb);
                      if (lp != 0.f && !isNull(emission)) {
                        float w = power_heuristic(matPDF, lp);
                        direct = direct + mv * emission * w / matPDF;
                      }
                    }
                  }
                }
              }
            }
          }
          L = L + T * direct;
        }
        if (queried) {
          // result = L now; the cache adds T * radiance after inference
        } else if (is_light) {
          // diffuse_light.cu:43-47 and raygen.cu:85-90
          if (dot(r.N, r.Wo) < 0.f) T = T * f.c0;
          if (depth == 0 || (flags & 1u)) L = L + T;
          res = make_float4(L.x, L.y, L.z, __uint_as_float(depth + 1));
        } else {
          v3 ndir;
          alive = scatter_and_roulette<MS>(f, r.Wo, r.N, seed, sig, T, L, ndir, flags, prev_pdf, depth, rc, res);
          if (alive) {
            no = make_float4(r.P.x, r.P.y, r.P.z, __uint_as_float(seed));
            nd = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float((dw & ~SPEC_BIT) | (flags << 31)));
            nT = make_float4(T.x, T.y, T.z, prev_pdf);
            nL = make_float4(L.x, L.y, L.z, __uint_as_float(sig));
          }
        }
      }
      if (!alive) {
        P.result[dw & SLOT_MASK] = res;
        P.sig_final[dw & SLOT_MASK] = sig;
      }
    }
    // wave-aggregated compaction of the survivors into set `out`
    uint64_t mask = __ballot(alive);
    if (mask) {
      int leader = __ffsll((unsigned long long)mask) - 1;
      uint32_t base = 0;
      if ((int)__lane_id() == leader) base = atomicAdd(out_count, (uint32_t)__popcll(mask));
      base = __shfl(base, leader);
      if (alive) {
        uint32_t j = base + lane_rank(mask);
        D.o[j] = no;
        D.d[j] = nd;
        D.T[j] = nT;
        D.L[j] = nL;
      }
    }
  }
  uint32_t a = tc.nodes, b = tc.prims, c = tc.insts;
  for (int off = 32; off > 0; off >>= 1) {
    shadow_rays += __shfl_down(shadow_rays, off);
    if (COUNT) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      c += __shfl_down(c, off);
    }
  }
  if (__lane_id() == 0) {
    if (shadow_rays) atomicAdd((unsigned long long*)&stats[STAT_SHADOW], (unsigned long long)shadow_rays);
    if (COUNT) {
      atomicAdd((unsigned long long*)&stats[STAT_SH_NODES], (unsigned long long)a);
      atomicAdd((unsigned long long*)&stats[STAT_SH_PRIMS], (unsigned long long)b);
      atomicAdd((unsigned long long*)&stats[STAT_SH_INSTS], (unsigned long long)c);
    }
  }
}

// End of a bounce: the consumed set becomes the next output; buckets reset.
__global__ void k_bounce_reset(DevPaths P, int in) {
  if (threadIdx.x == 0) P.counters[CTR_ACTIVE + in] = 0;
  if (threadIdx.x < 8) P.counters[CTR_BUCKETS + threadIdx.x] = 0;
}

// Paths alive after the last bounce end with color() = 0 (raygen.cu:121-122);
// only their signature is recorded.
__global__ __launch_bounds__(BLOCK) void k_exhausted(DevPaths P, int in, uint64_t* stats) {
  const uint32_t n = P.counters[CTR_ACTIVE + in];
  const PathState S = P.st[in];
  if (blockIdx.x == 0 && threadIdx.x == 0 && n)
    atomicAdd((unsigned long long*)&stats[STAT_EXHAUSTED], (unsigned long long)n);
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK)
    P.sig_final[__float_as_uint(S.d[i].w) & SLOT_MASK] = __float_as_uint(S.L[i].w);
}

// Radiance-cache training targets: log(1 + radiance leaving each recorded
// vertex toward the camera side) = log(1 + (L_final - L_before) / T), per
// channel, zero where T is zero or the value is not finite.
__global__ __launch_bounds__(BLOCK) void k_cache_targets(DevPaths P) {
  const uint32_t n = min(P.counters[CTR_TRAIN], P.t_cap);
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    float4 a = P.t_L[i], t = P.t_T[i];
    float4 fin = P.result[__float_as_uint(a.w) & SLOT_MASK];
    float L0[3] = {a.x, a.y, a.z}, T0[3] = {t.x, t.y, t.z}, Lf[3] = {fin.x, fin.y, fin.z};
    for (int c = 0; c < 3; c++) {
      float v = T0[c] > 0.f ? (Lf[c] - L0[c]) / T0[c] : 0.f;
      v = (v > 0.f && v < 3.0e38f) ? v : 0.f;
      P.t_tgt[3 * (size_t)i + c] = log1pf(v);
    }
  }
}

// result[slot] += T * (exp(network output) - 1) for every query of the batch
__global__ __launch_bounds__(BLOCK) void k_cache_apply(DevPaths P) {
  const uint32_t n = min(P.counters[CTR_QUERY], P.q_cap);
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    float4 m = P.q_meta[i];
    const uint32_t slot = __float_as_uint(m.w) & SLOT_MASK;
    float y0 = fmaxf(expm1f(P.q_out[3 * (size_t)i]), 0.f), y1 = fmaxf(expm1f(P.q_out[3 * (size_t)i + 1]), 0.f),
          y2 = fmaxf(expm1f(P.q_out[3 * (size_t)i + 2]), 0.f);
    float4 r = P.result[slot];
    r.x += m.x * y0;
    r.y += m.y * y1;
    r.z += m.z * y2;
    P.result[slot] = r;
  }
}

// raygen.cu:126-134 de_nan and raygen.cu:166 acc_buffer += (rgb, 1), in FP64
// and frame order; accumulation row 0 is the top of the image (raygen.cu:153).
__global__ __launch_bounds__(BLOCK) void k_accumulate(DevPaths P, RenderConsts rc, uint32_t nframes, double* accum) {
  for (uint32_t pix = blockIdx.x * BLOCK + threadIdx.x; pix < rc.WH; pix += gridDim.x * BLOCK) {
    uint32_t x = pix % rc.W, y = pix / rc.W;
    size_t o = 3 * ((size_t)(rc.H - 1 - y) * rc.W + x);
    double a0 = accum[o], a1 = accum[o + 1], a2 = accum[o + 2];
    for (uint32_t f = 0; f < nframes; f++) {
      float4 r = P.result[(size_t)f * rc.WH + pix];
      float cx = (r.x == r.x) ? r.x : 0.f, cy = (r.y == r.y) ? r.y : 0.f, cz = (r.z == r.z) ? r.z : 0.f;
      a0 += (double)cx;
      a1 += (double)cy;
      a2 += (double)cz;
    }
    accum[o] = a0;
    accum[o + 1] = a1;
    accum[o + 2] = a2;
  }
}

// raygen.cu:136-146 make_Color on the FP64 accumulation
__global__ __launch_bounds__(BLOCK) void k_display(const double* accum, uint32_t npix, uint32_t frames,
                                                   uchar4* rgba) {
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < npix; i += gridDim.x * BLOCK) {
    unsigned char c[3];
    for (int k = 0; k < 3; k++) {
      float avg = (float)(accum[3 * (size_t)i + k] / (double)frames);
      float s = clampf(sqrtf(avg), 0.f, 1.f);
      c[k] = (unsigned char)(int)(255.99 * (double)s);
    }
    rgba[i] = make_uchar4(c[0], c[1], c[2], 255);
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) p[i] = v;
}
__global__ void k_fill_f64(double* p, uint64_t n, double v) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) p[i] = v;
}

// --------------------------------------------------------- host launchers
void launch_trace_closest(const DevScene& sc, const pt_ray* rays, uint64_t n, pt_hit* hits, uint32_t* counts,
                          uint2* spill, uint32_t* overflow, int grid, hipStream_t st) {
  if (counts)
    hipLaunchKernelGGL(k_trace_closest<true>, dim3(grid), dim3(BLOCK), 0, st, sc, rays, n, hits, counts, spill,
                       overflow);
  else
    hipLaunchKernelGGL(k_trace_closest<false>, dim3(grid), dim3(BLOCK), 0, st, sc, rays, n, hits, counts, spill,
                       overflow);
}
void launch_trace_occluded(const DevScene& sc, const pt_ray* rays, uint64_t n, uint8_t* occ, uint2* spill,
                           uint32_t* overflow, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_trace_occluded, dim3(grid), dim3(BLOCK), 0, st, sc, rays, n, occ, spill, overflow);
}
__global__ void k_set_frame(uint32_t* f, uint32_t v) { *f = v; }
__global__ void k_next_frames(uint32_t* f, uint32_t step) { *f += step; }
void launch_set_frame(uint32_t* frame_dev, uint32_t value, hipStream_t st) {
  hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, st, frame_dev, value);
}
void launch_next_frames(uint32_t* frame_dev, uint32_t step, hipStream_t st) {
  hipLaunchKernelGGL(k_next_frames, dim3(1), dim3(1), 0, st, frame_dev, step);
}
void launch_raygen(const DevPaths& P, const RenderConsts& rc, uint32_t nslots, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_raygen, dim3(grid), dim3(BLOCK), 0, st, P, rc, nslots);
}
void launch_extend(const DevScene& sc, const DevPaths& P, int in, float tmin, uint2* spill, uint64_t* stats,
                   bool count_traversal, bool sort, int grid, hipStream_t st) {
  if (count_traversal && sort)
    hipLaunchKernelGGL((k_extend<true, true>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, tmin, spill, stats);
  else if (count_traversal)
    hipLaunchKernelGGL((k_extend<true, false>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, tmin, spill, stats);
  else if (sort)
    hipLaunchKernelGGL((k_extend<false, true>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, tmin, spill, stats);
  else
    hipLaunchKernelGGL((k_extend<false, false>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, tmin, spill, stats);
}
void launch_sort(const DevPaths& P, int in, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_sort_offsets, dim3(1), dim3(1), 0, st, P);
  hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(BLOCK), 0, st, P, in);
}
template <int MS, int FEAT>
static void launch_shade_v(const DevScene& sc, const DevPaths& P, int in, int src, const float4* hits, int out,
                           uint32_t depth, const RenderConsts& rc, uint2* spill, uint64_t* stats, int grid,
                           hipStream_t st) {
  hipLaunchKernelGGL((k_shade<MS, FEAT>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, src, hits, out, depth, rc, spill,
                     stats);
}
void launch_shade(const DevScene& sc, const DevPaths& P, int in, bool sorted, int out, uint32_t depth,
                  const RenderConsts& rc, int matset, bool count, uint2* spill, uint64_t* stats, int grid,
                  hipStream_t st) {
  const int src = sorted ? 2 : in;
  const float4* hits = sorted ? P.hit_sorted : P.hit;
  const bool env = (rc.flags & PT_FLAG_ENV_MIS) && sc.num_lights != sc.num_lights_plain;
  const bool cache = (rc.flags & PT_FLAG_RADIANCE_CACHE) != 0;
#define PT_SHADE(MSV, F) launch_shade_v<MSV, F>(sc, P, in, src, hits, out, depth, rc, spill, stats, grid, st)
  if (cache) {  // every material; no traversal counting
    if (env) PT_SHADE(MS_ALL, SH_CACHE | SH_ENV);
    else PT_SHADE(MS_ALL, SH_CACHE);
  } else if (env) {
    if (count) PT_SHADE(MS_ALL, SH_ENV | SH_COUNT);
    else PT_SHADE(MS_ALL, SH_ENV);
  } else if (matset == MS_DIFFUSE) {
    if (count) PT_SHADE(MS_DIFFUSE, SH_COUNT);
    else PT_SHADE(MS_DIFFUSE, 0);
  } else if (matset == MS_SPHERES) {
    if (count) PT_SHADE(MS_SPHERES, SH_COUNT);
    else PT_SHADE(MS_SPHERES, 0);
  } else {
    if (count) PT_SHADE(MS_ALL, SH_COUNT);
    else PT_SHADE(MS_ALL, 0);
  }
#undef PT_SHADE
}
int shade_material_set(uint32_t types_mask) {
  if ((types_mask & ~(uint32_t)MS_DIFFUSE) == 0) return MS_DIFFUSE;
  if ((types_mask & ~(uint32_t)MS_SPHERES) == 0) return MS_SPHERES;
  return MS_ALL;
}
void launch_bounce_reset(const DevPaths& P, int in, hipStream_t st) {
  hipLaunchKernelGGL(k_bounce_reset, dim3(1), dim3(64), 0, st, P, in);
}
void launch_cache_targets(const DevPaths& P, hipStream_t st) {
  hipLaunchKernelGGL(k_cache_targets, dim3(1024), dim3(BLOCK), 0, st, P);
}
void launch_cache_apply(const DevPaths& P, hipStream_t st) {
  hipLaunchKernelGGL(k_cache_apply, dim3(2048), dim3(BLOCK), 0, st, P);
}
void launch_exhausted(const DevPaths& P, int in, uint64_t* stats, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_exhausted, dim3(grid), dim3(BLOCK), 0, st, P, in, stats);
}
void launch_accumulate(const DevPaths& P, const RenderConsts& rc, uint32_t nframes, double* accum, int grid,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(BLOCK), 0, st, P, rc, nframes, accum);
}
void launch_display(const double* accum, uint32_t npix, uint32_t frames, uchar4* rgba, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_display, dim3(grid), dim3(BLOCK), 0, st, accum, npix, frames, rgba);
}
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
  int grid = (int)std::min<uint64_t>((n + BLOCK - 1) / BLOCK + 1, 1024);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(BLOCK), 0, st, p, n, v);
}
void launch_fill_f64(double* p, uint64_t n, double v, hipStream_t st) {
  int grid = (int)std::min<uint64_t>((n + BLOCK - 1) / BLOCK + 1, 1024);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid), dim3(BLOCK), 0, st, p, n, v);
}

}  // namespace pt
