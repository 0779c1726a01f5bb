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
// Wave-aggregated append of `value` to `list` where `pred` holds.
PTD void wave_append(bool pred, uint32_t value, uint32_t* list, uint32_t* count) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (pred) list[base + lane_rank(mask)] = value;
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
    P.ray_o[s] = make_float4(origin.x, origin.y, origin.z, 1e-6f);
    P.ray_d[s] = make_float4(dir.x, dir.y, dir.z, 0.f);
    P.thr[s] = make_float4(1.f, 1.f, 1.f, 0.f);
    P.rad[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.seed[s] = seed;
    P.sig[s] = 0x811C9DC5u;
    P.result[s] = make_float4(0.f, 0.f, 0.f, __uint_as_float(rc.max_depth));
    P.list_a[s] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.counters[CTR_ACTIVE_A] = nslots;
    P.counters[CTR_ACTIVE_B] = 0;
    P.counters[CTR_SHADOW] = 0;
  }
}

// closest-hit query of every active path (raygen.cu:77)
template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_extend(DevScene sc, DevPaths P, const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ count_ptr, uint2* spill,
                                                  uint64_t* stats) {
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  Stack st = make_stack(s_ref, s_tn, spill, P.counters + CTR_STACK_OVERFLOW);
  const uint32_t n = *count_ptr;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&stats[STAT_EXT], (unsigned long long)n);
  TraceCounts c{0, 0, 0};
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    uint32_t s = list[i];
    float4 o = P.ray_o[s], d = P.ray_d[s];
    HitOut h = trace_closest<COUNT>(sc, mk3(o), mk3(d), o.w, FINF, st, c);
    P.hit[s] = make_float4(h.t, __int_as_float(h.pos), __int_as_float(h.prim), __int_as_float(h.inst));
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

// Material sort: bucket 0 = miss, 1 + material type otherwise.  Counts and
// scatter cursors are aggregated per wave (ballot) and per block (LDS), so
// each block issues one global atomic per bucket.
PTD uint32_t material_key(const DevScene& sc, const DevPaths& P, uint32_t s) {
  float4 h = P.hit[s];
  int pos = __float_as_int(h.y);
  return pos < 0 ? 0u : 1u + (uint32_t)sc.mats[hit_material(sc, pos, __float_as_int(h.z))].type;
}
__global__ __launch_bounds__(BLOCK) void k_sort_count(DevScene sc, DevPaths P, const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count_ptr) {
  __shared__ uint32_t cnt[8];
  const uint32_t n = *count_ptr;
  if (threadIdx.x < 8) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    uint32_t key = 0xFFu;
    if (i < n) {
      key = material_key(sc, P, list[i]);
      P.keys[i] = key;
    }
#pragma unroll
    for (uint32_t b = 0; b < 8; b++) {
      uint64_t m = __ballot(key == b);
      if (m && key == b && lane_rank(m) == 0) atomicAdd(&cnt[b], (uint32_t)__popcll(m));
    }
  }
  __syncthreads();
  if (threadIdx.x < 8 && cnt[threadIdx.x]) atomicAdd(&P.counters[CTR_BUCKETS + threadIdx.x], cnt[threadIdx.x]);
}
__global__ void k_sort_offsets(DevPaths P) {
  uint32_t acc = 0;
  for (int b = 0; b < 8; b++) {
    P.counters[CTR_BUCKET_CURSOR + b] = acc;
    acc += P.counters[CTR_BUCKETS + b];
  }
}
__global__ __launch_bounds__(BLOCK) void k_sort_scatter(DevPaths P, const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count_ptr) {
  __shared__ uint32_t cnt[8], base[8];
  const uint32_t n = *count_ptr;
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    uint32_t key = i < n ? P.keys[i] : 0xFFu;
    uint32_t s = i < n ? list[i] : 0;
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
    if (threadIdx.x < 8) base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&P.counters[CTR_BUCKET_CURSOR + threadIdx.x], cnt[threadIdx.x]) : 0u;
    __syncthreads();
    if (key < 8) P.list_sorted[base[key] + wave_off + lane_rank(mine)] = s;
    __syncthreads();
  }
}

// Scatter of the materials without light sampling plus the continuation
// sample of Lambertian / Oren-Nayar, then Russian roulette (raygen.cu:96-118).
// Returns false when the path ended (result written).
PTD bool scatter_and_roulette(const Surface& f, v3 P, v3 Wo, v3 N, uint32_t& seed, uint32_t& sig, v3& T, v3 L,
                              v3& dir, uint32_t& flags, uint32_t depth, const RenderConsts& rc, float4& result) {
  bool spec = false;
  switch (f.type) {
    case PT_MAT_LAMBERTIAN:
    case PT_MAT_OREN_NAYAR: {  // lambertian.cu:62-72, oren_nayar.cu:59-69
      v3 Wi = brdf_sample(f, Wo, N, seed);
      float pdf = 1.f;
      v3 att = brdf_eval(f, Wo, Wi, N, pdf);
      dir = Wi;
      T = T * clamp3(att / pdf, 0.f, 1.f);
      break;
    }
    case PT_MAT_METAL: {  // metal.cu:38-56
      v3 refl = reflect(-Wo, N);
      dir = refl + f.p0 * random_in_unit_sphere(seed);
      T = T * f.c0;
      spec = true;
      break;
    }
    case PT_MAT_DIELECTRIC: {  // dielectric.cu:46-96
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
      break;
    }
    case PT_MAT_ASHIKHMIN_SHIRLEY: {  // ashikhmin_shirley.cu:48-68
      v3 Wi = brdf_sample(f, Wo, N, seed);
      float pdf = 1.f;
      v3 att = brdf_eval(f, Wo, Wi, N, pdf);
      dir = Wi;
      T = T * clamp3(att, 0.f, 1.f);
      spec = true;
      break;
    }
    default: {  // Torrance-Sparrow, torrance_sparrow.cu:39-59
      v3 Wi = brdf_sample(f, Wo, N, seed);
      float pdf = 1.f;
      v3 att = brdf_eval(f, Wo, Wi, N, pdf);
      dir = Wi;
      T = T * clamp3(att / pdf, 0.f, 1.f);
      spec = true;
      break;
    }
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

// Closest-hit / miss programs of every active path, first half of
// Direct_Light (light_sample.cuh:35-59) for materials that sample lights.
__global__ __launch_bounds__(BLOCK) void k_shade(DevScene sc, DevPaths P, const uint32_t* __restrict__ list,
                                                 const uint32_t* __restrict__ count_ptr, uint32_t* next_list,
                                                 uint32_t* next_count, uint32_t depth, RenderConsts rc) {
  const uint32_t n = *count_ptr;
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    bool alive = false, pending = false;
    uint32_t s = 0;
    if (i < n) {
      s = list[i];
      float4 hit = P.hit[s];
      float4 d4 = P.ray_d[s];
      float4 t4 = P.thr[s];
      float4 l4 = P.rad[s];
      v3 d = mk3(d4), T = mk3(t4), L = mk3(l4);
      uint32_t flags = __float_as_uint(t4.w);
      uint32_t seed = P.seed[s];
      int pos = __float_as_int(hit.y), prim = __float_as_int(hit.z), inst = __float_as_int(hit.w);
      uint32_t sig = sig_mix(P.sig[s], pos >= 0 ? (((uint32_t)inst << 20) ^ (uint32_t)prim) : 0xFFFFFFFFu);
      if (pos < 0) {
        // miss programs (miss.cu), then color() returns L + clamp(T) (raygen.cu:81-82)
        T = T * miss_color(sc, d);
        v3 r = L + clamp3(T, 0.f, 1.f);
        P.result[s] = make_float4(r.x, r.y, r.z, __uint_as_float(depth + 1));
      } else {
        float4 o4 = P.ray_o[s];
        int mat = hit_material(sc, pos, prim);
        const DevMaterial& M = sc.mats[mat];
        bool need_uv = tex_needs_uv(sc, M.tex0) || (M.type == PT_MAT_ASHIKHMIN_SHIRLEY && tex_needs_uv(sc, M.tex1));
        HitRec r = hit_record(sc, pos, prim, hit.x, mk3(o4), d, need_uv);
        Surface f = make_surface(sc, r);
        bool nee = (f.type == PT_MAT_LAMBERTIAN || f.type == PT_MAT_OREN_NAYAR || f.type == PT_MAT_DIFFUSE_LIGHT);
        bool is_light = f.type == PT_MAT_DIFFUSE_LIGHT;
        if (nee && sc.num_lights > 0) {
          int nl = sc.num_lights;
          int index = ((int)(rnd(seed) * nl)) % nl;
          if (!(is_light && nl == 1)) {
            const DevLight& Lt = sc.lights[index];
            v3 Wi = light_sample(Lt, r.P, seed);
            float lpdf = light_pdf(Lt, r.P, Wi);
            if (dot(Wi, r.N) < 0.f) {
              sig = sig_mix(sig, 0x33333333u);
            } else {
              pending = true;
              P.nee_p[s] = make_float4(r.P.x, r.P.y, r.P.z, __int_as_float(index));
              P.nee_n[s] = make_float4(r.N.x, r.N.y, r.N.z, __int_as_float(mat));
              P.nee_wo[s] = make_float4(r.Wo.x, r.Wo.y, r.Wo.z, lpdf);
              P.nee_wi[s] = make_float4(Wi.x, Wi.y, Wi.z, __uint_as_float(is_light ? 1u : 0u));
              P.nee_c[s] = make_float4(f.c0.x, f.c0.y, f.c0.z, 0.f);
            }
          }
        }
        if (!pending) {
          if (nee) L = L + T * mk3(0.f);  // L += T * direct with direct = 0
          if (is_light) {
            // diffuse_light.cu:43-47 and raygen.cu:85-90
            if (dot(r.N, r.Wo) < 0.f) T = T * f.c0;
            if (depth == 0 || (flags & 1u)) L = L + T;
            P.result[s] = make_float4(L.x, L.y, L.z, __uint_as_float(depth + 1));
          } else {
            v3 ndir;
            float4 res;
            alive = scatter_and_roulette(f, r.P, r.Wo, r.N, seed, sig, T, L, ndir, flags, depth, rc, res);
            if (alive) {
              P.ray_o[s] = make_float4(r.P.x, r.P.y, r.P.z, 1e-3f);
              P.ray_d[s] = make_float4(ndir.x, ndir.y, ndir.z, 0.f);
            } else {
              P.result[s] = res;
            }
          }
        }
        P.thr[s] = make_float4(T.x, T.y, T.z, __uint_as_float(flags));
        P.rad[s] = make_float4(L.x, L.y, L.z, 0.f);
        P.seed[s] = seed;
      }
      P.sig[s] = sig;
    }
    wave_append(alive, s, next_list, next_count);
    wave_append(pending, s, P.list_shadow, P.counters + CTR_SHADOW);
  }
}

// Shadow ray of Direct_Light (light_sample.cuh:61-72) under the rule: lit
// iff the closest hit in (1e-3, inf) is a diffuse light with dot(N, Wi) > 0,
// or nothing is hit; then the second half of Direct_Light
// (light_sample.cuh:74-102), the radiance update of the closest-hit program
// and the continuation bounce.
__global__ __launch_bounds__(BLOCK) void k_shadow_connect(DevScene sc, DevPaths P, uint32_t* next_list,
                                                          uint32_t* next_count, uint32_t depth, RenderConsts rc,
                                                          uint2* spill, uint64_t* stats) {
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  Stack st = make_stack(s_ref, s_tn, spill, P.counters + CTR_STACK_OVERFLOW);
  const uint32_t n = P.counters[CTR_SHADOW];
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&stats[STAT_SHADOW], (unsigned long long)n);
  TraceCounts tc{0, 0, 0};
  for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
    uint32_t i = i0 + threadIdx.x;
    bool alive = false;
    uint32_t s = 0;
    if (i < n) {
      s = P.list_shadow[i];
      float4 p4 = P.nee_p[s], n4 = P.nee_n[s], wo4 = P.nee_wo[s], wi4 = P.nee_wi[s], c4 = P.nee_c[s];
      v3 Pp = mk3(p4), N = mk3(n4), Wo = mk3(wo4), Wi = mk3(wi4);
      HitOut sh = trace_closest<false>(sc, Pp, Wi, 1e-3f, FINF, st, tc);
      bool occluded = false;
      if (sh.inst >= 0)
        occluded = !(sc.mats[hit_material(sc, sh.pos, sh.prim)].type == PT_MAT_DIFFUSE_LIGHT && dot(N, Wi) > 0.f);
      float4 t4 = P.thr[s], l4 = P.rad[s];
      v3 T = mk3(t4), L = mk3(l4);
      uint32_t flags = __float_as_uint(t4.w);
      uint32_t seed = P.seed[s];
      uint32_t sig = P.sig[s];
      int li = __float_as_int(p4.w), mat = __float_as_int(n4.w);
      bool is_light = (__float_as_uint(wi4.w) & 1u) != 0;
      float lightPDF = wo4.w;
      const DevMaterial& M = sc.mats[mat];
      Surface f;
      f.type = M.type;
      f.c0 = mk3(c4);
      f.c1 = mk3(0.f);
      f.p0 = M.p0;
      f.p1 = M.p1;
      sig = sig_mix(sig, occluded ? 0xA5A5A5A5u : 0x5A5A5A5Au);
      v3 direct = mk3(0.f);
      if (!occluded) {
        v3 emission = light_emission(sc.lights[li]);
        if (lightPDF != 0.f && !isNull(emission)) {
          float matPDF = 1.f;
          v3 mv = brdf_eval(f, Wo, Wi, N, matPDF);
          if (matPDF != 0.f && !isNull(mv)) {
            float w = power_heuristic(lightPDF, matPDF);
            direct = direct + mv * emission * w / lightPDF;
          }
        }
        v3 Wb = brdf_sample(f, Wo, N, seed);
        float matPDF = 1.f;
        v3 mv = brdf_eval(f, Wo, Wb, N, matPDF);
        if (matPDF != 0.f && !isNull(mv)) {
          float lp = light_pdf(sc.lights[li], Pp, Wb);
          if (lp != 0.f && !isNull(emission)) {
            float w = power_heuristic(matPDF, lp);
            direct = direct + mv * emission * w / matPDF;
          }
        }
      }
      L = L + T * direct;
      if (is_light) {
        if (dot(N, Wo) < 0.f) T = T * f.c0;
        if (depth == 0 || (flags & 1u)) L = L + T;
        P.result[s] = make_float4(L.x, L.y, L.z, __uint_as_float(depth + 1));
      } else {
        v3 ndir;
        float4 res;
        alive = scatter_and_roulette(f, Pp, Wo, N, seed, sig, T, L, ndir, flags, depth, rc, res);
        if (alive) {
          P.ray_o[s] = make_float4(Pp.x, Pp.y, Pp.z, 1e-3f);
          P.ray_d[s] = make_float4(ndir.x, ndir.y, ndir.z, 0.f);
        } else {
          P.result[s] = res;
        }
      }
      P.thr[s] = make_float4(T.x, T.y, T.z, __uint_as_float(flags));
      P.rad[s] = make_float4(L.x, L.y, L.z, 0.f);
      P.seed[s] = seed;
      P.sig[s] = sig;
    }
    wave_append(alive, s, next_list, next_count);
  }
}

// Zero the counters the next bounce appends to.
__global__ void k_bounce_reset(DevPaths P, int zero_ctr) {
  if (threadIdx.x == 0) {
    P.counters[zero_ctr] = 0;
    P.counters[CTR_SHADOW] = 0;
    for (int b = 0; b < 8; b++) P.counters[CTR_BUCKETS + b] = 0;
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
void launch_raygen(const DevPaths& P, const RenderConsts& rc, uint32_t nslots, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_raygen, dim3(grid), dim3(BLOCK), 0, st, P, rc, nslots);
}
void launch_extend(const DevScene& sc, const DevPaths& P, const uint32_t* list, const uint32_t* count, uint2* spill,
                   uint64_t* stats, bool count_traversal, int grid, hipStream_t st) {
  if (count_traversal)
    hipLaunchKernelGGL(k_extend<true>, dim3(grid), dim3(BLOCK), 0, st, sc, P, list, count, spill, stats);
  else
    hipLaunchKernelGGL(k_extend<false>, dim3(grid), dim3(BLOCK), 0, st, sc, P, list, count, spill, stats);
}
void launch_sort(const DevScene& sc, const DevPaths& P, const uint32_t* list, const uint32_t* count, int grid,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_sort_count, dim3(grid), dim3(BLOCK), 0, st, sc, P, list, count);
  hipLaunchKernelGGL(k_sort_offsets, dim3(1), dim3(1), 0, st, P);
  hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(BLOCK), 0, st, P, list, count);
}
void launch_shade(const DevScene& sc, const DevPaths& P, const uint32_t* list, const uint32_t* count,
                  uint32_t* next_list, uint32_t* next_count, uint32_t depth, const RenderConsts& rc, int grid,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_shade, dim3(grid), dim3(BLOCK), 0, st, sc, P, list, count, next_list, next_count, depth, rc);
}
void launch_shadow_connect(const DevScene& sc, const DevPaths& P, uint32_t* next_list, uint32_t* next_count,
                           uint32_t depth, const RenderConsts& rc, uint2* spill, uint64_t* stats, int grid,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_shadow_connect, dim3(grid), dim3(BLOCK), 0, st, sc, P, next_list, next_count, depth, rc, spill,
                     stats);
}
void launch_bounce_reset(const DevPaths& P, int zero_ctr, hipStream_t st) {
  hipLaunchKernelGGL(k_bounce_reset, dim3(1), dim3(64), 0, st, P, zero_ctr);
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
