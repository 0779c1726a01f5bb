// pt_render.hip -- gfx950 kernels of the per-sample render path: query
// kernels behind pt_trace_closest / pt_trace_occluded and the wavefront
// integrator that replaces the OptiX megakernel (programs/raygen.cu:148-168
// renderPixel and :66-123 color, closest-hit programs of
// programs/materials/*.cu, miss programs, shadow any-hit).
//
// Wavefront, one bounce at a time over SoA path state in HBM:
//   [extend (depth 0: path start + camera rays) -> (material sort)
//    -> shade + shadow ray + connect]* -> accumulate
// Path sets are split into wave segments (pt_layout.hpp): each wave reads
// its own segment and compacts its survivors into its own segment of the
// next set with a ballot, so the bounce loop takes no global atomic.  The
// counts live in device memory, so a whole batch is enqueued without host
// synchronisation.
#include "pt_kernels.hpp"
#include "pt_shade.hpp"

namespace pt {

// -------------------------------------------------------------- helpers
PTD uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// This wave's segment of a path set (layout in pt_layout.hpp).
struct Seg {
  uint32_t stride;  // G * BLOCK
  uint32_t base;    // b * BLOCK + v * 64
  uint32_t w;       // wave index in the grid
  PTD uint32_t pos(uint32_t j) const { return (j >> 6) * stride + base + (j & 63u); }
};
PTD Seg my_seg() {
  const uint32_t v = threadIdx.x >> 6;
  return Seg{gridDim.x * BLOCK, blockIdx.x * BLOCK + v * 64u, blockIdx.x * WAVES + v};
}
// positions of the segment below n: the segment count of a dense set
PTD uint32_t dense_count(uint32_t n, const Seg& g) {
  const uint32_t full = n / g.stride, r = n - full * g.stride;
  return full * 64u + (r > g.base ? min(r - g.base, 64u) : 0u);
}
PTD uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// live paths of this wave's segment of state set `set`
PTD uint32_t seg_count(const DevPaths& P, int set, const Seg& g) {
  return uniform(P.seg[(size_t)set * MAX_WAVES + g.w]);
}
// per-wave statistics row (one writer: lane 0 of the wave)
PTD void stat_add(uint64_t* stats, const Seg& g, int k, uint64_t v) {
  if (v) stats[(size_t)g.w * STAT_WORDS + k] += v;
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

template <int NL = STK_LDS>
PTD StackT<NL> make_stack(uint32_t* s_ref, float* s_tn, uint2* spill, uint32_t* overflow) {
  StackT<NL> st;
  st.ref = s_ref + threadIdx.x;
  st.tn = s_tn + threadIdx.x;
  st.spill = spill + (size_t)(blockIdx.x * BLOCK + threadIdx.x) * STK_TOTAL;
  st.sp = 0;
  st.overflow = overflow;
  return st;
}

// ------------------------------------------------------- query kernels
// EXT: volumes draw their free flight from the ray's index
template <bool COUNT, bool EXT>
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
    HitOut h = trace_closest<COUNT, EXT>(sc, mk3(a), mk3(b), a.w, b.w, st, c, EXT ? query_vseed(i) : 0u);
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

template <bool EXT>
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
    HitOut h = trace_closest<false, EXT>(sc, mk3(a), mk3(b), a.w, b.w, st, c, EXT ? query_vseed(i) : 0u);
    uint8_t o = 0;
    if (h.inst >= 0) o = sc.mats[hit_material(sc, h.pos, h.prim)].type == PT_MAT_DIFFUSE_LIGHT ? 0 : 1;
    occ[i] = o;
  }
}

// ------------------------------------------------------ wavefront kernels
// Slot of position p in a batch of nslots.  Slots come in 64-slot chunks
// (64 consecutive pixels of one row, coherent camera rays); the chunk index
// goes through a keyed 4-round Feistel permutation, cycle-walked into the Q
// full chunks, so that the chunks a wave owns come from all over the
// batch's frames and every wave keeps about the image-average share of
// surviving paths (with slot = position a wave would keep the same 64
// pixels in every frame, and the waves on the light would idle).  A partial
// last chunk keeps its place.
PTD uint32_t slot_of(uint32_t p, uint32_t nslots) {
  const uint32_t Q = nslots >> 6, q = p >> 6;
  if (Q < 2 || q >= Q) return p;
  int bits = 32 - __clz(Q - 1);  // 2^bits >= Q, made even
  bits += bits & 1;
  const int half = bits >> 1;
  const uint32_t mask = (1u << half) - 1u;
  uint32_t x = q;
  do {  // a permutation of [0, 2^bits): the walk from q < Q returns below Q
    uint32_t L = x >> half, R = x & mask;
    for (uint32_t r = 1; r <= 4; r++) {
      const uint32_t t = L ^ (mix32(R ^ (0x9E3779B9u * r)) & mask);
      L = R;
      R = t;
    }
    x = (L << half) | R;
  } while (x >= Q);
  return (x << 6) | (p & 63u);
}

// Throughput (w: pdf of the previous bounce) and radiance (w: path signature,
// the FNV-1a offset basis) of every path before its first bounce.  The first
// shading pass (depth 0) takes them from here, so the path start does not
// store them: state set 0 holds only o and d until depth 1 rewrites it
// (max_depth >= 1; the material sort copies the unset T and L of depth 0,
// which nothing reads).
PTD float4 path_T0() { return make_float4(1.f, 1.f, 1.f, 1.f); }
PTD float4 path_L0() { return make_float4(0.f, 0.f, 0.f, __uint_as_float(0x811C9DC5u)); }

// raygen.cu:148-168 renderPixel up to the first rtTrace, raygen.cu:50-64
// generateRay, raygen.cu:66-74 color() set-up.
// One camera path: slot s (frame `frame` of the scene, pixel `pix`) at
// position p of state set 0; in a batch of scenes the RNG state's top byte
// is replaced by the scene index.  Returns the stored ray and its time.
struct PathStart {
  float4 o, d;
  float tm;
};
PTD PathStart start_path(const DevPaths& P, const RenderConsts& rc, uint32_t p, uint32_t s, uint32_t frame,
                         uint32_t pix, bool batch, uint32_t sid) {
  const PathState S = P.st[0];
  uint32_t x = pix % rc.W, y = pix / rc.W;
  uint32_t seed = tea64(rc.W * y + x, frame);
  float u = ((float)x + rnd(seed)) / (float)rc.W;
  float v = ((float)y + rnd(seed)) / (float)rc.H;
  const DevCamera& c = rc.cam;
  v3 cu = mk3(c.u[0], c.u[1], c.u[2]), cv = mk3(c.v[0], c.v[1], c.v[2]);
  v3 rd = c.lens_radius * random_in_unit_disk(seed);
  v3 lens = cu * rd.x + cv * rd.y;
  v3 origin = mk3(c.org[0], c.org[1], c.org[2]) + lens;
  v3 dir = mk3(c.llc[0], c.llc[1], c.llc[2]) + u * mk3(c.hor[0], c.hor[1], c.hor[2]) +
           v * mk3(c.ver[0], c.ver[1], c.ver[2]) - origin;
  // prd.time (raygen.cu:69); kept per slot when the scene has moving spheres
  const float tm = c.time0 + rnd(seed) * (c.time1 - c.time0);
  if (rc.store_time) P.ptime[s] = tm;
  const float4 o4 =
      make_float4(origin.x, origin.y, origin.z, __uint_as_float(batch ? ((seed & 0xFFFFFFu) | (sid << 24)) : seed));
  uint32_t dw = s;
  if ((rc.flags & PT_FLAG_RADIANCE_CACHE) && train_hash(pix, frame) < rc.cache.train_threshold) dw |= TRAIN_BIT;
  const float4 d4 = make_float4(dir.x, dir.y, dir.z, __uint_as_float(dw));
  S.o[p] = o4;
  S.d[p] = d4;
  P.result[s] = make_float4(0.f, 0.f, 0.f, __uint_as_float(rc.max_depth));
  return PathStart{o4, d4, tm};
}
PTD void reset_segments(const DevPaths& P, uint32_t nslots, const Seg& g) {
  if (__lane_id() == 0) {
    P.seg[g.w] = dense_count(nslots, g);
    P.tseg[g.w] = 0u;
    P.qseg[g.w] = 0u;
  }
}

// material bucket of a hit: 0 = miss, 1 + material type otherwise
PTD uint32_t material_key(const DevScene& sc, float4 h) {
  int pos = __float_as_int(h.y);
  // the isotropic medium and the normal shader share the last bucket
  return pos < 0 ? 0u : 1u + (uint32_t)min(sc.mats[hit_material(sc, pos, __float_as_int(h.z))].type, 6);
}

// Closest-hit query of every live path (raygen.cu:77).  With SORT the
// material bucket of each hit is recorded and counted per wave and bucket.
// 5 waves per SIMD (at most 96 VGPRs, no spills): latency-bound traversal
// of large meshes gains 3.5 % over 4 waves; the Cornell box is unchanged.
//
// START (depth 0, in == 0): the query starts the nstart camera paths
// itself (raygen.cu:148-168, 50-74): position i of the dense state set 0
// is slot slot_of(i, nstart), its ray is stored for the shading pass and
// traced at once.  A batch of scenes (MULTI) finds the slot's item
// (scene and frame) in the BatchStart tables.  The path start is VALU-bound (the 64
// round seed hash) and the traversal latency-bound, so one launch hides
// most of the former under the latter.  A separate instantiation: the
// path-start code costs the traversal of the other depths a few spills.
// Occupancy and LDS stack entries per lane of a closest-hit instantiation:
// without meshes (short stacks, fewer registers) the bounce-ray kernels fit
// 6 waves per SIMD with 12 LDS entries and no spills (Cornell box 10.2 ->
// 10.1 ms, random spheres 2.48 -> 2.38 ms per step); the camera-ray ones
// spill at 6 and keep 5.
template <int KINDS, bool START>
constexpr int extend_waves() { return (KINDS & 8) || START ? PT_EXTEND_WAVES : 6; }
template <int KINDS>
constexpr int extend_stk() { return (KINDS & 8) ? STK_LDS_EXTEND : 12; }
template <bool COUNT, bool SORT, bool MULTI, bool EXT, int KINDS = KIND_ALL, bool START = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(extend_waves<KINDS, START>(), 8))) void k_extend(
    DevScene sc0, DevPaths P, int in, float tmin, uint2* spill, uint64_t* stats, const DevScene* __restrict__ scenes,
    RenderConsts rc, uint32_t nstart, BatchStart bs) {
  constexpr int NS = extend_stk<KINDS>();
  __shared__ uint32_t s_ref[NS * BLOCK];
  __shared__ float s_tn[NS * BLOCK];
  StackT<NS> st = make_stack<NS>(s_ref, s_tn, spill, P.counters + CTR_STACK_OVERFLOW);
  const Seg g = my_seg();
  constexpr bool starts = START;
  uint32_t n;
  if (starts) {
    if (!MULTI && rc.frame_dev) rc.frame0 = *rc.frame_dev;  // replayed graphs take the batch's frame from memory
    n = dense_count(nstart, g);
    reset_segments(P, nstart, g);
  } else {
    n = seg_count(P, in, g);
  }
  const PathState S = P.st[in];
  uint32_t bc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  TraceCounts c{0, 0, 0};
  constexpr bool REFILL = PT_EXTEND_REFILL > 0 && !START && !MULTI && (KINDS & 8) != 0;
  if (REFILL) {
    // Bounce rays in scenes with meshes differ widely in length (the sky
    // after one node, a mesh after dozens): each lane keeps tracing paths
    // of the wave's segment one after another, and the idle lanes take the
    // next paths together once PT_EXTEND_REFILL of them wait (few ray
    // set-ups per refill).  Every path is traced exactly as alone, so hits
    // and counts do not depend on which lane takes it.  (Scenes of a few
    // primitives, whose rays are all short, run faster in lock-step: the
    // lanes' steps stay alike.)
    Tracer<COUNT, EXT, KINDS> tr;
    uint32_t j = __lane_id();
    bool busy = j < n;
    uint32_t next = 64;  // wave-uniform: the next position to hand out
    auto take = [&](uint32_t jj) {
      const uint32_t i = g.pos(jj);
      const float4 o = S.o[i], d = S.d[i];
      const float tm = (EXT && sc0.mov_instances) ? P.ptime[__float_as_uint(d.w) & SLOT_MASK] : 0.f;
      tr.begin(sc0, mk3(o), mk3(d), tmin, FINF, st, __float_as_uint(o.w), tm);
    };
    if (busy) take(j);
    while (true) {
      if (busy && !tr.step(sc0, st, c)) {
        P.hit[g.pos(j)] = make_float4(tr.h.t, __int_as_float(tr.h.pos), __int_as_float(tr.h.prim),
                                      __int_as_float(tr.h.inst));
        busy = false;
      }
      const uint64_t idle = __ballot(!busy);
      if (next >= n) {
        if (idle == ~0ull) break;
      } else if (__popcll(idle) >= (uint32_t)PT_EXTEND_REFILL) {
        if (!busy) {
          const uint32_t jn = next + lane_rank(idle);
          if (jn < n) {
            j = jn;
            take(j);
            busy = true;
          }
        }
        next += (uint32_t)__popcll(idle);
      }
    }
    if (SORT) {
      // material keys and bucket counts of the wave's hits, 64 at a time
      // (the material look-ups' latency paid once per 64 paths); the
      // fence makes the hits other lanes stored visible to these loads
      __threadfence_block();
      for (uint32_t j0 = 0; j0 < n; j0 += 64) {
        const uint32_t jj = j0 + __lane_id();
        uint32_t key = 0xFFu;
        if (jj < n) {
          const uint32_t i = g.pos(jj);
          key = material_key(sc0, P.hit[i]);
          P.keys[i] = key;
        }
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) bc[b] += (uint32_t)__popcll(__ballot(key == b));
      }
    }
  }
  for (uint32_t j0 = 0; j0 < (REFILL ? 0u : n); j0 += 64) {
    const uint32_t j = j0 + __lane_id(), i = g.pos(j);
    uint32_t key = 0xFFu;
    if (j < n) {
      float4 o, d;
      float tm = 0.f;
      if (starts) {
        const uint32_t s = slot_of(i, nstart);
        PathStart ps;
        if (MULTI) {
          uint32_t lo = 0, hi = bs.nitems - 1;  // last item starting at or before s
          while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (bs.off[mid] <= s) lo = mid;
            else hi = mid - 1;
          }
          const uint32_t sid = bs.scene[lo];
          ps = start_path(P, bs.rcs[sid], i, s, bs.frame[lo], s - bs.off[lo], true, sid);
        } else {
          const uint32_t f = s / rc.WH;
          ps = start_path(P, rc, i, s, rc.frame0 + f, s - f * rc.WH, false, 0u);
        }
        o = ps.o;
        d = ps.d;
        if (EXT && sc0.mov_instances) tm = ps.tm;
      } else {
        o = S.o[i];
        d = S.d[i];
      }
      const DevScene& sc = MULTI ? scenes[__float_as_uint(o.w) >> 24] : sc0;
      if (EXT && sc.mov_instances && !starts) tm = P.ptime[__float_as_uint(d.w) & SLOT_MASK];
      HitOut h = trace_closest<COUNT, EXT, KINDS>(sc, mk3(o), mk3(d), tmin, FINF, st, c, __float_as_uint(o.w), tm);
      float4 hv = make_float4(h.t, __int_as_float(h.pos), __int_as_float(h.prim), __int_as_float(h.inst));
      P.hit[i] = hv;
      if (SORT) {
        key = material_key(sc, hv);
        P.keys[i] = key;
      }
    }
    if (SORT) {
#pragma unroll
      for (uint32_t b = 0; b < 8; b++) bc[b] += (uint32_t)__popcll(__ballot(key == b));
    }
  }
  uint32_t a = c.nodes, b = c.prims, d = c.insts;
  if (COUNT) {
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      d += __shfl_down(d, off);
    }
  }
  if (__lane_id() == 0) {
    if (SORT) {
#pragma unroll
      for (int k = 0; k < 8; k++) P.bcount[(size_t)k * (gridDim.x * WAVES) + g.w] = bc[k];
    }
    stat_add(stats, g, STAT_EXT, n);
    if (COUNT) {
      stat_add(stats, g, STAT_NODES, a);
      stat_add(stats, g, STAT_PRIMS, b);
      stat_add(stats, g, STAT_INSTS, d);
    }
  }
}

// Exclusive scan of the (bucket, wave) counts in bucket-major order: the
// first sorted position of each wave's paths of each bucket, and the size
// of the sorted set.  The table holds 8 rows of `waves` entries (entry
// k * waves + w: bucket k of wave w, `waves` the launch grid's waves).
// Each of the block's 16 waves owns a contiguous chunk of whole 64-entry
// rows, read coalesced 16 loads at a time: a first pass sums the chunk,
// and after the chunk totals are exchanged a second pass (re-reading from
// L2) scans 64 entries at a time across the wave's lanes and writes the
// positions.
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_BATCH = 16;
__global__ __launch_bounds__(SCAN_THREADS) void k_sort_scan(DevPaths P, uint32_t waves) {
  constexpr int NW = SCAN_THREADS / 64;
  __shared__ uint32_t wsum[NW];
  const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  const uint32_t total = 8u * waves, rows = (total + 63u) / 64u;
  const uint32_t per = ((rows + NW - 1) / NW + SCAN_BATCH - 1) / SCAN_BATCH * SCAN_BATCH;  // rows per wave
  const uint32_t e0 = wv * per * 64u + lane;
  auto count = [&](uint32_t e) -> uint32_t { return e < total ? P.bcount[e] : 0u; };
  uint32_t sum = 0;
#pragma unroll 1
  for (uint32_t b0 = 0; b0 < per; b0 += SCAN_BATCH) {
    uint32_t v[SCAN_BATCH];
#pragma unroll
    for (int i = 0; i < SCAN_BATCH; i++) v[i] = count(e0 + 64u * (b0 + i));
#pragma unroll
    for (int i = 0; i < SCAN_BATCH; i++) sum += v[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if (lane == 0) wsum[wv] = sum;
  __syncthreads();
  uint32_t run = 0, all = 0;
#pragma unroll
  for (int k = 0; k < NW; k++) {
    const uint32_t t = wsum[k];
    run += (uint32_t)k < wv ? t : 0u;
    all += t;
  }
#pragma unroll 1
  for (uint32_t b0 = 0; b0 < per; b0 += SCAN_BATCH) {
    uint32_t v[SCAN_BATCH];
#pragma unroll
    for (int i = 0; i < SCAN_BATCH; i++) v[i] = count(e0 + 64u * (b0 + i));
#pragma unroll
    for (int i = 0; i < SCAN_BATCH; i++) {
      uint32_t s = v[i];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(s, off);
        if (lane >= (uint32_t)off) s += t;
      }
      const uint32_t e = e0 + 64u * (b0 + i);
      if (e < total) P.bbase[e] = run + s - v[i];
      run += __shfl(s, 63);
    }
  }
  if (threadIdx.x == 0) P.counters[CTR_SORTED] = all;
}
// Move path state and hit into material-bucket order (state set 2, a dense
// set): each wave writes its bucket-b paths from its scanned base on.
__global__ __launch_bounds__(BLOCK) void k_sort_scatter(DevPaths P, int in) {
  const Seg g = my_seg();
  const uint32_t n = seg_count(P, in, g);
  const PathState S = P.st[in], D = P.st[2];
  uint32_t cur[8];
#pragma unroll
  for (int k = 0; k < 8; k++) cur[k] = uniform(P.bbase[(size_t)k * (gridDim.x * WAVES) + g.w]);
  for (uint32_t j0 = 0; j0 < n; j0 += 64) {
    const uint32_t j = j0 + __lane_id(), i = g.pos(j);
    const uint32_t key = j < n ? P.keys[i] : 0xFFu;
    uint32_t dst = 0;
#pragma unroll
    for (uint32_t b = 0; b < 8; b++) {
      const uint64_t m = __ballot(key == b);
      if (key == b) dst = cur[b] + lane_rank(m);
      cur[b] += (uint32_t)__popcll(m);
    }
    if (key < 8) {
      D.o[dst] = S.o[i];
      D.d[dst] = S.d[i];
      D.T[dst] = S.T[i];
      D.L[dst] = S.L[i];
      P.hit_sorted[dst] = P.hit[i];
    }
  }
  if (__lane_id() == 0) P.seg[2 * MAX_WAVES + g.w] = dense_count(P.counters[CTR_SORTED], g);
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
  } else if (has_mat<MS>(PT_MAT_ISOTROPIC) && f.type == PT_MAT_ISOTROPIC) {  // isotropic.cu:42-50
    v3 Wi = brdf_sample<MS>(f, Wo, N, seed);
    float pdf = 1.f;
    v3 att = brdf_eval<MS>(f, Wo, Wi, N, pdf);
    dir = Wi;
    T = T * (att / pdf);
    prev_pdf = pdf;
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
// lights, SH_EXT traces shadow rays through volumes; each is compiled only
// into the variants that use it.
constexpr int SH_COUNT = 1, SH_CACHE = 2, SH_ENV = 4, SH_MULTI = 8, SH_EXT = 16;
template <int MS, int FEAT, int KINDS = KIND_ALL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_shade(
    DevScene sc0, DevPaths P, int in, int src, const float4* __restrict__ hits, int out, uint32_t depth,
    RenderConsts rc0, uint2* spill, uint64_t* stats, const DevScene* __restrict__ scenes,
    const RenderConsts* __restrict__ rcs) {
  constexpr bool MULTI = (FEAT & SH_MULTI) != 0;
  constexpr bool EXT = (FEAT & SH_EXT) != 0;
  __shared__ uint32_t s_ref[STK_LDS * BLOCK];
  __shared__ float s_tn[STK_LDS * BLOCK];
  // each lane's T and L wait here while its shadow ray is traced: fewer
  // registers live across the traversal (fewer scratch spills; with the
  // stack, 38 KB per block keeps 4 blocks of 4 waves per CU)
  __shared__ float s_keep[6 * BLOCK];
  Stack st = make_stack(s_ref, s_tn, spill, P.counters + CTR_STACK_OVERFLOW);
  const Seg g = my_seg();
  const uint32_t n = seg_count(P, src, g);
  const PathState S = P.st[src], D = P.st[out];
  uint32_t survivors = 0;  // wave-uniform: paths written to this wave's segment of `out`
  constexpr bool COUNT = (FEAT & SH_COUNT) != 0;
  const bool cache_on = !MULTI && (FEAT & SH_CACHE) && (rc0.flags & PT_FLAG_RADIANCE_CACHE) != 0;
  TraceCounts tc{0, 0, 0};
  uint32_t shadow_rays = 0, exhausted = 0;
  // radiance-cache records of this wave (wave-uniform running counts):
  // training vertices continue the batch's count, queries start at zero
  uint32_t tcount = 0, qcount = 0, tcap = 0;
  if (cache_on) {
    tcount = uniform(P.tseg[g.w]);
    tcap = dense_count(P.t_cap, g);
  }
  for (uint32_t j0 = 0; j0 < n; j0 += 64) {
    const uint32_t i = g.pos(j0 + __lane_id());
    bool alive = false, did_rec = false, did_ask = false;
    float4 no = make_float4(0.f, 0.f, 0.f, 0.f), nd = no, nT = no, nL = no;
    if (j0 + __lane_id() < n) {
      float4 hit = hits[i];
      float4 o4 = S.o[i], d4 = S.d[i];
      const float4 t4 = depth == 0 ? path_T0() : S.T[i], l4 = depth == 0 ? path_L0() : S.L[i];
      v3 d = mk3(d4), T = mk3(t4), L = mk3(l4);
      uint32_t seed = __float_as_uint(o4.w), dw = __float_as_uint(d4.w);
      // a batch of scenes keeps the path's scene in the RNG state's top
      // byte (the 24-bit outputs never depend on it)
      const uint32_t sid = MULTI ? (seed >> 24) : 0u;
      const DevScene& sc = MULTI ? scenes[sid] : sc0;
      const RenderConsts& rc = MULTI ? rcs[sid] : rc0;
      // light table: environment lights take part only with PT_FLAG_ENV_MIS
      const bool env_mis = (FEAT & SH_ENV) && (rc.flags & PT_FLAG_ENV_MIS) != 0;
      const DevLight* lights = env_mis ? sc.lights : sc.lights_plain;
      const int nl = env_mis ? sc.num_lights : sc.num_lights_plain;
      const bool env_light = env_mis && sc.env_total != 0 && sc.num_lights != sc.num_lights_plain;
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
        const float tm = (EXT && sc.mov_instances) ? P.ptime[dw & SLOT_MASK] : 0.f;
        HitRec r = hit_record<KINDS, EXT>(sc, pos, prim, hit.x, mk3(o4), d, need_uv, tm);
        Surface f = make_surface(sc, r);
        bool is_light = has_mat<MS>(PT_MAT_DIFFUSE_LIGHT) && f.type == PT_MAT_DIFFUSE_LIGHT;
        // the isotropic medium samples lights without the hemisphere test of
        // light_sample.cuh:60 and :71 (its +X normal is not a surface's)
        const bool iso = has_mat<MS>(PT_MAT_ISOTROPIC) && f.type == PT_MAT_ISOTROPIC;
        bool nee = ((has_mat<MS>(PT_MAT_LAMBERTIAN) || has_mat<MS>(PT_MAT_OREN_NAYAR)) &&
                    (f.type == PT_MAT_LAMBERTIAN || f.type == PT_MAT_OREN_NAYAR)) || is_light || iso;
        bool queried = false;
        if (cache_on) {
          // training paths record every vertex (radiance so far, throughput);
          // other paths end at the query depth with a cache lookup
          // (every lane with rec or ask is active here, so the ballots rank
          // exactly the lanes the converged counts below add up)
          const bool rec = (dw & TRAIN_BIT) != 0;
          const bool ask = !rec && depth == rc.cache.query_depth;
          const uint32_t tj = tcount + lane_rank(__ballot(rec));
          const uint32_t qj = qcount + lane_rank(__ballot(ask));
          did_rec = rec;
          did_ask = ask;
          if (rec && tj < tcap) {
            const uint32_t k = g.pos(tj);
            write_features(P.t_feat + (size_t)k * 12, rc.cache, r.P, r.Wo, r.N, f);
            P.t_L[k] = make_float4(L.x, L.y, L.z, __uint_as_float(dw & SLOT_MASK));
            P.t_T[k] = make_float4(T.x, T.y, T.z, 0.f);
          }
          if (ask) {
            // queries never outnumber the wave's inputs: the position exists
            queried = true;
            const uint32_t k = g.pos(qj);
            write_features(P.q_feat + (size_t)k * 12, rc.cache, r.P, r.Wo, r.N, f);
            P.q_meta[k] = make_float4(T.x, T.y, T.z, __uint_as_float(dw & SLOT_MASK));
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
              if (!iso && dot(Wi, r.N) < 0.f) {
                sig = sig_mix(sig, 0x33333333u);
              } else {
                shadow_rays++;
                const uint32_t kt = threadIdx.x;
                s_keep[kt] = T.x;
                s_keep[BLOCK + kt] = T.y;
                s_keep[2 * BLOCK + kt] = T.z;
                s_keep[3 * BLOCK + kt] = L.x;
                s_keep[4 * BLOCK + kt] = L.y;
                s_keep[5 * BLOCK + kt] = L.z;
                // an environment sample is lit iff its shadow ray hits
                // nothing (any hit occludes): the first hit found decides it
                HitOut sh = trace_closest<COUNT, EXT, KINDS, true>(sc, r.P, Wi, 1e-3f, FINF, st, tc, seed, tm, env);
                T = mk3(s_keep[kt], s_keep[BLOCK + kt], s_keep[2 * BLOCK + kt]);
                L = mk3(s_keep[3 * BLOCK + kt], s_keep[4 * BLOCK + kt], s_keep[5 * BLOCK + kt]);
                const v3 emission = env ? env_radiance(sc, Wi) : light_emission(Lt);
                bool occluded = false;
                if (sh.inst >= 0)
                  occluded = env || !(sc.mats[hit_material(sc, sh.pos, sh.prim)].type == PT_MAT_DIFFUSE_LIGHT &&
                                      (iso || dot(r.N, Wi) > 0.f));
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
        } else if (has_mat<MS>(PT_MAT_NORMAL) && f.type == PT_MAT_NORMAL) {
          // normal_shader.cu:22-34: the radiance becomes 0.5 N + 0.5 and the
          // path is cancelled (color() returns it)
          v3 nn = f.p0 != 0.f ? r.N : geometric_normal(sc, pos, prim, r.N);
          v3 c = nn * 0.5f + mk3(0.5f);
          res = make_float4(c.x, c.y, c.z, __uint_as_float(depth + 1));
        } else if (is_light) {
          // diffuse_light.cu:43-47 and raygen.cu:85-90
          if (dot(r.N, r.Wo) < 0.f) T = T * f.c0;
          if (depth == 0 || (flags & 1u)) L = L + T;
          res = make_float4(L.x, L.y, L.z, __uint_as_float(depth + 1));
        } else {
          v3 ndir;
          alive = scatter_and_roulette<MS>(f, r.Wo, r.N, seed, sig, T, L, ndir, flags, prev_pdf, depth, rc, res);
          if (alive) {
            no = make_float4(r.P.x, r.P.y, r.P.z, __uint_as_float(MULTI ? ((seed & 0xFFFFFFu) | (sid << 24)) : seed));
            nd = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float((dw & ~SPEC_BIT) | (flags << 31)));
            nT = make_float4(T.x, T.y, T.z, prev_pdf);
            nL = make_float4(L.x, L.y, L.z, __uint_as_float(sig));
          }
        }
      }
      if (MULTI && alive && depth + 1 >= rc.max_depth) {
        // the last bounce of this path's scene: it ends like the paths that
        // outlive a single scene's loop (k_exhausted), result left as the
        // path start set it
        alive = false;
        exhausted++;
        P.sig_final[dw & SLOT_MASK] = sig;
      } else if (!alive) {
        P.result[dw & SLOT_MASK] = res;
        P.sig_final[dw & SLOT_MASK] = sig;
      }
    }
    // compaction of the survivors into this wave's segment of set `out`
    // (never more than the wave's inputs, so every position exists)
    const uint64_t mask = __ballot(alive);
    if (alive) {
      const uint32_t k = g.pos(survivors + lane_rank(mask));
      D.o[k] = no;
      D.d[k] = nd;
      D.T[k] = nT;
      D.L[k] = nL;
    }
    survivors += (uint32_t)__popcll(mask);
    if (cache_on) {
      tcount = min(tcount + (uint32_t)__popcll(__ballot(did_rec)), tcap);
      qcount += (uint32_t)__popcll(__ballot(did_ask));
    }
  }
  uint32_t a = tc.nodes, b = tc.prims, c = tc.insts;
  for (int off = 32; off > 0; off >>= 1) {
    shadow_rays += __shfl_down(shadow_rays, off);
    if (MULTI) exhausted += __shfl_down(exhausted, off);
    if (COUNT) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      c += __shfl_down(c, off);
    }
  }
  if (__lane_id() == 0) {
    P.seg[(size_t)out * MAX_WAVES + g.w] = survivors;
    if (cache_on) {
      P.tseg[g.w] = tcount;
      if (depth == rc0.cache.query_depth) P.qseg[g.w] = qcount;
    }
    stat_add(stats, g, STAT_SHADOW, shadow_rays);
    if (MULTI) stat_add(stats, g, STAT_EXHAUSTED, exhausted);
    if (COUNT) {
      stat_add(stats, g, STAT_SH_NODES, a);
      stat_add(stats, g, STAT_SH_PRIMS, b);
      stat_add(stats, g, STAT_SH_INSTS, c);
    }
  }
}

// Paths alive after the last bounce end with color() = 0 (raygen.cu:121-122);
// only their signature is recorded.
__global__ __launch_bounds__(BLOCK) void k_exhausted(DevPaths P, int in, uint64_t* stats) {
  const Seg g = my_seg();
  const uint32_t n = seg_count(P, in, g);
  const PathState S = P.st[in];
  if (__lane_id() == 0) stat_add(stats, g, STAT_EXHAUSTED, n);
  for (uint32_t j = __lane_id(); j < n; j += 64) {
    const uint32_t i = g.pos(j);
    P.sig_final[__float_as_uint(S.d[i].w) & SLOT_MASK] = __float_as_uint(S.L[i].w);
  }
}

// Sum of the per-wave statistics rows of a render (one block per word).
__global__ __launch_bounds__(BLOCK) void k_stats_reduce(const uint64_t* __restrict__ part, uint64_t* __restrict__ total) {
  __shared__ unsigned long long red[BLOCK];
  unsigned long long v = 0;
  for (uint32_t w = threadIdx.x; w < (uint32_t)MAX_WAVES; w += BLOCK) v += part[(size_t)w * STAT_WORDS + blockIdx.x];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = BLOCK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) total[blockIdx.x] = red[0];
}

// Radiance-cache records after a batch: exclusive prefix sums of the
// per-wave training and query counts (waves of the batch's grid) and their
// totals.  One block of SCAN_THREADS threads.
__global__ __launch_bounds__(SCAN_THREADS) void k_cache_scan(DevPaths P, uint32_t waves) {
  __shared__ uint32_t part[SCAN_THREADS];
  const uint32_t per = (waves + SCAN_THREADS - 1) / SCAN_THREADS;
  const uint32_t e0 = min(threadIdx.x * per, waves), e1 = min(e0 + per, waves);
  for (int which = 0; which < 2; which++) {
    const uint32_t* cnt = which == 0 ? P.tseg : P.qseg;
    uint32_t* pre = which == 0 ? P.tpre : P.qpre;
    uint32_t sum = 0;
    for (uint32_t e = e0; e < e1; e++) sum += cnt[e];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
      uint32_t v = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t e = e0; e < e1; e++) {
      pre[e] = run;
      run += cnt[e];
    }
    if (threadIdx.x == SCAN_THREADS - 1) P.counters[which == 0 ? CTR_TRAIN : CTR_QUERY] = part[SCAN_THREADS - 1];
    __syncthreads();
  }
}

// Dense copies of `rows` records (which 0: training vertices, 1: queries)
// spread evenly over the batch's `total` records (all of them when rows ==
// total), row r taking the record of rank (r * mul) mod rows (mul coprime to
// rows: with mul ~ 0.618 rows every run of rows, such as one Adam batch,
// samples the whole image instead of a few waves' pixels): features, and
// for training vertices the target
//   radiance leaving the vertex toward the camera side
//     = (L_final - L_before) / T
// per channel, zero where T is zero or the value is not finite.  vals (9
// per row: L, T, target for training vertices; T, 0, network output for
// queries) and slots are optional.
__global__ __launch_bounds__(BLOCK) void k_cache_gather(DevPaths P, int which, uint32_t rows, uint32_t mul,
                                                        uint32_t half_step, uint32_t waves, uint32_t stride,
                                                        uint4* __restrict__ feat, float* __restrict__ tgt,
                                                        float* __restrict__ vals, uint32_t* __restrict__ slots) {
  const uint32_t total = P.counters[which == 0 ? CTR_TRAIN : CTR_QUERY];
  const uint32_t* pre = which == 0 ? P.tpre : P.qpre;
  // half_step = k: ranks move by total / (2 k) (records between those a
  // gather of k rows takes; the same ones when total < 2 k)
  const uint32_t shift = half_step ? total / (2u * half_step) : 0u;
  for (uint32_t r = blockIdx.x * BLOCK + threadIdx.x; r < rows && r < total; r += gridDim.x * BLOCK) {
    const uint32_t rank = (uint32_t)((uint64_t)r * mul % rows);
    uint32_t idx = rows >= total ? rank : (uint32_t)((uint64_t)rank * total / rows);
    idx = (uint32_t)(((uint64_t)idx + shift) % total);
    uint32_t lo = 0, hi = waves - 1;  // last wave whose prefix is <= idx
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= idx) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t j = idx - pre[lo];
    const size_t k = (size_t)(j >> 6) * stride + (size_t)lo * 64 + (j & 63u);
    const uint4* src = (which == 0 ? P.t_feat : P.q_feat) + k * 12;
    for (int q = 0; q < 12; q++) feat[(size_t)r * 12 + q] = src[q];
    float out[9];
    uint32_t slot;
    if (which == 0) {
      const float4 a = P.t_L[k], t = P.t_T[k];
      slot = __float_as_uint(a.w) & SLOT_MASK;
      const float4 fin = P.result[slot];
      const float L0[3] = {a.x, a.y, a.z}, T0[3] = {t.x, t.y, t.z}, Lf[3] = {fin.x, fin.y, fin.z};
      for (int c = 0; c < 3; c++) {
        float v = T0[c] > 0.f ? (Lf[c] - L0[c]) / T0[c] : 0.f;
        v = (v > 0.f && v < 3.0e38f) ? v : 0.f;
        out[c] = L0[c];
        out[3 + c] = T0[c];
        out[6 + c] = v;
      }
      if (tgt)
        for (int c = 0; c < 3; c++) tgt[(size_t)r * 3 + c] = out[6 + c];
    } else {
      const float4 m = P.q_meta[k];
      slot = __float_as_uint(m.w) & SLOT_MASK;
      out[0] = m.x;
      out[1] = m.y;
      out[2] = m.z;
      out[3] = out[4] = out[5] = 0.f;
      for (int c = 0; c < 3; c++) out[6 + c] = P.q_out[k * 3 + c];
    }
    if (vals)
      for (int c = 0; c < 9; c++) vals[(size_t)r * 9 + c] = out[c];
    if (slots) slots[r] = slot;
  }
}

// The queries' answers: result += T * cache_radiance(y) per channel, over
// each wave's query segment (same grid as the batch).
__global__ __launch_bounds__(BLOCK) void k_cache_apply(DevPaths P) {
  const Seg g = my_seg();
  const uint32_t n = uniform(P.qseg[g.w]);
  for (uint32_t j = __lane_id(); j < n; j += 64) {
    const uint32_t i = g.pos(j);
    float4 m = P.q_meta[i];
    const uint32_t slot = __float_as_uint(m.w) & SLOT_MASK;
    const float y0 = cache_radiance(P.q_out[3 * (size_t)i]), y1 = cache_radiance(P.q_out[3 * (size_t)i + 1]),
                y2 = cache_radiance(P.q_out[3 * (size_t)i + 2]);
    float4 r = P.result[slot];
    r.x += m.x * y0;
    r.y += m.y * y1;
    r.z += m.z * y2;
    P.result[slot] = r;
  }
}

// The cache's per-batch judge, on n training vertices the plain tracer
// finished (vals: 9 per row, T at 3..5 and the unbiased radiance sample t
// at 6..8) and the network's outputs y for them:
//   res[0] = sum T^2 (cache_radiance(y) - t)^2   squared error the answers
//                                                would add to the image
//   res[1] = sum T^2 t^2                          the same for answering 0
//                                                (ending the paths there)
//   res[2] = outputs that are not finite
// per channel, summed in FP64 in a fixed order (one block).
__global__ __launch_bounds__(BLOCK) void k_cache_judge(const float* __restrict__ vals, const float* __restrict__ y,
                                                       uint32_t n, double* __restrict__ res) {
  __shared__ double s_num[BLOCK], s_den[BLOCK], s_bad[BLOCK];
  double num = 0.0, den = 0.0, bad = 0.0;
  for (uint32_t r = threadIdx.x; r < n; r += BLOCK) {
    for (int c = 0; c < 3; c++) {
      const float yc = y[3 * (size_t)r + c];
      const double T = vals[9 * (size_t)r + 3 + c], t = vals[9 * (size_t)r + 6 + c];
      if (!(fabsf(yc) <= 3.0e38f)) bad += 1.0;
      const double e = (double)cache_radiance(yc) - t;
      num += T * T * e * e;
      den += T * T * t * t;
    }
  }
  s_num[threadIdx.x] = num;
  s_den[threadIdx.x] = den;
  s_bad[threadIdx.x] = bad;
  __syncthreads();
  for (int off = BLOCK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      s_num[threadIdx.x] += s_num[threadIdx.x + off];
      s_den[threadIdx.x] += s_den[threadIdx.x + off];
      s_bad[threadIdx.x] += s_bad[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[0] = s_num[0];
    res[1] = s_den[0];
    res[2] = s_bad[0];
  }
}

// raygen.cu:126-134 de_nan and raygen.cu:166 acc_buffer += (rgb, 1), in FP64
// and frame order; accumulation row 0 is the top of the image (raygen.cu:153).
// Paths with a NaN channel are counted into the wave's statistics row.
__global__ __launch_bounds__(BLOCK) void k_accumulate(DevPaths P, RenderConsts rc, uint32_t nframes, double* accum,
                                                      uint64_t* stats) {
  uint32_t nans = 0;
  for (uint32_t pix = blockIdx.x * BLOCK + threadIdx.x; pix < rc.WH; pix += gridDim.x * BLOCK) {
    uint32_t x = pix % rc.W, y = pix / rc.W;
    size_t o = 3 * ((size_t)(rc.H - 1 - y) * rc.W + x);
    double a0 = accum[o], a1 = accum[o + 1], a2 = accum[o + 2];
    for (uint32_t f = 0; f < nframes; f++) {
      float4 r = P.result[(size_t)f * rc.WH + pix];
      const bool ok = (r.x == r.x) & (r.y == r.y) & (r.z == r.z);
      nans += ok ? 0u : 1u;
      float cx = (r.x == r.x) ? r.x : 0.f, cy = (r.y == r.y) ? r.y : 0.f, cz = (r.z == r.z) ? r.z : 0.f;
      a0 += (double)cx;
      a1 += (double)cy;
      a2 += (double)cz;
    }
    accum[o] = a0;
    accum[o + 1] = a1;
    accum[o + 2] = a2;
  }
  for (int d = 32; d >= 1; d >>= 1) nans += __shfl_xor(nans, d);
  if (__lane_id() == 0) stat_add(stats, my_seg(), STAT_NAN, nans);
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
#define PT_TRACE(C, V) \
  hipLaunchKernelGGL((k_trace_closest<C, V>), dim3(grid), dim3(BLOCK), 0, st, sc, rays, n, hits, counts, spill, overflow)
  const bool vol = sc.ext_instances != 0;
  if (counts) {
    if (vol) PT_TRACE(true, true);
    else PT_TRACE(true, false);
  } else {
    if (vol) PT_TRACE(false, true);
    else PT_TRACE(false, false);
  }
#undef PT_TRACE
}
void launch_trace_occluded(const DevScene& sc, const pt_ray* rays, uint64_t n, uint8_t* occ, uint2* spill,
                           uint32_t* overflow, int grid, hipStream_t st) {
  if (sc.ext_instances != 0)
    hipLaunchKernelGGL(k_trace_occluded<true>, dim3(grid), dim3(BLOCK), 0, st, sc, rays, n, occ, spill, overflow);
  else
    hipLaunchKernelGGL(k_trace_occluded<false>, dim3(grid), dim3(BLOCK), 0, st, sc, rays, n, occ, spill, overflow);
}
__global__ void k_set_frame(uint32_t* f, uint32_t v) { *f = v; }
__global__ void k_next_frames(uint32_t* f, uint32_t step) { *f += step; }
void launch_set_frame(uint32_t* frame_dev, uint32_t value, hipStream_t st) {
  hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, st, frame_dev, value);
}
void launch_next_frames(uint32_t* frame_dev, uint32_t step, hipStream_t st) {
  hipLaunchKernelGGL(k_next_frames, dim3(1), dim3(1), 0, st, frame_dev, step);
}
// instantiation for the scene's kinds: rects + boxes (the Cornell box),
// rects + meshes (meshes on a ground rect), spheres, analytic, all
static int extend_kinds(uint32_t k) {
  if (k == 6u || k == 10u || k == 1u) return (int)k;
  if ((k & 8u) == 0) return 7;
  return KIND_ALL;
}
void launch_extend(const DevScene& sc, const DevPaths& P, int in, float tmin, uint2* spill, uint64_t* stats,
                   bool count_traversal, bool sort, int grid, hipStream_t st, const DevScene* scenes,
                   const RenderConsts* start, uint32_t nstart, bool multi_ext, const BatchStart* bstart) {
  const RenderConsts rc = start ? *start : RenderConsts{};
  const bool starts = start != nullptr && scenes == nullptr && nstart != 0;
  const bool bstarts = bstart != nullptr && scenes != nullptr && nstart != 0 && bstart->nitems != 0;
  const BatchStart bs = bstarts ? *bstart : BatchStart{};
#define PT_EXTEND(C, S, M, E, K, ST)                                                                                \
  hipLaunchKernelGGL((k_extend<C, S, M, E, K, ST>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, tmin, spill, stats, \
                     scenes, rc, nstart, bs)
#define PT_EXTEND4(M, E, K, ST)                                          \
  do {                                                                   \
    if (count_traversal && sort) PT_EXTEND(true, true, M, E, K, ST);     \
    else if (count_traversal) PT_EXTEND(true, false, M, E, K, ST);       \
    else if (sort) PT_EXTEND(false, true, M, E, K, ST);                  \
    else PT_EXTEND(false, false, M, E, K, ST);                           \
  } while (0)
// the camera-ray pass never sorts (camera hits are shaded in pixel order)
#define PT_EXTEND2(M, E, K, ST)                                     \
  do {                                                              \
    if (count_traversal) PT_EXTEND(true, false, M, E, K, ST);       \
    else PT_EXTEND(false, false, M, E, K, ST);                      \
  } while (0)
#define PT_EXTEND_KINDS(ST, X)                                                                           \
  do {                                                                                                   \
    if (sc.ext_instances != 0) { /* volumes, cylinders, moving spheres */                              \
      X(false, true, KIND_ALL, ST);                                                                      \
    } else {                                                                                             \
      switch (extend_kinds(sc.kinds)) {                                                                  \
        case 6: X(false, false, 6, ST); break;                                                           \
        case 10: X(false, false, 10, ST); break;                                                         \
        case 1: X(false, false, 1, ST); break;                                                           \
        case 7: X(false, false, 7, ST); break;                                                           \
        default: X(false, false, KIND_ALL, ST);                                                          \
      }                                                                                                  \
    }                                                                                                    \
  } while (0)
  if (bstarts && multi_ext) {
    PT_EXTEND2(true, true, KIND_ALL, true);
  } else if (bstarts) {
    PT_EXTEND2(true, false, KIND_ALL, true);
  } else if (scenes && multi_ext) {  // a batch holding volumes, cylinders or moving spheres
    PT_EXTEND4(true, true, KIND_ALL, false);
  } else if (scenes) {
    PT_EXTEND4(true, false, KIND_ALL, false);
  } else if (starts) {
    PT_EXTEND_KINDS(true, PT_EXTEND2);
  } else {
    PT_EXTEND_KINDS(false, PT_EXTEND4);
  }
#undef PT_EXTEND_KINDS
#undef PT_EXTEND2
#undef PT_EXTEND4
#undef PT_EXTEND
}
void launch_sort(const DevPaths& P, int in, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(SCAN_THREADS), 0, st, P, (uint32_t)(grid * WAVES));
  hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(BLOCK), 0, st, P, in);
}
template <int MS, int FEAT, int KINDS = KIND_ALL>
static void launch_shade_v(const DevScene& sc, const DevPaths& P, int in, int src, const float4* hits, int out,
                           uint32_t depth, const RenderConsts& rc, uint2* spill, uint64_t* stats, int grid,
                           hipStream_t st, const DevScene* scenes, const RenderConsts* rcs) {
  hipLaunchKernelGGL((k_shade<MS, FEAT, KINDS>), dim3(grid), dim3(BLOCK), 0, st, sc, P, in, src, hits, out, depth, rc, spill,
                     stats, scenes, rcs);
}
void launch_shade(const DevScene& sc, const DevPaths& P, int in, bool sorted, int out, uint32_t depth,
                  const RenderConsts& rc, int matset, bool count, uint2* spill, uint64_t* stats, int grid,
                  hipStream_t st, const DevScene* scenes, const RenderConsts* rcs, bool batch_env, bool batch_ext) {
  const int src = sorted ? 2 : in;
  const float4* hits = sorted ? P.hit_sorted : P.hit;
  const bool env = (rc.flags & PT_FLAG_ENV_MIS) && sc.num_lights != sc.num_lights_plain;
  const bool cache = (rc.flags & PT_FLAG_RADIANCE_CACHE) != 0;
#define PT_SHADE(MSV, F) \
  launch_shade_v<MSV, F>(sc, P, in, src, hits, out, depth, rc, spill, stats, grid, st, scenes, rcs)
#define PT_SHADE_K(MSV, F, K) \
  launch_shade_v<MSV, F, K>(sc, P, in, src, hits, out, depth, rc, spill, stats, grid, st, scenes, rcs)
  if (scenes && batch_ext) {  // a batch with media, cylinders, moving spheres or the normal shader
    if (batch_env) {
      if (count) PT_SHADE(MS_FULL, SH_MULTI | SH_EXT | SH_ENV | SH_COUNT);
      else PT_SHADE(MS_FULL, SH_MULTI | SH_EXT | SH_ENV);
    } else {
      if (count) PT_SHADE(MS_FULL, SH_MULTI | SH_EXT | SH_COUNT);
      else PT_SHADE(MS_FULL, SH_MULTI | SH_EXT);
    }
  } else if (scenes) {  // a batch of scenes: every material, per-path scene
    if (batch_env) {
      if (count) PT_SHADE(MS_ALL, SH_MULTI | SH_ENV | SH_COUNT);
      else PT_SHADE(MS_ALL, SH_MULTI | SH_ENV);
    } else {
      if (count) PT_SHADE(MS_ALL, SH_MULTI | SH_COUNT);
      else PT_SHADE(MS_ALL, SH_MULTI);
    }
  } else if (matset == MS_FULL || sc.ext_instances != 0) {
    // media or the normal shader (single scenes, no radiance cache)
    if (env) {
      if (count) PT_SHADE(MS_FULL, SH_EXT | SH_ENV | SH_COUNT);
      else PT_SHADE(MS_FULL, SH_EXT | SH_ENV);
    } else {
      if (count) PT_SHADE(MS_FULL, SH_EXT | SH_COUNT);
      else PT_SHADE(MS_FULL, SH_EXT);
    }
  } else if (cache) {  // every material; no traversal counting
    if (env) PT_SHADE(MS_ALL, SH_CACHE | SH_ENV);
    else PT_SHADE(MS_ALL, SH_CACHE);
  } else if (env && sc.kinds == 10u) {  // meshes on a ground rect
    if (count) PT_SHADE_K(MS_ALL, SH_ENV | SH_COUNT, 10);
    else PT_SHADE_K(MS_ALL, SH_ENV, 10);
  } else if (env) {
    if (count) PT_SHADE(MS_ALL, SH_ENV | SH_COUNT);
    else PT_SHADE(MS_ALL, SH_ENV);
  } else if (matset == MS_DIFFUSE && (sc.kinds & 8u) == 0) {
    // analytic only (a sphere-free k_shade allocates its registers worse)
    if (count) PT_SHADE_K(MS_DIFFUSE, SH_COUNT, 7);
    else PT_SHADE_K(MS_DIFFUSE, 0, 7);
  } else if (matset == MS_SPHERES && (sc.kinds & 8u) == 0) {
    if (count) PT_SHADE_K(MS_SPHERES, SH_COUNT, 7);
    else PT_SHADE_K(MS_SPHERES, 0, 7);
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
#undef PT_SHADE_K
#undef PT_SHADE
}
int shade_material_set(uint32_t types_mask) {
  if ((types_mask & ~(uint32_t)MS_DIFFUSE) == 0) return MS_DIFFUSE;
  if ((types_mask & ~(uint32_t)MS_SPHERES) == 0) return MS_SPHERES;
  if ((types_mask & ~(uint32_t)MS_ALL) == 0) return MS_ALL;
  return MS_FULL;
}
void launch_cache_scan(const DevPaths& P, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_cache_scan, dim3(1), dim3(SCAN_THREADS), 0, st, P, (uint32_t)(grid * WAVES));
}
void launch_cache_gather(const DevPaths& P, int which, uint32_t rows, uint32_t mul, int grid, uint4* feat, float* tgt,
                         float* vals, uint32_t* slots, hipStream_t st, uint32_t half_step) {
  const int g = (int)std::max<uint32_t>(1, std::min<uint32_t>((rows + BLOCK - 1) / BLOCK, 2048));
  hipLaunchKernelGGL(k_cache_gather, dim3(g), dim3(BLOCK), 0, st, P, which, rows, mul, half_step,
                     (uint32_t)(grid * WAVES), (uint32_t)(grid * BLOCK), feat, tgt, vals, slots);
}
void launch_cache_judge(const float* vals, const float* y, uint32_t n, double* res, hipStream_t st) {
  hipLaunchKernelGGL(k_cache_judge, dim3(1), dim3(BLOCK), 0, st, vals, y, n, res);
}
void launch_cache_apply(const DevPaths& P, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_cache_apply, dim3(grid), dim3(BLOCK), 0, st, P);
}
void launch_exhausted(const DevPaths& P, int in, uint64_t* stats, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_exhausted, dim3(grid), dim3(BLOCK), 0, st, P, in, stats);
}
void launch_stats_reduce(const uint64_t* part, uint64_t* total, hipStream_t st) {
  hipLaunchKernelGGL(k_stats_reduce, dim3(STAT_WORDS), dim3(BLOCK), 0, st, part, total);
}
void launch_accumulate(const DevPaths& P, const RenderConsts& rc, uint32_t nframes, double* accum, uint64_t* stats,
                       int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(BLOCK), 0, st, P, rc, nframes, accum, stats);
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
