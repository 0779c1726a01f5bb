// pt_bvh.hip -- GPU LBVH build, replacing OptiX's createAcceleration("Trbvh")
// (host_includes/scenes.hpp:31-32, transforms.hpp:59) and the bounds
// programs it runs (sphere.cu:96-100, aarect.cu:117-136, box.cu:84-88,
// triangle.cu:61-83).
//
//   bounds -> centroid bounds -> 30-bit Morton codes -> stable LSD radix sort
//   (8-bit digits, LDS histograms, wave64 ballot ranking) -> Karras 2012
//   binary hierarchy -> bottom-up bounds -> breadth-first collapse to 4-wide
//   128-byte nodes.
//
// Every step is integer work or exact min/max, so the result is the same
// bit for bit on every run and equals a sequential CPU build.
#include "pt_device.hpp"
#include "pt_kernels.hpp"

namespace pt {

constexpr int BB = 256;              // block size of the build kernels
constexpr int SORT_ROUNDS = 8;       // keys per thread per tile
constexpr int TILE = BB * SORT_ROUNDS;

PTD uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
PTD float ord2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u); }

static inline int grid_for(uint64_t n) { return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + BB - 1) / BB, 4096)); }

// ------------------------------------------------------------- bounds
// Each box is widened outward by x -+ (|x| 2^-23 + 2^-126), at least one ulp,
// so that a ray lying exactly in a bounding plane with a zero direction
// component never meets 0 * inf in the slab test and no primitive whose
// (inclusive) intersection test accepts it is culled.
PTD float widen_lo(float x) { return x - (fabsf(x) * 0x1p-23f + 0x1p-126f); }
PTD float widen_hi(float x) { return x + (fabsf(x) * 0x1p-23f + 0x1p-126f); }
PTD void put_box(float4* lo, float4* hi, uint32_t i, float lx, float ly, float lz, float hx, float hy, float hz) {
  lo[i] = make_float4(widen_lo(lx), widen_lo(ly), widen_lo(lz), 0.f);
  hi[i] = make_float4(widen_hi(hx), widen_hi(hy), widen_hi(hz), 0.f);
}
__global__ void k_bounds_spheres(const float4* sph, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 s = sph[i];
    put_box(lo, hi, i, s.x - s.w, s.y - s.w, s.z - s.w, s.x + s.w, s.y + s.w, s.z + s.w);
  }
}
__global__ void k_bounds_rects(const float4* rect, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 A = rect[2 * i], B = rect[2 * i + 1];
    int axis = (int)__float_as_uint(B.y);
    float k = B.x;
    if (axis == 0) put_box(lo, hi, i, k - 0.0001f, A.x, A.z, k + 0.0001f, A.y, A.w);
    else if (axis == 1) put_box(lo, hi, i, A.x, k - 0.0001f, A.z, A.y, k + 0.0001f, A.w);
    else put_box(lo, hi, i, A.x, A.z, k - 0.0001f, A.y, A.w, k + 0.0001f);
  }
}
__global__ void k_bounds_boxes(const float4* box, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 a = box[2 * i], b = box[2 * i + 1];
    put_box(lo, hi, i, a.x - 0.001f, a.y - 0.001f, a.z - 0.001f, b.x + 0.001f, b.y + 0.001f, b.z + 0.001f);
  }
}
// volume_sphere.cu:111-115 (center -+ radius), volume_box.cu:90-94 (-+1e-4)
__global__ void k_bounds_volumes(const float4* vol, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 a = vol[2 * i], b = vol[2 * i + 1];
    if (__float_as_uint(b.w) >> 31)
      put_box(lo, hi, i, a.x - 0.0001f, a.y - 0.0001f, a.z - 0.0001f, b.x + 0.0001f, b.y + 0.0001f, b.z + 0.0001f);
    else
      put_box(lo, hi, i, a.x - b.x, a.y - b.x, a.z - b.x, a.x + b.x, a.y + b.x, a.z + b.x);
  }
}
// cylinder.cu:100-107: origin and origin + length (0, 1, 0), -+ radius
__global__ void k_bounds_cylinders(const float4* cyl, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 a = cyl[2 * i], b = cyl[2 * i + 1];
    float ox = a.x + a.w * 0.f, oy = a.y + a.w * 1.f, oz = a.z + a.w * 0.f;
    put_box(lo, hi, i, fminf(a.x, ox) - b.x, fminf(a.y, oy) - b.x, fminf(a.z, oz) - b.x, fmaxf(a.x, ox) + b.x,
            fmaxf(a.y, oy) + b.x, fmaxf(a.z, oz) + b.x);
  }
}
// moving_sphere.cu:137-147: the union of the boxes at time0 and time1
__global__ void k_bounds_moving(const float4* mov, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 a = mov[3 * i], b = mov[3 * i + 1], c = mov[3 * i + 2];
    v3 p = moving_center(a, b, c, c.x), q = moving_center(a, b, c, c.y);
    const float r = a.w;
    put_box(lo, hi, i, fminf(p.x - r, q.x - r), fminf(p.y - r, q.y - r), fminf(p.z - r, q.z - r),
            fmaxf(p.x + r, q.x + r), fmaxf(p.y + r, q.y + r), fmaxf(p.z + r, q.z + r));
  }
}
__global__ void k_bounds_tris(const float4* vtx, const uint4* idx, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    uint4 t = idx[i];
    float4 a = vtx[t.x], b = vtx[t.y], c = vtx[t.z];
    put_box(lo, hi, i, fminf(fminf(a.x, b.x), c.x) - 0.0001f, fminf(fminf(a.y, b.y), c.y) - 0.0001f,
            fminf(fminf(a.z, b.z), c.z) - 0.0001f, fmaxf(fmaxf(a.x, b.x), c.x) + 0.0001f,
            fmaxf(fmaxf(a.y, b.y), c.y) + 0.0001f, fmaxf(fmaxf(a.z, b.z), c.z) + 0.0001f);
  }
}
// world bounds of an instance: its structure's root box, 8 corners through
// the object-to-world matrix (identity instances keep the root box)
__global__ void k_bounds_instances(const uint4* insts, const float4* blas_root, uint32_t n, float4* lo, float4* hi) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    const uint4* ir = insts + (size_t)i * 8;
    uint32_t blas = ir[7].w;
    float4 rl = blas_root[2 * blas], rh = blas_root[2 * blas + 1];
    if ((ir[3].z >> 8) == 0) {
      lo[i] = rl;
      hi[i] = rh;
      continue;
    }
    uint4 q4 = ir[4], q5 = ir[5], q6 = ir[6];
    float4 m0 = make_float4(__uint_as_float(q4.x), __uint_as_float(q4.y), __uint_as_float(q4.z), __uint_as_float(q4.w));
    float4 m1 = make_float4(__uint_as_float(q5.x), __uint_as_float(q5.y), __uint_as_float(q5.z), __uint_as_float(q5.w));
    float4 m2 = make_float4(__uint_as_float(q6.x), __uint_as_float(q6.y), __uint_as_float(q6.z), __uint_as_float(q6.w));
    v3 wl = mk3(FINF), wh = mk3(-FINF);
    for (int c = 0; c < 8; c++) {
      v3 p = mk3((c & 1) ? rh.x : rl.x, (c & 2) ? rh.y : rl.y, (c & 4) ? rh.z : rl.z);
      v3 q = xf_point(m0, m1, m2, p);
      wl = mk3(fminf(wl.x, q.x), fminf(wl.y, q.y), fminf(wl.z, q.z));
      wh = mk3(fmaxf(wh.x, q.x), fmaxf(wh.y, q.y), fmaxf(wh.z, q.z));
    }
    put_box(lo, hi, i, wl.x, wl.y, wl.z, wh.x, wh.y, wh.z);
  }
}

// ------------------------------------------------------- centroid bounds
__global__ void k_init_cbounds(uint32_t* cb) {
  if (threadIdx.x < 3) cb[threadIdx.x] = 0xFFFFFFFFu;       // min
  else if (threadIdx.x < 6) cb[threadIdx.x] = 0u;           // max
}
__global__ void k_centroid_bounds(const float4* lo, const float4* hi, uint32_t n, uint32_t* cb) {
  __shared__ uint32_t s[6][BB];
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 a = lo[i], b = hi[i];
    uint32_t c[3] = {f2ord((a.x + b.x) * 0.5f), f2ord((a.y + b.y) * 0.5f), f2ord((a.z + b.z) * 0.5f)};
    for (int k = 0; k < 3; k++) { mn[k] = min(mn[k], c[k]); mx[k] = max(mx[k], c[k]); }
  }
  for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
  __syncthreads();
  for (int w = BB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = min(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        s[3 + k][threadIdx.x] = max(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) { atomicMin(&cb[k], s[k][0]); atomicMax(&cb[3 + k], s[3 + k][0]); }
}

// ---------------------------------------------------------- Morton codes
PTD uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
PTD uint32_t quantize10(float c, float lo, float scale) {
  float f = (c - lo) * scale;
  return f >= 1023.0f ? 1023u : (uint32_t)f;
}
__global__ void k_morton(const float4* lo, const float4* hi, uint32_t n, const uint32_t* cb, uint32_t* code,
                         uint32_t* idx) {
  float lx = ord2f(cb[0]), ly = ord2f(cb[1]), lz = ord2f(cb[2]);
  float ex = ord2f(cb[3]) - lx, ey = ord2f(cb[4]) - ly, ez = ord2f(cb[5]) - lz;
  float sx = ex > 0.0f ? 1024.0f / ex : 0.0f;
  float sy = ey > 0.0f ? 1024.0f / ey : 0.0f;
  float sz = ez > 0.0f ? 1024.0f / ez : 0.0f;
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) {
    float4 a = lo[i], b = hi[i];
    float cx = (a.x + b.x) * 0.5f, cy = (a.y + b.y) * 0.5f, cz = (a.z + b.z) * 0.5f;
    code[i] = (spread10(quantize10(cx, lx, sx)) << 2) | (spread10(quantize10(cy, ly, sy)) << 1) |
              spread10(quantize10(cz, lz, sz));
    idx[i] = i;
  }
}

// ------------------------------------------------------------------ scan
// Exclusive scan of uint32 in place; 4 items per thread, 1024 per block.
__global__ void k_scan_block(uint32_t* data, uint64_t n, uint32_t* block_sums) {
  __shared__ uint32_t s[BB];
  uint64_t base = (uint64_t)blockIdx.x * BB * 4 + threadIdx.x * 4;
  uint32_t v[4], sum = 0;
  for (int k = 0; k < 4; k++) {
    v[k] = (base + k < n) ? data[base + k] : 0u;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < BB; off <<= 1) {
    uint32_t t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t excl = s[threadIdx.x] - sum;
  for (int k = 0; k < 4; k++) {
    if (base + k < n) data[base + k] = excl;
    excl += v[k];
  }
  if (threadIdx.x == BB - 1 && block_sums) block_sums[blockIdx.x] = s[BB - 1];
}
__global__ void k_scan_add(uint32_t* data, uint64_t n, const uint32_t* block_sums) {
  uint64_t base = (uint64_t)blockIdx.x * BB * 4 + threadIdx.x * 4;
  uint32_t add = block_sums[blockIdx.x];
  for (int k = 0; k < 4; k++)
    if (base + k < n) data[base + k] += add;
}
// scratch needs ceil(n / 1024) + ceil(n / 1024^2) + ... words
static void exclusive_scan(uint32_t* data, uint64_t n, uint32_t* tmp, hipStream_t st) {
  uint64_t nb = (n + BB * 4 - 1) / (BB * 4);
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(BB), 0, st, data, n, (uint32_t*)nullptr);
    return;
  }
  hipLaunchKernelGGL(k_scan_block, dim3((uint32_t)nb), dim3(BB), 0, st, data, n, tmp);
  exclusive_scan(tmp, nb, tmp + nb, st);
  hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)nb), dim3(BB), 0, st, data, n, tmp);
}

// ------------------------------------------------------------ radix sort
__global__ void k_radix_hist(const uint32_t* code, uint32_t n, int shift, uint32_t nblk, uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * TILE;
  for (int r = 0; r < SORT_ROUNDS; r++) {
    uint32_t i = base + r * BB + threadIdx.x;
    if (i < n) atomicAdd(&h[(code[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}
// stable scatter: keys of a tile keep their order within a digit
__global__ void k_radix_scatter(const uint32_t* code_in, const uint32_t* idx_in, uint32_t* code_out,
                                uint32_t* idx_out, uint32_t n, int shift, uint32_t nblk, const uint32_t* offs) {
  __shared__ uint32_t base_d[256];
  __shared__ uint32_t wcnt[BB / 64][256];
  base_d[threadIdx.x] = offs[threadIdx.x * nblk + blockIdx.x];
  for (int w = 0; w < BB / 64; w++) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t wave = threadIdx.x / 64;
  const uint64_t lt = (threadIdx.x % 64) == 0 ? 0ull : (~0ull >> (64 - (threadIdx.x % 64)));
  for (int r = 0; r < SORT_ROUNDS; r++) {
    uint32_t i = blockIdx.x * TILE + r * BB + threadIdx.x;
    bool valid = i < n;
    uint32_t key = valid ? code_in[i] : 0u;
    uint32_t val = valid ? idx_in[i] : 0u;
    uint32_t d = (key >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base_d[d] + rank;
      for (uint32_t w = 0; w < wave; w++) pos += wcnt[w][d];
      code_out[pos] = key;
      idx_out[pos] = val;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int w = 0; w < BB / 64; w++) { tot += wcnt[w][threadIdx.x]; wcnt[w][threadIdx.x] = 0; }
    base_d[threadIdx.x] += tot;
    __syncthreads();
  }
}

// ------------------------------------------------------ Karras hierarchy
PTD int kdelta(const uint32_t* code, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j > n - 1) return -1;
  uint64_t ki = ((uint64_t)code[i] << 32) | (uint64_t)i;
  uint64_t kj = ((uint64_t)code[j] << 32) | (uint64_t)j;
  return __clzll(ki ^ kj);
}
__global__ void k_karras(const uint32_t* code, uint32_t n, uint32_t* left, uint32_t* right, uint32_t* first,
                         uint32_t* last, uint32_t* parent) {
  const int64_t nn = n;
  for (uint32_t ii = blockIdx.x * BB + threadIdx.x; ii + 1 < n; ii += gridDim.x * BB) {
    int64_t i = ii;
    int d = (kdelta(code, nn, i, i + 1) - kdelta(code, nn, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(code, nn, i, i - d);
    int64_t lmax = 2;
    while (kdelta(code, nn, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (kdelta(code, nn, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = kdelta(code, nn, i, j);
    int64_t s = 0, div = 2, t;
    do {
      t = (l + div - 1) / div;
      if (kdelta(code, nn, i, i + (s + t) * d) > dnode) s += t;
      div *= 2;
    } while (t > 1);
    int64_t gamma = i + s * d + min(d, 0);
    int64_t lo = min(i, j), hi = max(i, j);
    uint32_t lref = (lo == gamma) ? (LEAF_BIT | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t rref = (hi == gamma + 1) ? (LEAF_BIT | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    left[i] = lref;
    right[i] = rref;
    first[i] = (uint32_t)lo;
    last[i] = (uint32_t)hi;
    parent[(lref & LEAF_BIT) ? (n - 1 + (lref & ~LEAF_BIT)) : lref] = (uint32_t)i;
    parent[(rref & LEAF_BIT) ? (n - 1 + (rref & ~LEAF_BIT)) : rref] = (uint32_t)i;
  }
}

// Bottom-up bounds: the second child to arrive at a node merges both.
// A Karras node's index is one end of its leaf range, so a node whose range
// lies inside a block's span of REFIT_SPAN leaves is indexed inside that
// span: phase 1 (k_refit_local) climbs from every leaf of the span with
// arrival flags and boxes in LDS, stopping below the first node whose range
// leaves the span (that node's inside child is appended to a list); phase 2
// (k_refit_top, one workgroup) climbs from the listed children through the
// straddling nodes to the root with arrival flags in global memory.  Every
// node box is min/max of its left and right child's, so the result does not
// depend on who merges.  No device-scope fence is needed: LDS arrivals are
// ordered at workgroup scope, and phase 2 runs in one workgroup.
constexpr uint32_t REFIT_SPAN = 2048;
constexpr uint32_t NO_PARENT = 0xFFFFFFFFu;

PTD void leaf_box(uint32_t ref, const uint32_t* idx, const float4* plo, const float4* phi, float4& lo, float4& hi) {
  uint32_t p = idx[ref & ~LEAF_BIT];
  lo = plo[p];
  hi = phi[p];
}
PTD void merge(const float4& al, const float4& ah, const float4& bl, const float4& bh, float4& lo, float4& hi) {
  lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.f);
  hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
}

__global__ void __launch_bounds__(BB) k_refit_local(uint32_t n, const uint32_t* idx, const float4* plo,
                                                    const float4* phi, const uint32_t* left, const uint32_t* right,
                                                    const uint32_t* first, const uint32_t* last,
                                                    const uint32_t* parent, float4* nlo, float4* nhi,
                                                    uint32_t* list, uint32_t* list_count) {
  __shared__ uint32_t arrived[REFIT_SPAN];
  __shared__ float4 blo[REFIT_SPAN], bhi[REFIT_SPAN];
  const uint32_t base = blockIdx.x * REFIT_SPAN;
  const uint32_t end = min(n, base + REFIT_SPAN);
  for (uint32_t k = threadIdx.x; k < REFIT_SPAN; k += BB) arrived[k] = 0u;
  __syncthreads();
  auto inside = [&](uint32_t node) { return first[node] >= base && last[node] < end; };
  for (uint32_t l = base + threadIdx.x; l < end; l += BB) {
    uint32_t child = LEAF_BIT | l;
    uint32_t node = parent[n - 1 + l];
    for (int climb = 0; climb < 4096 && node != NO_PARENT; climb++) {
      if (!inside(node)) {
        list[atomicAdd(list_count, 1u)] = child;
        break;
      }
      __threadfence_block();
      if (atomicAdd(&arrived[node - base], 1u) == 0u) break;
      __threadfence_block();
      const uint32_t lr = left[node], rr = right[node];
      float4 al, ah, bl, bh, lo, hi;
      if (lr & LEAF_BIT) leaf_box(lr, idx, plo, phi, al, ah);
      else { al = blo[lr - base]; ah = bhi[lr - base]; }
      if (rr & LEAF_BIT) leaf_box(rr, idx, plo, phi, bl, bh);
      else { bl = blo[rr - base]; bh = bhi[rr - base]; }
      merge(al, ah, bl, bh, lo, hi);
      blo[node - base] = lo;
      bhi[node - base] = hi;
      nlo[node] = lo;
      nhi[node] = hi;
      child = node;
      node = parent[node];
    }
  }
}

__global__ void __launch_bounds__(1024) k_refit_top(uint32_t n, const uint32_t* idx, const float4* plo,
                                                    const float4* phi, const uint32_t* left, const uint32_t* right,
                                                    const uint32_t* parent, uint32_t* arrived, float4* nlo,
                                                    float4* nhi, const uint32_t* list, const uint32_t* list_count) {
  const uint32_t count = *list_count;
  for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
    const uint32_t start = list[k];
    uint32_t node = parent[(start & LEAF_BIT) ? (n - 1 + (start & ~LEAF_BIT)) : start];
    for (int climb = 0; climb < 4096 && node != NO_PARENT; climb++) {
      __threadfence_block();
      if (atomicAdd(&arrived[node], 1u) == 0u) break;
      __threadfence_block();
      float4 al, ah, bl, bh, lo, hi;
      const uint32_t refs[2] = {left[node], right[node]};
      for (int c = 0; c < 2; c++) {
        float4 l4, h4;
        if (refs[c] & LEAF_BIT) {
          leaf_box(refs[c], idx, plo, phi, l4, h4);
        } else {
          const volatile float* a = reinterpret_cast<const volatile float*>(nlo + refs[c]);
          const volatile float* b = reinterpret_cast<const volatile float*>(nhi + refs[c]);
          l4 = make_float4(a[0], a[1], a[2], 0.f);
          h4 = make_float4(b[0], b[1], b[2], 0.f);
        }
        if (c == 0) { al = l4; ah = h4; } else { bl = l4; bh = h4; }
      }
      merge(al, ah, bl, bh, lo, hi);
      volatile float* wl = reinterpret_cast<volatile float*>(nlo + node);
      volatile float* wh = reinterpret_cast<volatile float*>(nhi + node);
      wl[0] = lo.x; wl[1] = lo.y; wl[2] = lo.z; wl[3] = 0.f;
      wh[0] = hi.x; wh[1] = hi.y; wh[2] = hi.z; wh[3] = 0.f;
      node = parent[node];
    }
  }
}

// --------------------------------------------------- collapse to BVH4
struct BinTree {
  const uint32_t *left, *right, *first, *last, *idx;
  const float4 *plo, *phi, *nlo, *nhi;
};
PTD uint32_t ref_size(const BinTree& t, uint32_t r) {
  return (r & LEAF_BIT) ? 1u : (t.last[r] - t.first[r] + 1u);
}
PTD uint32_t ref_first(const BinTree& t, uint32_t r) { return (r & LEAF_BIT) ? (r & ~LEAF_BIT) : t.first[r]; }
PTD float half_area(float4 lo, float4 hi) {
  float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}
PTD void ref_box(const BinTree& t, uint32_t r, float4& lo, float4& hi) {
  if (r & LEAF_BIT) {
    uint32_t p = t.idx[r & ~LEAF_BIT];
    lo = t.plo[p];
    hi = t.phi[p];
  } else {
    lo = t.nlo[r];
    hi = t.nhi[r];
  }
}
// expansion: repeatedly open the expandable slot of largest half area
PTD uint32_t collapse_expand(const BinTree& t, uint32_t m, int leaf_max, uint32_t s[4]) {
  s[0] = t.left[m];
  s[1] = t.right[m];
  s[2] = s[3] = EMPTY_CHILD;
  int ns = 2;
  while (ns < 4) {
    int best = -1;
    float best_area = -1.0f;
    for (int k = 0; k < ns; k++) {
      uint32_t r = s[k];
      if ((r & LEAF_BIT) || ref_size(t, r) <= (uint32_t)leaf_max) continue;
      float a = half_area(t.nlo[r], t.nhi[r]);
      if (a > best_area) { best_area = a; best = k; }
    }
    if (best < 0) break;
    uint32_t r = s[best];
    for (int k = ns; k > best + 1; k--) s[k] = s[k - 1];
    s[best] = t.left[r];
    s[best + 1] = t.right[r];
    ns++;
  }
  uint32_t ni = 0;
  for (int k = 0; k < 4; k++)
    if (s[k] != EMPTY_CHILD && !(s[k] & LEAF_BIT) && ref_size(t, s[k]) > (uint32_t)leaf_max) ni++;
  return ni;
}
// one BVH4 node from its four slots; internal children are numbered from
// next_base + o in slot order and appended to the next frontier
PTD void collapse_emit(const BinTree& t, const uint32_t s[4], uint32_t o, int leaf_max, uint32_t next_base,
                       uint4* nd, uint32_t* next_frontier, uint32_t cap) {
  uint32_t wl[6][4], ch[4], cn[4];
  for (int k = 0; k < 4; k++) {
    uint32_t r = s[k];
    if (r == EMPTY_CHILD) {
      wl[0][k] = __float_as_uint(FINF); wl[1][k] = __float_as_uint(-FINF);
      wl[2][k] = __float_as_uint(FINF); wl[3][k] = __float_as_uint(-FINF);
      wl[4][k] = __float_as_uint(FINF); wl[5][k] = __float_as_uint(-FINF);
      ch[k] = EMPTY_CHILD;
      cn[k] = 0;
      continue;
    }
    float4 lo, hi;
    ref_box(t, r, lo, hi);
    wl[0][k] = __float_as_uint(lo.x); wl[1][k] = __float_as_uint(hi.x);
    wl[2][k] = __float_as_uint(lo.y); wl[3][k] = __float_as_uint(hi.y);
    wl[4][k] = __float_as_uint(lo.z); wl[5][k] = __float_as_uint(hi.z);
    uint32_t sz = ref_size(t, r);
    if ((r & LEAF_BIT) || sz <= (uint32_t)leaf_max) {
      ch[k] = ref_first(t, r);
      cn[k] = sz;
    } else {
      ch[k] = next_base + o;
      cn[k] = 0;
      if (o < cap) next_frontier[o] = r;
      o++;
    }
  }
  for (int q = 0; q < 6; q++) nd[q] = make_uint4(wl[q][0], wl[q][1], wl[q][2], wl[q][3]);
  nd[6] = make_uint4(ch[0], ch[1], ch[2], ch[3]);
  nd[7] = make_uint4(cn[0], cn[1], cn[2], cn[3]);
}

// Breadth-first collapse of the binary hierarchy into BVH4 nodes: expand
// each frontier node into up to four slots, exclusive-scan the numbers of
// internal children (which numbers the next level in frontier then slot
// order), emit the nodes.  The narrow top levels run in one workgroup
// (k_collapse_top: no launch or host round trip per level); once a level
// holds more than COLLAPSE_WIDE nodes the workgroup stops and the wide
// bottom levels (a few: each is about 3.5 times the last) run as multi-block
// expand / scan / emit launches.  The node order is the same either way.
// k_collapse_top's state: result[0] = 0 malformed (more than 128 levels or
// more than n nodes), 1 done, 2 continue; [1] frontier size, [2] level
// base, [3] next level base, [4] levels so far, [5] 1 if the frontier is in
// fb.
constexpr int CB = 1024;
constexpr uint32_t COLLAPSE_WIDE = 4096;
__global__ void __launch_bounds__(CB) k_collapse_top(BinTree t, uint32_t n, int leaf_max, uint32_t cap,
                                                     uint32_t* fa, uint32_t* fb, uint32_t* slots, uint32_t* nchild,
                                                     uint4* nodes, uint32_t* result) {
  __shared__ uint32_t wsum[CB / 64];
  const uint32_t tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  if (tid == 0) fa[0] = 0u;  // the binary root
  __syncthreads();
  uint32_t F = 1, level_base = 0, next_base = 1;
  uint32_t *cur = fa, *nxt = fb;
  bool ok = true;
  int levels = 0;
  while (F > 0 && F <= COLLAPSE_WIDE) {
    if (++levels > 128 || next_base > n || F > cap) {
      ok = false;
      break;
    }
    for (uint32_t w = tid; w < F; w += CB) {
      uint32_t sl[4];
      nchild[w] = collapse_expand(t, cur[w], leaf_max, sl);
      for (int k = 0; k < 4; k++) slots[4 * w + k] = sl[k];
    }
    // exclusive scan of nchild[0, F) in chunks of CB (each thread scans the
    // entries it wrote above, so no other thread's writes are read)
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < F; b0 += CB) {
      const uint32_t i = b0 + tid;
      const uint32_t v = i < F ? nchild[i] : 0u;
      uint32_t x = v;
      for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
      }
      if (lane == 63) wsum[wave] = x;
      __syncthreads();
      if (wave == 0) {
        uint32_t ws = lane < CB / 64 ? wsum[lane] : 0u;
        for (int d = 1; d < CB / 64; d <<= 1) {
          uint32_t y = __shfl_up(ws, d, 64);
          if (lane >= (uint32_t)d) ws += y;
        }
        if (lane < CB / 64) wsum[lane] = ws;
      }
      __syncthreads();
      const uint32_t woff = wave ? wsum[wave - 1] : 0u;
      if (i < F) nchild[i] = carry + woff + x - v;
      carry += wsum[CB / 64 - 1];
      __syncthreads();
    }
    const uint32_t total = carry;
    for (uint32_t w = tid; w < F; w += CB) {
      uint32_t sl[4];
      for (int k = 0; k < 4; k++) sl[k] = slots[4 * w + k];
      collapse_emit(t, sl, nchild[w], leaf_max, next_base, nodes + (size_t)(level_base + w) * 8, nxt, cap);
    }
    __syncthreads();
    level_base = next_base;
    next_base += total;
    F = total;
    uint32_t* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (tid == 0) {
    result[0] = !ok ? 0u : (F == 0 ? 1u : 2u);
    result[1] = F;
    result[2] = level_base;
    result[3] = next_base;
    result[4] = (uint32_t)levels;
    result[5] = cur == fb ? 1u : 0u;
  }
}
// one wide level, multi-block
__global__ void k_collapse_expand(BinTree t, const uint32_t* frontier, uint32_t F, int leaf_max, uint32_t* slots,
                                  uint32_t* nchild) {
  for (uint32_t w = blockIdx.x * BB + threadIdx.x; w < F; w += gridDim.x * BB) {
    uint32_t sl[4];
    nchild[w] = collapse_expand(t, frontier[w], leaf_max, sl);
    for (int k = 0; k < 4; k++) slots[4 * w + k] = sl[k];
  }
}
__global__ void k_collapse_emit(BinTree t, uint32_t F, int leaf_max, const uint32_t* slots, const uint32_t* offs,
                                uint32_t level_base, uint32_t next_base, uint4* nodes, uint32_t* next_frontier,
                                uint32_t cap) {
  for (uint32_t w = blockIdx.x * BB + threadIdx.x; w < F; w += gridDim.x * BB) {
    uint32_t sl[4];
    for (int k = 0; k < 4; k++) sl[k] = slots[4 * w + k];
    collapse_emit(t, sl, offs[w], leaf_max, next_base, nodes + (size_t)(level_base + w) * 8, next_frontier, cap);
  }
}
// single-leaf root (n <= leaf_max)
__global__ void k_single_leaf(const float4* lo_box, const float4* hi_box, uint32_t n, uint4* nodes, float4* root) {
  if (threadIdx.x != 0) return;
  float4 lo = lo_box[0], hi = hi_box[0];
  uint4 e_lo = make_uint4(__float_as_uint(FINF), __float_as_uint(FINF), __float_as_uint(FINF), __float_as_uint(FINF));
  uint4 e_hi = make_uint4(__float_as_uint(-FINF), __float_as_uint(-FINF), __float_as_uint(-FINF),
                          __float_as_uint(-FINF));
  nodes[0] = make_uint4(__float_as_uint(lo.x), e_lo.y, e_lo.z, e_lo.w);
  nodes[1] = make_uint4(__float_as_uint(hi.x), e_hi.y, e_hi.z, e_hi.w);
  nodes[2] = make_uint4(__float_as_uint(lo.y), e_lo.y, e_lo.z, e_lo.w);
  nodes[3] = make_uint4(__float_as_uint(hi.y), e_hi.y, e_hi.z, e_hi.w);
  nodes[4] = make_uint4(__float_as_uint(lo.z), e_lo.y, e_lo.z, e_lo.w);
  nodes[5] = make_uint4(__float_as_uint(hi.z), e_hi.y, e_hi.z, e_hi.w);
  nodes[6] = make_uint4(0u, EMPTY_CHILD, EMPTY_CHILD, EMPTY_CHILD);
  nodes[7] = make_uint4(n, 0u, 0u, 0u);
  root[0] = lo;
  root[1] = hi;
}
// box of all primitives for the single-leaf case: merge of the n boxes
__global__ void k_merge_boxes(const uint32_t* idx, const float4* plo, const float4* phi, uint32_t n, float4* out) {
  if (threadIdx.x != 0) return;
  float4 lo = make_float4(FINF, FINF, FINF, 0.f), hi = make_float4(-FINF, -FINF, -FINF, 0.f);
  for (uint32_t i = 0; i < n; i++) {
    float4 a = plo[idx[i]], b = phi[idx[i]];
    lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.f);
    hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.f);
  }
  out[0] = lo;
  out[1] = hi;
}
__global__ void k_copy_root(const float4* nlo, const float4* nhi, float4* root) {
  if (threadIdx.x == 0) {
    root[0] = nlo[0];
    root[1] = nhi[0];
  }
}
__global__ void k_gather_u32(const uint32_t* src, uint32_t n, uint32_t* dst) {
  for (uint32_t i = blockIdx.x * BB + threadIdx.x; i < n; i += gridDim.x * BB) dst[i] = src[i];
}

// --------------------------------------------------------- build sequence
uint32_t build_bvh4(BuildScratch& sb, uint32_t n, int leaf_max, uint4* nodes_out, uint32_t* order_out,
                    uint32_t* morton_out, float4* root_lo_hi, hipStream_t st) {
  if (n == 0 || n > sb.capacity) return 0;
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_init_cbounds, dim3(1), dim3(64), 0, st, sb.cbounds);
  hipLaunchKernelGGL(k_centroid_bounds, dim3(g), dim3(BB), 0, st, sb.lo, sb.hi, n, sb.cbounds);
  hipLaunchKernelGGL(k_morton, dim3(g), dim3(BB), 0, st, sb.lo, sb.hi, n, sb.cbounds, sb.code_a, sb.idx_a);
  // stable LSD radix sort on 30 bits: 4 passes of 8 bits
  const uint32_t nblk = (n + TILE - 1) / TILE;
  uint32_t *ci = sb.code_a, *co = sb.code_b, *ii = sb.idx_a, *io = sb.idx_b;
  for (int pass = 0; pass < 4; pass++) {
    int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(BB), 0, st, ci, n, shift, nblk, sb.hist);
    exclusive_scan(sb.hist, (uint64_t)256 * nblk, sb.scan_tmp, st);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(BB), 0, st, ci, ii, co, io, n, shift, nblk, sb.hist);
    std::swap(ci, co);
    std::swap(ii, io);
  }
  // after an even number of passes the sorted data is back in *_a
  hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(BB), 0, st, ci, n, morton_out);
  hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(BB), 0, st, ii, n, order_out);
  BinTree t{sb.left, sb.right, sb.first, sb.last, ii, sb.lo, sb.hi, sb.nlo, sb.nhi};
  if (n == 1) {
    hipLaunchKernelGGL(k_merge_boxes, dim3(1), dim3(64), 0, st, ii, sb.lo, sb.hi, n, root_lo_hi);
    hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(64), 0, st, root_lo_hi, root_lo_hi + 1, n, nodes_out,
                       root_lo_hi);
    return 1;
  }
  launch_fill_u32(sb.flags, n - 1, 0u, st);
  launch_fill_u32(sb.parent, 2ull * n - 1, 0xFFFFFFFFu, st);
  hipLaunchKernelGGL(k_karras, dim3(grid_for(n - 1)), dim3(BB), 0, st, ci, n, sb.left, sb.right, sb.first, sb.last,
                     sb.parent);
  // refit: block-local climbs, then the nodes straddling the blocks' spans
  // (list of their inside children in frontier_b, its length in misc[0])
  launch_fill_u32(sb.misc, 1, 0u, st);
  hipLaunchKernelGGL(k_refit_local, dim3((n + REFIT_SPAN - 1) / REFIT_SPAN), dim3(BB), 0, st, n, ii, sb.lo, sb.hi,
                     sb.left, sb.right, sb.first, sb.last, sb.parent, sb.nlo, sb.nhi, sb.frontier_b, sb.misc);
  hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, st, n, ii, sb.lo, sb.hi, sb.left, sb.right, sb.parent,
                     sb.flags, sb.nlo, sb.nhi, sb.frontier_b, sb.misc);
  hipLaunchKernelGGL(k_copy_root, dim3(1), dim3(64), 0, st, sb.nlo, sb.nhi, root_lo_hi);
  if (n <= (uint32_t)leaf_max) {
    hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(64), 0, st, sb.nlo, sb.nhi, n, nodes_out, root_lo_hi);
    return 1;
  }
  // breadth-first collapse to BVH4 nodes: the narrow levels in one
  // workgroup (state in misc[1..6]), then the wide ones level by level
  hipLaunchKernelGGL(k_collapse_top, dim3(1), dim3(CB), 0, st, t, n, leaf_max, (uint32_t)sb.capacity, sb.frontier_a,
                     sb.frontier_b, sb.slots, sb.nchild, nodes_out, sb.misc + 1);
  uint32_t state[6] = {0, 0, 0, 0, 0, 0};
  if (hipMemcpyAsync(state, sb.misc + 1, sizeof(state), hipMemcpyDeviceToHost, st) != hipSuccess) return 0;
  if (hipStreamSynchronize(st) != hipSuccess) return 0;
  if (state[0] == 0) return 0;  // malformed hierarchy
  uint32_t F = state[1], level_base = state[2], next_base = state[3];
  int levels = (int)state[4];
  uint32_t* fa = state[5] ? sb.frontier_b : sb.frontier_a;
  uint32_t* fb = state[5] ? sb.frontier_a : sb.frontier_b;
  const uint32_t cap = (uint32_t)sb.capacity;
  while (F > 0) {
    if (++levels > 128 || next_base > n || F > cap) return 0;  // malformed hierarchy
    const int gf = grid_for(F);
    hipLaunchKernelGGL(k_collapse_expand, dim3(gf), dim3(BB), 0, st, t, fa, F, leaf_max, sb.slots, sb.nchild);
    // total = last exclusive prefix + last count
    uint32_t last_cnt = 0, last_off = 0;
    if (hipMemcpyAsync(&last_cnt, sb.nchild + (F - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess) return 0;
    exclusive_scan(sb.nchild, F, sb.scan_tmp, st);
    if (hipMemcpyAsync(&last_off, sb.nchild + (F - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess) return 0;
    hipLaunchKernelGGL(k_collapse_emit, dim3(gf), dim3(BB), 0, st, t, F, leaf_max, sb.slots, sb.nchild, level_base,
                       next_base, nodes_out, fb, cap);
    if (hipStreamSynchronize(st) != hipSuccess) return 0;
    const uint32_t total = last_off + last_cnt;
    level_base = next_base;
    next_base += total;
    F = total;
    std::swap(fa, fb);
  }
  return next_base;
}

// ------------------------------------------------ sorted record pools
// mode 0: records of `stride` float4 in BVH order; mode 1: original index
// (first + local) of each sorted position as uint32
__global__ void k_gather_records(const float4* src, const uint32_t* order, uint32_t n, uint32_t stride, float4* dst,
                                 uint32_t first, int mode) {
  for (uint32_t k = blockIdx.x * BB + threadIdx.x; k < n; k += gridDim.x * BB) {
    uint32_t o = order[k];
    if (mode == 1) {
      reinterpret_cast<uint32_t*>(dst)[k] = first + o;
    } else {
      for (uint32_t j = 0; j < stride; j++) dst[(size_t)k * stride + j] = src[(size_t)o * stride + j];
    }
  }
}
// triangle records (v0.xyz, mesh-local index), v1, v2 in BVH order
__global__ void k_gather_tris(const float4* vtx, const uint4* idx, const uint32_t* order, uint32_t n, float4* dst) {
  for (uint32_t k = blockIdx.x * BB + threadIdx.x; k < n; k += gridDim.x * BB) {
    uint32_t o = order[k];
    uint4 t = idx[o];
    float4 a = vtx[t.x], b = vtx[t.y], c = vtx[t.z];
    dst[3 * (size_t)k] = make_float4(a.x, a.y, a.z, __uint_as_float(o));
    dst[3 * (size_t)k + 1] = make_float4(b.x, b.y, b.z, 0.f);
    dst[3 * (size_t)k + 2] = make_float4(c.x, c.y, c.z, 0.f);
  }
}
__global__ void k_gather_instances(const uint4* src, const uint32_t* order, uint32_t n, uint4* dst) {
  for (uint32_t k = blockIdx.x * BB + threadIdx.x; k < n; k += gridDim.x * BB) {
    uint32_t o = order[k];
    for (int j = 0; j < 8; j++) dst[(size_t)k * 8 + j] = src[(size_t)o * 8 + j];
  }
}
void launch_gather_records(const float4* src, const uint32_t* order, uint32_t n, uint32_t stride, float4* dst,
                           uint32_t first, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_records, dim3(grid_for(n)), dim3(BB), 0, st, src, order, n, stride, dst, first, mode);
}
void launch_gather_tris(const float4* vtx, const uint4* idx, const uint32_t* order, uint32_t n, float4* dst,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_gather_tris, dim3(grid_for(n)), dim3(BB), 0, st, vtx, idx, order, n, dst);
}
void launch_gather_instances(const uint4* src, const uint32_t* order, uint32_t n, uint4* dst, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_instances, dim3(grid_for(n)), dim3(BB), 0, st, src, order, n, dst);
}

// ------------------------------------------------------------ launchers
void launch_bounds_spheres(const float4* sph, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_spheres, dim3(grid_for(n)), dim3(BB), 0, st, sph, n, lo, hi);
}
void launch_bounds_rects(const float4* rect, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_rects, dim3(grid_for(n)), dim3(BB), 0, st, rect, n, lo, hi);
}
void launch_bounds_boxes(const float4* box, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_boxes, dim3(grid_for(n)), dim3(BB), 0, st, box, n, lo, hi);
}
void launch_bounds_volumes(const float4* vol, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_volumes, dim3(grid_for(n)), dim3(BB), 0, st, vol, n, lo, hi);
}
void launch_bounds_cylinders(const float4* cyl, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_cylinders, dim3(grid_for(n)), dim3(BB), 0, st, cyl, n, lo, hi);
}
void launch_bounds_moving(const float4* mov, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_moving, dim3(grid_for(n)), dim3(BB), 0, st, mov, n, lo, hi);
}
void launch_bounds_tris(const float4* vtx, const uint4* idx, uint32_t n, float4* lo, float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_tris, dim3(grid_for(n)), dim3(BB), 0, st, vtx, idx, n, lo, hi);
}
void launch_bounds_instances(const uint4* insts_unsorted, const float4* blas_root, uint32_t n, float4* lo,
                             float4* hi, hipStream_t st) {
  hipLaunchKernelGGL(k_bounds_instances, dim3(grid_for(n)), dim3(BB), 0, st, insts_unsorted, blas_root, n, lo, hi);
}

}  // namespace pt
