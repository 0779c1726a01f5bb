// pt_env.hip -- importance-sampling tables of the environment map (no
// counterpart in the reference, which only looks the map up on a miss,
// programs/miss.cu:64-95).
//
// Texel weight = FP64 luminance (Rec. 709) x sin(theta of the row centre),
// scaled by 2^(40 - e) with e the binary exponent of the largest weight and
// truncated to a 64-bit integer.  The tables are inclusive prefix sums per
// row (cond, H x W) and over the row totals (marg, H).  Integer sums are
// exact, so the block-parallel scans below give the same tables as a
// sequential CPU build, bit for bit.
//
//   k_env_max   one pass over the texels: the largest weight (as ordered
//               double bits, one atomic per block)
//   k_env_rows  one block per row: per-thread chunk sums, an LDS scan of the
//               256 chunk sums, then the row's prefix sums and its total
//   k_env_marg  one block: prefix sums of the row totals
//
// Algorithmic bytes: 16 B read per texel per pass (two passes in k_env_rows
// plus k_env_max) and 8 B written per texel.
#include <algorithm>

#include "pt_device.hpp"
#include "pt_kernels.hpp"

namespace pt {

constexpr int EB = 256;

PTD double env_weight(float4 t, uint32_t j, uint32_t H) {
  double lum = 0.2126 * (double)t.x + 0.7152 * (double)t.y + 0.0722 * (double)t.z;
  float st = psin(PI_F * (((float)j + 0.5f) / (float)H));
  double w = lum * (double)st;
  return (w > 0.0) ? w : 0.0;  // NaN and negative texels get no weight
}

__global__ __launch_bounds__(EB) void k_env_max(const float4* tex, uint32_t W, uint32_t H,
                                                unsigned long long* max_bits) {
  __shared__ double red[EB];
  double m = 0.0;
  const uint64_t n = (uint64_t)W * H;
  for (uint64_t i = (uint64_t)blockIdx.x * EB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * EB)
    m = fmax(m, env_weight(tex[i], (uint32_t)(i / W), H));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = EB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  // non-negative doubles order like their bit patterns
  if (threadIdx.x == 0) atomicMax(max_bits, (unsigned long long)__double_as_longlong(red[0]));
}

PTD unsigned long long quantize_weight(double w, double mx, double scale) {
  return (mx > 0.0) ? (unsigned long long)(w * scale) : 0ull;
}

// Bits per quantised texel weight: 40, fewer for maps of more than 2^23
// texels, so that the total of W * H weights stays below 2^63.
PTD int env_weight_bits(uint32_t W, uint32_t H) {
  const unsigned long long n = (unsigned long long)W * H;
  int lg = 0;
  while (lg < 63 && (1ull << lg) < n) lg++;
  return min(40, 63 - lg);
}

// inclusive scan of one value per thread over the block (Hillis-Steele in LDS)
PTD unsigned long long block_inclusive_scan(unsigned long long v, unsigned long long* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < EB; off <<= 1) {
    unsigned long long a = threadIdx.x >= (unsigned)off ? lds[threadIdx.x - off] : 0ull;
    __syncthreads();
    lds[threadIdx.x] += a;
    __syncthreads();
  }
  return lds[threadIdx.x];
}

__global__ __launch_bounds__(EB) void k_env_rows(const float4* tex, uint32_t W, uint32_t H,
                                                 const unsigned long long* max_bits, unsigned long long* cond,
                                                 unsigned long long* row_sum) {
  __shared__ unsigned long long lds[EB];
  const uint32_t j = blockIdx.x;
  const double mx = __longlong_as_double((long long)*max_bits);
  int ex = 0;
  (void)frexp(mx, &ex);
  const double scale = ldexp(1.0, env_weight_bits(W, H) - ex);
  const uint32_t chunk = (W + EB - 1) / EB;
  const uint32_t b = threadIdx.x * chunk, e = min(W, b + chunk);
  const float4* row = tex + (size_t)j * W;
  unsigned long long local = 0;
  for (uint32_t i = b; i < e; i++) local += quantize_weight(env_weight(row[i], j, H), mx, scale);
  unsigned long long incl = block_inclusive_scan(local, lds);
  unsigned long long run = incl - local;
  unsigned long long* out = cond + (size_t)j * W;
  for (uint32_t i = b; i < e; i++) {
    run += quantize_weight(env_weight(row[i], j, H), mx, scale);
    out[i] = run;
  }
  if (threadIdx.x == EB - 1) row_sum[j] = incl;
}

__global__ __launch_bounds__(EB) void k_env_marg(const unsigned long long* row_sum, uint32_t H,
                                                 unsigned long long* marg) {
  __shared__ unsigned long long lds[EB];
  const uint32_t chunk = (H + EB - 1) / EB;
  const uint32_t b = threadIdx.x * chunk, e = min(H, b + chunk);
  unsigned long long local = 0;
  for (uint32_t i = b; i < e; i++) local += row_sum[i];
  unsigned long long incl = block_inclusive_scan(local, lds);
  unsigned long long run = incl - local;
  for (uint32_t i = b; i < e; i++) {
    run += row_sum[i];
    marg[i] = run;
  }
}

// Guide tables (Chen & Asau 1974) of the sampler's two searches: for an
// array a of n inclusive prefix sums with total a[n-1], entry b < G is the
// first index whose prefix exceeds floor(b * total / G), entry G is n - 1.
// A target x in [floor(b total / G), floor((b+1) total / G)) has its first
// prefix above it between entries b and b + 1, so the sampler searches
// only there and finds the same index as a search of the whole array.
// floor(b * total / 64) = b (total >> 6) + (b (total & 63)) >> 6, exact in
// 64-bit integers.
PTD unsigned long long guide_bound(unsigned long long total, uint32_t b) {
  return (total >> 6) * b + (((total & 63ull) * b) >> 6);
}
PTD uint32_t first_above(const unsigned long long* a, uint32_t lo, uint32_t hi, unsigned long long x) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}
static_assert(ENV_GUIDE == 64, "guide_bound divides by 64");
__global__ void k_env_guides(const unsigned long long* cond, const unsigned long long* marg, uint32_t W, uint32_t H,
                             uint32_t* guide) {
  const uint64_t total_entries = (uint64_t)(H + 1) * (ENV_GUIDE + 1);
  for (uint64_t e = (uint64_t)blockIdx.x * EB + threadIdx.x; e < total_entries; e += (uint64_t)gridDim.x * EB) {
    const uint32_t r = (uint32_t)(e / (ENV_GUIDE + 1)), b = (uint32_t)(e % (ENV_GUIDE + 1));
    const unsigned long long* a = r < H ? cond + (size_t)r * W : marg;
    const uint32_t n = r < H ? W : H;
    const unsigned long long total = r < H ? a[n - 1] : marg[H - 1];
    guide[e] = b == ENV_GUIDE ? n - 1 : first_above(a, 0, n - 1, guide_bound(total, b));
  }
}
void launch_env_guides(const unsigned long long* cond, const unsigned long long* marg, uint32_t W, uint32_t H,
                       uint32_t* guide, hipStream_t st) {
  const uint64_t n = (uint64_t)(H + 1) * (ENV_GUIDE + 1);
  const int grid = (int)std::min<uint64_t>(2048, (n + EB - 1) / EB);
  hipLaunchKernelGGL(k_env_guides, dim3(grid), dim3(EB), 0, st, cond, marg, W, H, guide);
}

void launch_env_tables(const float4* tex, uint32_t W, uint32_t H, unsigned long long* cond,
                       unsigned long long* marg, unsigned long long* scratch, hipStream_t st) {
  // scratch: [0] max weight bits, [1 .. H] row totals
  launch_fill_u32((uint32_t*)scratch, 2, 0u, st);
  uint64_t n = (uint64_t)W * H;
  int grid = (int)std::min<uint64_t>(2048, (n + EB - 1) / EB);
  hipLaunchKernelGGL(k_env_max, dim3(grid), dim3(EB), 0, st, tex, W, H, scratch);
  hipLaunchKernelGGL(k_env_rows, dim3(H), dim3(EB), 0, st, tex, W, H, (const unsigned long long*)scratch, cond,
                     scratch + 1);
  hipLaunchKernelGGL(k_env_marg, dim3(1), dim3(EB), 0, st, (const unsigned long long*)(scratch + 1), H, marg);
}

}  // namespace pt
