// pt_kernels.hpp -- host-side launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "pt_layout.hpp"

namespace pt {

// radiance cache: scene box for the position encoding, query depth and the
// training-path threshold (24-bit hash < threshold)
struct CacheConsts {
  float lo[3], inv_ext[3];
  uint32_t query_depth;
  uint32_t train_threshold;
};

struct RenderConsts {
  uint32_t W, H, WH;
  uint32_t max_depth;
  int32_t russian;
  uint32_t rr_min_depth;
  uint32_t frame0;
  uint32_t flags;
  DevCamera cam;
  CacheConsts cache;
  const uint32_t* frame_dev;  // non-null: first frame of the batch read from device memory
  uint32_t store_time;        // keep each path's ray time (the scene has moving spheres)
};

// render path (pt_render.hip)
void launch_trace_closest(const DevScene& sc, const pt_ray* rays, uint64_t n, pt_hit* hits, uint32_t* counts,
                          uint2* spill, uint32_t* overflow, int grid, hipStream_t st);
void launch_trace_occluded(const DevScene& sc, const pt_ray* rays, uint64_t n, uint8_t* occ, uint2* spill,
                           uint32_t* overflow, int grid, hipStream_t st);
// frame_dev[0] = value (k_set_frame) / frame_dev[0] += step (k_next_frames)
void launch_set_frame(uint32_t* frame_dev, uint32_t value, hipStream_t st);
void launch_next_frames(uint32_t* frame_dev, uint32_t step, hipStream_t st);
// The camera paths of a batch of scenes (device tables): slots
// [off[k], off[k + 1]) are frame frame[k] of scene scene[k], k < nitems.
struct BatchStart {
  const RenderConsts* rcs;
  const uint32_t* off;
  const uint32_t* scene;
  const uint32_t* frame;
  uint32_t nitems;
};
// `scenes` (device table, non-null for a batch of scenes): each path's
// scene comes from its RNG state's top byte instead of `sc`.
// `start` (single scenes, depth 0, in == 0): the launch first starts the
// nstart camera paths of the dense state set 0 with these constants.
// `bstart` (a batch of scenes, depth 0, in == 0): the same for the nstart
// paths of the batch's items.
void launch_extend(const DevScene& sc, const DevPaths& P, int in, float tmin, uint2* spill, uint64_t* stats,
                   bool count_traversal, bool sort, int grid, hipStream_t st, const DevScene* scenes = nullptr,
                   const RenderConsts* start = nullptr, uint32_t nstart = 0, bool multi_ext = false,
                   const BatchStart* bstart = nullptr);
void launch_sort(const DevPaths& P, int in, int grid, hipStream_t st);
void launch_shade(const DevScene& sc, const DevPaths& P, int in, bool sorted, int out, uint32_t depth,
                  const RenderConsts& rc, int matset, bool count, uint2* spill, uint64_t* stats, int grid,
                  hipStream_t st, const DevScene* scenes = nullptr, const RenderConsts* rcs = nullptr,
                  bool batch_env = false, bool batch_ext = false);
int shade_material_set(uint32_t types_mask);
// environment importance-sampling tables of a W x H float4 map; scratch
// holds 1 + H 64-bit words
void launch_env_tables(const float4* tex, uint32_t W, uint32_t H, unsigned long long* cond,
                       unsigned long long* marg, unsigned long long* scratch, hipStream_t st);
void launch_env_guides(const unsigned long long* cond, const unsigned long long* marg, uint32_t W, uint32_t H,
                       uint32_t* guide, hipStream_t st);
// radiance cache in the render loop (pt_render.hip); `grid` is the batch's
// grid (its waves own the record segments)
void launch_cache_scan(const DevPaths& P, int grid, hipStream_t st);
// dense copies of `rows` records spread over the batch's records (which 0:
// training vertices with targets, 1: queries), row r taking rank
// (r * mul) mod rows (mul coprime to rows); tgt, vals, slots optional.
// half_step = k > 0 moves every rank by total / (2 k): the records halfway
// between those a gather of k rows takes.
void launch_cache_gather(const DevPaths& P, int which, uint32_t rows, uint32_t mul, int grid, uint4* feat, float* tgt,
                         float* vals, uint32_t* slots, hipStream_t st, uint32_t half_step = 0);
// judge sums of n gathered training rows (vals) against the network's
// outputs y: res[0] error of the answers, res[1] error of answering 0,
// res[2] non-finite outputs (FP64, one block)
void launch_cache_judge(const float* vals, const float* y, uint32_t n, double* res, hipStream_t st);
void launch_cache_apply(const DevPaths& P, int grid, hipStream_t st);
void launch_exhausted(const DevPaths& P, int in, uint64_t* stats, int grid, hipStream_t st);
// per-wave statistics rows [MAX_WAVES][STAT_WORDS] -> totals [STAT_WORDS]
void launch_stats_reduce(const uint64_t* part, uint64_t* total, hipStream_t st);
void launch_accumulate(const DevPaths& P, const RenderConsts& rc, uint32_t nframes, double* accum, uint64_t* stats,
                       int grid, hipStream_t st);
void launch_display(const double* accum, uint32_t npix, uint32_t frames, uchar4* rgba, int grid, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);
void launch_fill_f64(double* p, uint64_t n, double v, hipStream_t st);

// radiance-cache network (pt_mlp.hip)
void launch_mlp_pack(const float* params, uint16_t* fw, uint16_t* bw, float* bias, hipStream_t st);
// rows 0..n-1 dense, or (seg non-null) positions of a wave-segmented set
// with the given stride and per-wave counts
void launch_mlp_infer(const uint16_t* feats, uint32_t n, const uint16_t* fw, const float* bias, float* out, int grid,
                      hipStream_t st, const uint32_t* seg = nullptr, uint32_t stride = 0);
uint32_t mlp_padded(uint32_t n);
uint32_t mlp_fwd_bwd_waves(uint32_t n);
uint32_t mlp_chunks(uint32_t n);
void launch_mlp_train(const uint16_t* feats, const float* targets, uint32_t n, const uint16_t* fw, const uint16_t* bw,
                      const float* bias_in, uint16_t* X, uint16_t* dZ, float* db_partial, float* loss_partial,
                      float* dw_partial, float* params, float* m1, float* m2, float* grad_out, float lr,
                      uint32_t step, uint16_t* fw_out, uint16_t* bw_out, float* bias_out, hipStream_t st);
constexpr int MLP_FW_UINT4 = 2432;   // packed forward operands (uint4 = 8 BF16)
constexpr int MLP_BW_UINT4 = 1792;   // packed backward operands
constexpr int MLP_BIAS_WORDS = 4 * 64 + 16;
constexpr int MLP_X_ROWS = 352;      // unit-major activation rows
constexpr int MLP_DZ_ROWS = 272;     // unit-major output-gradient rows
constexpr int MLP_DB_WORDS = 4 * 64 + 3;

// acceleration-structure build (pt_bvh.hip)
struct BuildScratch {
  // sized for the largest structure of the scene
  float4* lo;         // primitive bounds
  float4* hi;
  uint32_t* code_a;   // Morton codes (ping-pong)
  uint32_t* code_b;
  uint32_t* idx_a;    // primitive order (ping-pong)
  uint32_t* idx_b;
  uint32_t* hist;     // radix-sort digit counts
  uint32_t* left;     // Karras binary hierarchy
  uint32_t* right;
  uint32_t* first;
  uint32_t* last;
  uint32_t* parent;   // parent of internal nodes [0, n-1) and leaves [n-1, 2n-1)
  uint32_t* flags;    // refit arrival counters
  float4* nlo;        // binary node bounds
  float4* nhi;
  uint32_t* frontier_a;  // collapse frontier (binary node per wide node)
  uint32_t* frontier_b;
  uint32_t* slots;    // 4 binary refs per frontier node
  uint32_t* nchild;   // internal children per frontier node (scanned)
  uint32_t* scan_tmp; // block sums of the scans
  uint32_t* cbounds;  // centroid bounds (6 ordered uints)
  uint32_t* misc;     // device scalars
  uint64_t capacity;
};

// Build one BVH4 over `n` primitive boxes already in scratch.lo/hi.
// Writes nodes to `nodes_out` (relative child indices), the sorted order to
// `order_out` and the sorted Morton codes to `morton_out`; returns the node
// count (host-synchronous), or 0 on error.
uint32_t build_bvh4(BuildScratch& sb, uint32_t n, int leaf_max, uint4* nodes_out, uint32_t* order_out,
                    uint32_t* morton_out, float4* root_lo_hi, hipStream_t st);
// Primitive bounds kernels (fill scratch.lo/hi)
void launch_bounds_spheres(const float4* sph, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_rects(const float4* rect, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_boxes(const float4* box, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_volumes(const float4* vol, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_cylinders(const float4* cyl, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_moving(const float4* mov, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_tris(const float4* vtx, const uint4* idx, uint32_t n, float4* lo, float4* hi, hipStream_t st);
void launch_bounds_instances(const uint4* insts_unsorted, const float4* blas_root, uint32_t n, float4* lo,
                             float4* hi, hipStream_t st);

}  // namespace pt
