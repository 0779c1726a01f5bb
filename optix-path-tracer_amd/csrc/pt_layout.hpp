// pt_layout.hpp -- HBM layout of the scene, acceleration structure and path
// queues shared by the host orchestration and the gfx950 kernels.
#pragma once

#include <stdint.h>

#include "../../include/pt_api.h"

namespace pt {

// ---------------------------------------------------------------- BVH4
// One node = 128 bytes = 32 words, one cache line:
//   words  0.. 3  lo.x[4]     words  4.. 7  hi.x[4]
//   words  8..11  lo.y[4]     words 12..15  hi.y[4]
//   words 16..19  lo.z[4]     words 20..23  hi.z[4]
//   words 24..27  child[4]: internal -> node index within its structure,
//                 leaf -> first sorted primitive, empty -> 0xFFFFFFFF
//   words 28..31  count[4]: 0 internal / empty, else leaf primitive count
constexpr int NODE_WORDS = 32;
constexpr int BLAS_LEAF_MAX = 4;
constexpr int TLAS_LEAF_MAX = 1;
// Traversal stack references: internal node index, or
// LEAF_BIT | (count - 1) << 28 | first.
constexpr uint32_t LEAF_BIT = 1u << 31;
constexpr uint32_t EMPTY_CHILD = 0xFFFFFFFFu;

// ------------------------------------------------------------ records
// Sorted (BVH-order) primitive records used by traversal:
//   sphere   float4 (c.xyz, r)                           16 B  + uint32 index
//   rect     float4 (a0, a1, b0, b1), float4 (k, axis, flip, index)  32 B
//   box      float4 (lo.xyz, index),  float4 (hi.xyz, 0)            32 B
//   triangle float4 (v0.xyz, index), float4 (v1.xyz, 0), float4 (v2, 0) 48 B
// Instance record, 8 x 16 B (sorted in top-level order):
//   [0..2] world-to-object rows, [3] (root node, prim base, kind | xf << 8,
//   instance id), [4..6] object-to-world rows, [7] (first, count, mesh,
//   bottom-level structure)
constexpr int INST_WORDS = 32;

// Block size of the path kernels and the traversal stack: STK_LDS entries per
// lane in LDS, STK_SPILL more per thread in HBM.
constexpr int BLOCK = 256;
constexpr int STK_LDS = 16;
// every thread owns STK_TOTAL entries in all: the ones beyond its kernel's
// LDS entries spill to its HBM slice of STK_TOTAL entries
constexpr int STK_TOTAL = 64;
constexpr int STK_SPILL = STK_TOTAL - STK_LDS;
// buckets per guide table of the environment sampler's searches
constexpr uint32_t ENV_GUIDE = 64;
// closest-hit kernel (k_extend) in scenes with meshes and in the camera-ray
// pass: LDS entries per lane and occupancy target (waves per SIMD); 16
// entries x 8 B x 256 lanes is 32 KB per block, which with up to 96 VGPRs
// holds 5 waves per SIMD (the other instantiations: extend_waves)
#ifndef PT_EXTEND_STK
#define PT_EXTEND_STK 16
#endif
#ifndef PT_EXTEND_WAVES
#define PT_EXTEND_WAVES 5
#endif
// bounce rays (k_extend after depth 0): a lane whose ray is done takes the
// wave's next path once at least this many of the wave's lanes are idle
// (0: the wave's lanes take 64 paths at a time, each group waiting for its
// longest ray)
#ifndef PT_EXTEND_REFILL
#define PT_EXTEND_REFILL 16
#endif
constexpr int STK_LDS_EXTEND = PT_EXTEND_STK;
// Grid of the path kernels (blocks of BLOCK / 64 waves).  Pools of at
// least GRID_CAP_PATHS paths take up to GRID_CAP blocks: 5120 blocks are
// whole rounds of both the closest-hit kernel (5 waves per SIMD: 1280
// resident blocks on 256 CUs) and the shading kernel (4 waves: 1024), so
// neither ends a launch with part of the chip idle.  Smaller pools take
// up to GRID_SMALL blocks: their deep bounces are short launches, which
// run faster in fewer, fuller wave segments.  MAX_GRID sizes the per-wave
// tables.
constexpr int MAX_GRID = 8192;
constexpr int GRID_CAP = 5120;
constexpr int GRID_SMALL = 2048;
constexpr uint64_t GRID_CAP_PATHS = 1ull << 27;
constexpr int WAVES = BLOCK / 64;
constexpr int MAX_WAVES = MAX_GRID * WAVES;

struct DevMaterial {
  int32_t type;
  int32_t tex0, tex1;
  float p0, p1, p2, p3;
  int32_t pad;
};

struct DevTexture {
  int32_t type;
  float r, g, b;
  uint32_t w, h;
  uint32_t offset;  // first float4 texel in the texel pool
  uint32_t pad;
};

struct DevLight {
  int32_t type, axis;
  float a0, a1, b0, b1, k;
  float cx, cy, cz, r;
  float ex, ey, ez;
  float pad[2];
};

struct DevMesh {
  uint32_t vbase;   // first vertex in the vertex pool
  uint32_t ibase;   // first index triple (triangle) in the index pool
  uint32_t ntri;
  int32_t has_n, has_uv;
  int32_t mat;      // mesh material when fmat_base < 0
  int32_t fmat_base;  // first per-face material, or -1
  int32_t fidx_base;  // first per-face vector-texture index, or -1
};

// Everything a kernel needs to see of the scene, passed by value.
struct DevScene {
  const uint4* nodes;  // all BVH4 nodes (every BLAS, then the TLAS)
  uint32_t tlas_root;  // absolute node index of the TLAS root
  uint32_t num_instances;
  const uint4* insts;  // instance records in top-level order
  // sorted primitive records (per-kind pools, BLAS after BLAS)
  const float4* sph;
  const uint32_t* sph_idx;
  const float4* rect;
  const float4* box;
  const float4* tri;
  // original-order data for hit records and materials
  const float4* sph_o;   // (c, r)
  const float4* rect_o;  // 2 per rect, as the sorted record
  const float4* box_o;   // 2 per box
  const int32_t* sph_mat;
  const int32_t* rect_mat;
  const int32_t* box_mat;
  // constant-density media: 2 float4 per volume, (a, density) and
  // (b, original index | box << 31); sorted pool and original order
  const float4* vol;
  const float4* vol_o;
  const int32_t* vol_mat;
  // open cylinders: (origin, length), (radius, 0, 0, original index)
  const float4* cyl;
  const float4* cyl_o;
  const int32_t* cyl_mat;
  // moving spheres: (center0, radius), (center1, original index), (time0, time1, 0, 0)
  const float4* mov;
  const float4* mov_o;
  const int32_t* mov_mat;
  uint32_t ext_instances;  // instances of the kinds volumes, cylinders, moving spheres (0: none)
  uint32_t mov_instances;  // instances of moving spheres (0: paths keep no time)
  uint32_t kinds;          // bit k: instances of PT_GEOM kind k < 4 occur
  const DevMesh* meshes;
  const float4* vtx;     // vertex pool (xyz)
  const float4* nrm;     // normal pool (xyz)
  const float2* uv;      // uv pool
  const uint4* idx;      // index pool (x, y, z: mesh-local vertex ids)
  const int32_t* fmat;   // per-face materials
  const int32_t* fidx;   // per-face vector-texture indices
  const DevMaterial* mats;
  const DevTexture* texs;
  const float4* texels;
  const DevLight* lights;        // every light, in scene order
  int32_t num_lights;
  const DevLight* lights_plain;  // the same without environment lights
  int32_t num_lights_plain;
  pt_miss_desc miss;
  // environment importance sampling (env_total == 0: none): inclusive
  // prefix sums of 64-bit fixed-point texel weights per row (env_cond,
  // env_h x env_w) and over row totals (env_marg)
  const unsigned long long* env_cond;
  const unsigned long long* env_marg;
  unsigned long long env_total;
  uint32_t env_w, env_h;
  // guide tables of the two searches: (env_h + 1) rows of ENV_GUIDE + 1
  // indices, row j < env_h for env_cond's row j, row env_h for env_marg
  const uint32_t* env_guide;
};

struct DevCamera {
  float llc[3], hor[3], ver[3], org[3], u[3], v[3];
  float lens_radius, time0, time1;
};

// Path state in HBM: 64 bytes per live path, structure of arrays, compacted
// every bounce into the next ping-pong set so that every kernel streams it
// contiguously.
//
// Wave segments: a state set is split among the waves of the batch's grid
// (G blocks).  Wave w = (block b, wave v of the block) owns the positions
//   pos(j) = (j / 64) * G * BLOCK + b * BLOCK + v * 64 + j % 64,  j = 0 .. n_w - 1
// and n_w is seg[set * MAX_WAVES + w].  Every kernel of a batch runs on the
// same grid; a wave reads its own segment of the input set and compacts its
// survivors (never more than its inputs) into its own segment of the output
// set, so the hot loop takes no global atomic.  A dense set of n positions
// is the case n_w = (positions of the segment below n).
constexpr uint32_t SLOT_MASK = 0x3FFFFFFFu;
constexpr uint32_t TRAIN_BIT = 1u << 30;  // radiance-cache training path
constexpr uint32_t SPEC_BIT = 1u << 31;   // last bounce was specular
struct PathState {
  float4* o;  // (origin.xyz, RNG state bits)
  float4* d;  // (direction.xyz, slot | TRAIN_BIT | SPEC_BIT; slot =
              //  frame-in-batch * W*H + launch index)
  float4* T;  // (throughput.xyz, pdf of the last diffuse continuation sample)
  float4* L;  // (radiance.xyz, decision-signature bits)
};

struct DevPaths {
  PathState st[3];      // [0], [1] ping-pong, [2] material-sorted copy
  float4* hit;          // per active path of the input set: (t, instance position, prim, instance id)
  float4* hit_sorted;
  uint32_t* keys;       // material bucket per active path (sorted shading)
  float4* result;       // per slot: (color() value, bounces bits)
  float* ptime;         // per slot: the path's ray time (scenes with moving spheres)
  uint32_t* sig_final;  // per slot: decision signature
  uint32_t* counters;   // see CTR_*
  uint32_t* seg;        // [3][MAX_WAVES] live paths per wave segment of each state set
  uint32_t* bcount;     // [8][MAX_WAVES] paths per material bucket and wave (sorted shading)
  uint32_t* bbase;      // [8][MAX_WAVES] first sorted position of each (bucket, wave)
  // radiance-cache records (PT_FLAG_RADIANCE_CACHE), kept in wave segments
  // like the path state: queries of the query-depth bounce in each wave's
  // segment of the q_* arrays (counts qseg), training vertices of the whole
  // batch in each wave's segment of the t_* arrays (running counts tseg,
  // capped by t_cap); tpre / qpre are the exclusive prefix sums over waves
  uint4* q_feat;        // per query: 96 BF16 features (12 uint4)
  float4* q_meta;       // per query: (T.xyz, slot bits)
  float* q_out;         // per query: network output (3 floats)
  uint32_t q_cap;
  uint4* t_feat;        // per training vertex: 96 BF16 features
  float4* t_L;          // per training vertex: (L before the vertex, slot bits)
  float4* t_T;          // per training vertex: (T at the vertex, 0)
  uint32_t t_cap;
  uint32_t* tseg;       // [MAX_WAVES]
  uint32_t* qseg;       // [MAX_WAVES]
  uint32_t* tpre;       // [MAX_WAVES]
  uint32_t* qpre;       // [MAX_WAVES]
};

// device counters (uint32)
constexpr int CTR_SORTED = 0;          // paths in the material-sorted set
constexpr int CTR_STACK_OVERFLOW = 20;
constexpr int CTR_QUERY = 21;          // radiance-cache queries of the batch
constexpr int CTR_TRAIN = 22;          // radiance-cache training vertices of the batch
constexpr int CTR_WORDS = 32;
// 64-bit statistics: [0] extension rays, [1] shadow rays, [2] nodes,
// [3] prims, [4] instances, ... kept per wave ([MAX_WAVES][STAT_WORDS],
// each row updated by its own wave only) and summed once per render.
constexpr int STAT_EXT = 0, STAT_SHADOW = 1, STAT_NODES = 2, STAT_PRIMS = 3, STAT_INSTS = 4;
constexpr int STAT_SH_NODES = 5, STAT_SH_PRIMS = 6, STAT_SH_INSTS = 7, STAT_EXHAUSTED = 8, STAT_NAN = 9;
constexpr int STAT_WORDS = 16;

// ------------------------------------------------------- radiance cache
// 96 -> 64 -> 64 -> 64 -> 64 -> 3 network (ReLU hidden layers); parameters
// in the order W1 (64 x 96), b1, W2, b2, W3, b3, W4, b4, W5 (3 x 64), b5.
constexpr int MLP_IN = 96;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 3;
constexpr int MLP_PARAMS = 18883;
constexpr int MLP_CHUNK = 1024;  // samples per split-K chunk of the weight-gradient GEMM

}  // namespace pt
