// pt_context.cpp -- implementation of the C ABI declared in include/pt_api.h.
//
// Host orchestration only: validation, HBM allocation, uploads, the LBVH
// build sequence and the wavefront bounce loop.  All per-ray and
// per-primitive work runs in the gfx950 kernels of pt_render.hip and
// pt_bvh.hip; there is no CPU fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pt_api.h"
#include "pt_kernels.hpp"
#include "pt_layout.hpp"
#include "../host/png_write.hpp"
#include "../host/rgbe.hpp"

using namespace pt;

namespace {

// Paths in flight per batch when the caller leaves it to the library: as
// many as half of the device's memory holds (PATH_BYTES of path state
// each), at most 2^29 (135 GB on the 288 GB MI355X).  Every batch ends in a
// dozen bounces whose launches last as long as their slowest waves (a few
// paths trapped in glass or between meshes), whatever the number of paths,
// so the fewer batches a render takes, the less of that tail it pays: the
// 1080p 256-spp mesh scene took 487 ms in batches of 8 frames (16.6 M
// paths) and 233 ms in batches of 128 frames (265 M paths, 67 GB); all 256
// frames fit one batch of 2^29.
constexpr uint64_t PATHS_MAX = 1ull << 29;
constexpr uint64_t PATH_BYTES = 3 * 64 + 16 + 16 + 4 + 16 + 4 + 4;  // per path, ensure_paths

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct BlasInfo {
  int kind;
  uint32_t first, count;
  uint32_t node_base, node_count;
  uint32_t prim_base;
};

}  // namespace

struct pt_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // ---- scene (host copies)
  bool has_scene = false;
  std::vector<pt_sphere_desc> spheres;
  std::vector<pt_rect_desc> rects;
  std::vector<pt_box_desc> boxes;
  std::vector<pt_volume_desc> volumes;
  std::vector<pt_cylinder_desc> cylinders;
  std::vector<pt_moving_sphere_desc> moving;
  struct MeshHost {
    std::vector<float> v, n, uv;
    std::vector<uint32_t> idx;
    std::vector<int32_t> fmat;
    std::vector<int32_t> fidx;
    int32_t mat;
  };
  std::vector<MeshHost> meshes;
  std::vector<pt_instance_desc> instances;
  std::vector<pt_material_desc> materials;
  std::vector<pt_texture_desc> textures;
  std::vector<std::vector<float>> texels;
  std::vector<pt_light_desc> lights;
  pt_miss_desc miss{};

  // ---- device scene buffers
  std::vector<DevBuf> scene_bufs;  // original-order data, freed on set_scene
  std::vector<DevBuf> accel_bufs;  // acceleration data, freed on rebuild
  DevScene dsc{};
  bool accel_ready = false;
  std::vector<BlasInfo> blas;
  uint64_t total_nodes = 0, total_prims_sorted = 0;
  uint32_t tlas_nodes = 0;
  uint4* nodes_dev = nullptr;
  uint32_t* order_dev = nullptr;   // every BLAS order, then the TLAS order
  uint32_t* morton_dev = nullptr;
  uint64_t accel_bytes = 0;

  // ---- camera and render state
  bool has_camera = false;
  pt_camera cam{};
  uint32_t W = 0, H = 0;
  uint32_t frames = 0;
  double* accum = nullptr;
  uchar4* rgba = nullptr;
  uint32_t last_fpb = 0;        // frames of the last batch
  uint32_t last_frame_slot = 0; // slot base of the last frame in its batch
  bool paths_valid = false;
  int shading_kinds = 0;  // distinct non-emitter material types of the scene
  uint32_t mat_types = 0;  // bit t set when material type t occurs

  // ---- radiance cache network
  bool cache_ready = false;
  std::vector<DevBuf> cache_bufs;   // parameters, Adam moments, packed operands
  float* cp = nullptr;              // FP32 master parameters
  float* cm1 = nullptr;             // Adam first moments
  float* cm2 = nullptr;             // Adam second moments
  float* cgrad = nullptr;           // gradient of the last training step
  uint16_t* cfw = nullptr;          // packed forward operands (BF16)
  uint16_t* cbw = nullptr;          // packed backward operands (BF16)
  float* cbias = nullptr;           // biases (FP32, padded)
  uint32_t cache_step = 0;
  // render-loop use (PT_FLAG_RADIANCE_CACHE)
  uint32_t cache_query_depth = 2;
  float cache_train_fraction = 0.03f;
  uint32_t cache_batch = 16384;     // samples per Adam step
  uint32_t cache_steps = 4;         // Adam steps per render batch
  float cache_lr = 1e-3f;
  uint32_t cache_seed = 0;          // of the last pt_cache_init (a diverged network restarts from it)
  // per-batch judge: training vertices the plain tracer finished, checked
  // against the network before its answers are used
  float cache_judge_max = 1.f;      // fall back when error ratio >= this
  uint32_t cache_judge_rows = 16384;
  uint32_t cjudge_cap = 0;
  std::vector<DevBuf> cjudge_bufs;
  uint4* cjfeat = nullptr;          // [rows][12] features
  float* cjvals = nullptr;          // [rows][9] L, T, target
  float* cjout = nullptr;           // [rows][3] network outputs
  double* cjres = nullptr;          // judge sums
  double cache_last_judge = -1.0;   // error ratio of the last judged batch (-1: none)
  // the next batch renders plain (no queries; training paths still train
  // and the network is still judged): a new network, or the last judge
  // failed it
  bool cache_hold = true;
  uint32_t cache_last_fallbacks = 0;
  uint32_t cache_last_plain_batches = 0;
  float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {1, 1, 1};  // top-level box, for the position encoding
  uint64_t cache_rec_cap = 0;       // slots the record buffers were sized for
  std::vector<DevBuf> crec_bufs;
  // batch of scenes (pt_render_batch run by this context): scene and
  // constant tables, work-item tables
  std::vector<DevBuf> batch_bufs;
  size_t batch_scene_cap = 0, batch_item_cap = 0;
  DevScene* d_scenes = nullptr;
  RenderConsts* d_rcs = nullptr;
  uint32_t *d_item_off = nullptr, *d_item_scene = nullptr, *d_item_frame = nullptr;
  uint32_t cache_last_train = 0, cache_last_query = 0, cache_last_steps = 0;
  uint32_t cache_last_batch = 0;    // samples of the last Adam step
  int cache_last_grid = 0;          // grid of the last cache batch (its waves own the records)
  uint32_t cwork_cap = 0;           // samples the work buffers hold
  std::vector<DevBuf> cwork_bufs;
  uint16_t* cfeat = nullptr;        // features, BF16 [n][96]
  float* ctgt = nullptr;            // targets / outputs [n][3]
  uint16_t* cX = nullptr;           // unit-major activations
  uint16_t* cdZ = nullptr;          // unit-major output gradients
  float* cdb = nullptr;             // per-wave bias gradients, then their sums
  float* closs = nullptr;           // per-wave loss sums
  float* cdw = nullptr;             // per-chunk weight gradients

  // ---- batch graph (PT_FLAG_GRAPH) and autotuned launch shape
  uint32_t* frame_dev = nullptr;    // first frame of the next replayed batch
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint8_t> graph_key;   // launch state the graph was captured for
  // pt_render_batch run by this context: its captured graph and tuned shape
  hipGraph_t batch_graph = nullptr;
  hipGraphExec_t batch_graph_exec = nullptr;
  std::vector<uint8_t> batch_graph_key;
  struct BatchTuned {
    bool valid = false;
    uint64_t paths = 0;
    bool sort = true, graph = false;
    std::vector<uint8_t> key;       // the batch composition it was tuned for
  } batch_tuned;
  struct Tuned {
    bool valid = false;
    uint32_t W = 0, H = 0, max_depth = 0, flags = 0;
    uint32_t frames_per_batch = 0;
    bool sort = false, graph = false;
    double ms_per_frame = 0.0;
  } tuned;

  // ---- path queues
  uint64_t paths_target = 1u << 24;  // automatic paths in flight per batch (set at pt_create)
  uint64_t path_capacity = 0;
  std::vector<DevBuf> path_bufs;
  DevPaths paths{};
  uint2* spill = nullptr;
  uint64_t* stats = nullptr;        // per-wave rows [MAX_WAVES][STAT_WORDS]
  uint64_t* stats_total = nullptr;  // [STAT_WORDS]
  uint32_t* qcounters = nullptr;  // counters of the query kernels
};

namespace {

pt_status fail(pt_context* c, pt_status s, const std::string& m) {
  if (c) c->err = m;
  return s;
}
pt_status hip_fail(pt_context* c, hipError_t e, const char* what) {
  return fail(c, PT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(ctx, call)                                  \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

hipError_t dalloc(std::vector<DevBuf>& list, void** out, size_t bytes) {
  if (bytes == 0) bytes = 16;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return e;
  list.push_back({p, bytes});
  *out = p;
  return hipSuccess;
}
void dfree(std::vector<DevBuf>& list) {
  for (auto& b : list) (void)hipFree(b.p);
  list.clear();
}
template <class T>
hipError_t upload(pt_context* c, std::vector<DevBuf>& list, T** out, const void* src, size_t bytes) {
  void* p = nullptr;
  hipError_t e = dalloc(list, &p, bytes);
  if (e != hipSuccess) return e;
  if (bytes) {
    e = hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return e;
  }
  *out = static_cast<T*>(p);
  return hipSuccess;
}

inline uint32_t fb(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bf(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

int grid_for(uint64_t n) {
  const uint64_t cap = n >= GRID_CAP_PATHS ? (uint64_t)GRID_CAP : (uint64_t)GRID_SMALL;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + BLOCK - 1) / BLOCK, cap));
}

// an odd stride near 0.618 n that is coprime to n (a full-period
// multiplicative order of 0..n-1)
uint32_t golden_stride(uint32_t n) {
  if (n < 3) return 1u;
  uint32_t a = (uint32_t)((double)n * 0.6180339887498949) | 1u;
  while (std::gcd(a, n) != 1u) a += 2u;
  return a % n ? a % n : 1u;
}
}  // namespace

// gather kernels of pt_bvh.hip: BVH-ordered primitive and instance records
namespace pt {
void launch_gather_records(const float4* src, const uint32_t* order, uint32_t n, uint32_t stride, float4* dst,
                           uint32_t first, int mode, hipStream_t st);
void launch_gather_tris(const float4* vtx, const uint4* idx, const uint32_t* order, uint32_t n, float4* dst,
                        hipStream_t st);
void launch_gather_instances(const uint4* src, const uint32_t* order, uint32_t n, uint4* dst, hipStream_t st);
}  // namespace pt

// Scratch of pt_trace_closest_device_noalloc: the traversal-stack spill
// of NOALLOC_GRID blocks (NOALLOC_SPILL_BYTES) followed by 256 bytes of
// counters.  Every GPU the library runs on holds one copy of it until the
// process ends; host code addresses it only through the pointer
// hipGetSymbolAddress returned for that device (g_scratch), never through
// the array's name.
namespace {
constexpr uint32_t NOALLOC_GRID = 384;
constexpr size_t NOALLOC_SPILL_BYTES = (size_t)NOALLOC_GRID * BLOCK * STK_TOTAL * sizeof(uint2);
constexpr size_t SCRATCH_BYTES = NOALLOC_SPILL_BYTES + 256;
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES] = {};
}  // namespace
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// once per device, with that device current
static hipError_t resolve_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch[device] = ptr;
  return hipSuccess;
}

extern "C" {

int32_t pt_api_version(void) { return PT_API_VERSION; }

static std::string g_create_err;

pt_status pt_create(int32_t device, pt_context** out) {
  if (!out) return PT_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    g_create_err = std::string("no HIP device available: ") + hipGetErrorString(e);
    return PT_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    g_create_err = "device index out of range";
    return PT_ERR_INVALID_ARG;
  }
  if (device >= MAX_DEVICES) {
    g_create_err = "device index beyond the library's device table";
    return PT_ERR_INVALID_ARG;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return PT_ERR_DEVICE;
  }
  if ((e = resolve_scratch(device)) != hipSuccess) {
    g_create_err = std::string("hipGetSymbolAddress (scratch): ") + hipGetErrorString(e);
    return PT_ERR_DEVICE;
  }
  pt_context* c = new pt_context();
  c->device = device;
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
      c->paths_target = std::max<uint64_t>(1u << 20, std::min<uint64_t>(PATHS_MAX, total_b / 2 / PATH_BYTES));
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    delete c;
    return PT_ERR_DEVICE;
  }
  if ((e = hipMalloc(&c->stats, (size_t)MAX_WAVES * STAT_WORDS * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc(&c->stats_total, STAT_WORDS * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc(&c->qcounters, CTR_WORDS * sizeof(uint32_t))) != hipSuccess) {
    g_create_err = std::string("hipMalloc: ") + hipGetErrorString(e);
    pt_destroy(c);
    return PT_ERR_OUT_OF_MEMORY;
  }
  launch_fill_u32(c->qcounters, CTR_WORDS, 0u, c->stream);
  *out = c;
  return PT_OK;
}

void pt_destroy(pt_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->scene_bufs);
  dfree(c->accel_bufs);
  dfree(c->path_bufs);
  dfree(c->cache_bufs);
  dfree(c->cjudge_bufs);
  dfree(c->cwork_bufs);
  dfree(c->crec_bufs);
  dfree(c->batch_bufs);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->batch_graph_exec) (void)hipGraphExecDestroy(c->batch_graph_exec);
  if (c->batch_graph) (void)hipGraphDestroy(c->batch_graph);
  if (c->frame_dev) (void)hipFree(c->frame_dev);
  if (c->accum) (void)hipFree(c->accum);
  if (c->rgba) (void)hipFree(c->rgba);
  if (c->spill) (void)hipFree(c->spill);
  if (c->stats) (void)hipFree(c->stats);
  if (c->stats_total) (void)hipFree(c->stats_total);
  if (c->qcounters) (void)hipFree(c->qcounters);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* pt_last_error(const pt_context* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

pt_status pt_set_scene(pt_context* c, const pt_scene_desc* d) {
  if (!c || !d) return PT_ERR_INVALID_ARG;
  (void)hipSetDevice(c->device);
  // ---- validation
  auto bad = [&](const std::string& m) { return fail(c, PT_ERR_INVALID_ARG, m); };
  if ((d->num_spheres && !d->spheres) || (d->num_rects && !d->rects) || (d->num_boxes && !d->boxes) ||
      (d->num_meshes && !d->meshes) || (d->num_instances && !d->instances) || (d->num_materials && !d->materials) ||
      (d->num_textures && !d->textures) || (d->num_lights && !d->lights) || (d->num_volumes && !d->volumes) ||
      (d->num_cylinders && !d->cylinders) || (d->num_moving_spheres && !d->moving_spheres))
    return bad("scene array pointer is NULL with a non-zero count");
  if (d->num_instances == 0) return bad("scene has no instances");
  auto mat_ok = [&](int32_t m) { return m >= 0 && (uint32_t)m < d->num_materials; };
  auto tex_ok = [&](int32_t t) { return t >= 0 && (uint32_t)t < d->num_textures; };
  for (uint32_t i = 0; i < d->num_spheres; i++)
    if (!mat_ok(d->spheres[i].material)) return bad("sphere material index out of range");
  for (uint32_t i = 0; i < d->num_rects; i++) {
    if (!mat_ok(d->rects[i].material)) return bad("rect material index out of range");
    if (d->rects[i].axis < 0 || d->rects[i].axis > 2) return bad("invalid rect axis");
  }
  for (uint32_t i = 0; i < d->num_boxes; i++)
    if (!mat_ok(d->boxes[i].material)) return bad("box material index out of range");
  for (uint32_t i = 0; i < d->num_volumes; i++) {
    const pt_volume_desc& v = d->volumes[i];
    if (!mat_ok(v.material)) return bad("volume material index out of range");
    if (v.shape != PT_VOLUME_SPHERE && v.shape != PT_VOLUME_BOX) return bad("invalid volume shape");
    if (!(v.density > 0.f) || !std::isfinite(v.density)) return bad("volume density must be positive and finite");
  }
  for (uint32_t i = 0; i < d->num_cylinders; i++)
    if (!mat_ok(d->cylinders[i].material)) return bad("cylinder material index out of range");
  for (uint32_t i = 0; i < d->num_moving_spheres; i++) {
    if (!mat_ok(d->moving_spheres[i].material)) return bad("moving sphere material index out of range");
    if (!(d->moving_spheres[i].time1 != d->moving_spheres[i].time0)) return bad("moving sphere needs time1 != time0");
  }
  for (uint32_t i = 0; i < d->num_meshes; i++) {
    const pt_mesh_desc& m = d->meshes[i];
    if (m.num_triangles == 0 || !m.vertices || !m.indices) return bad("mesh without vertices or triangles");
    for (uint64_t k = 0; k < 3ull * m.num_triangles; k++)
      if (m.indices[k] >= m.num_vertices) return bad("mesh index out of range");
    if (m.face_materials) {
      for (uint32_t k = 0; k < m.num_triangles; k++)
        if (!mat_ok(m.face_materials[k])) return bad("mesh face material out of range");
    } else if (!mat_ok(m.material)) {
      return bad("mesh material index out of range");
    }
  }
  for (uint32_t i = 0; i < d->num_instances; i++) {
    const pt_instance_desc& in = d->instances[i];
    uint64_t lim = in.kind == PT_GEOM_SPHERES ? d->num_spheres
                   : in.kind == PT_GEOM_RECTS ? d->num_rects
                   : in.kind == PT_GEOM_BOXES ? d->num_boxes
                   : in.kind == PT_GEOM_MESH ? d->num_meshes
                   : in.kind == PT_GEOM_VOLUMES ? d->num_volumes
                   : in.kind == PT_GEOM_CYLINDERS ? d->num_cylinders
                   : in.kind == PT_GEOM_MOVING_SPHERES ? d->num_moving_spheres
                                                       : 0;
    if (in.kind < 0 || in.kind > PT_GEOM_MOVING_SPHERES) return bad("invalid instance kind");
    if (in.kind == PT_GEOM_MESH) {
      if (in.first >= lim) return bad("instance mesh index out of range");
    } else if (in.count == 0 || (uint64_t)in.first + in.count > lim) {
      return bad("instance primitive range out of range");
    }
  }
  for (uint32_t i = 0; i < d->num_materials; i++) {
    const pt_material_desc& m = d->materials[i];
    if (m.type < 0 || m.type >= PT_MAT_COUNT) return bad("invalid material type");
    if (!tex_ok(m.tex[0]) && !(m.type == PT_MAT_NORMAL && m.tex[0] == -1))
      return bad("material texture index out of range");
    if ((m.type == PT_MAT_ASHIKHMIN_SHIRLEY || m.type == PT_MAT_DIELECTRIC) && !tex_ok(m.tex[1]))
      return bad("material second texture index out of range");
  }
  for (uint32_t i = 0; i < d->num_textures; i++) {
    const pt_texture_desc& t = d->textures[i];
    if (t.type == PT_TEX_IMAGE && (t.width == 0 || t.height == 0 || !t.texels)) return bad("empty image texture");
    if (t.type < PT_TEX_CONSTANT || t.type > PT_TEX_VECTOR) return bad("invalid texture type");
    if (t.type == PT_TEX_VECTOR) {
      if (t.width == 0 || !t.texels) return bad("empty vector texture");
      for (uint32_t k = 0; k < t.width; k++) {
        const float q = t.texels[k];
        if (!(q >= 0.f && q < (float)d->num_textures) || q != (float)(int)q) return bad("vector texture child out of range");
        const int32_t ty = d->textures[(int)q].type;
        if (ty == PT_TEX_VECTOR) return bad("nested vector textures are not supported");
        if (ty == PT_TEX_CHECKER) {
          const pt_texture_desc& ck = d->textures[(int)q];
          if (ck.width >= d->num_textures || ck.height >= d->num_textures) return bad("checker texture index out of range");
        }
      }
    }
    if ((t.type == PT_TEX_NOISE || t.type == PT_TEX_GRADIENT) && !t.texels) return bad("procedural texture without tables");
    if (t.type == PT_TEX_NOISE) {
      for (int k = 0; k < 3 * 256; k++) {
        const float q = t.texels[1024 + k];
        if (!(q >= 0.f && q <= 255.f) || q != (float)(int)q) return bad("noise permutation entry not in 0..255");
      }
    }
    if (t.type == PT_TEX_CHECKER) {
      if (t.width >= d->num_textures || t.height >= d->num_textures) return bad("checker texture index out of range");
      if (d->textures[t.width].type >= PT_TEX_CHECKER && d->textures[t.width].type != PT_TEX_NOISE &&
          d->textures[t.width].type != PT_TEX_GRADIENT)
        return bad("a checker's children are constant, image, noise or gradient textures");
      if (d->textures[t.height].type >= PT_TEX_CHECKER && d->textures[t.height].type != PT_TEX_NOISE &&
          d->textures[t.height].type != PT_TEX_GRADIENT)
        return bad("a checker's children are constant, image, noise or gradient textures");
    }
  }
  if (d->miss.type < 0 || d->miss.type > 3) return bad("invalid miss program");
  if ((d->miss.type == PT_MISS_IMAGE || d->miss.type == PT_MISS_HDR) && !tex_ok(d->miss.texture))
    return bad("miss texture index out of range");
  for (uint32_t i = 0; i < d->num_lights; i++) {
    const int32_t t = d->lights[i].type;
    if (t != PT_LIGHT_RECT && t != PT_LIGHT_SPHERE && t != PT_LIGHT_ENV)
      return fail(c, PT_ERR_UNSUPPORTED, "light type not supported");
    if (t == PT_LIGHT_ENV && (d->miss.type != PT_MISS_HDR || !d->miss.spherical ||
                              d->textures[d->miss.texture].type != PT_TEX_IMAGE))
      return bad("an environment light needs a spherical HDR miss texture");
  }

  // ---- host copies
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(c->scene_bufs);
  dfree(c->accel_bufs);
  c->accel_ready = false;
  c->paths_valid = false;
  c->spheres.assign(d->spheres, d->spheres + d->num_spheres);
  c->rects.assign(d->rects, d->rects + d->num_rects);
  c->boxes.assign(d->boxes, d->boxes + d->num_boxes);
  c->volumes.assign(d->volumes, d->volumes + d->num_volumes);
  c->cylinders.assign(d->cylinders, d->cylinders + d->num_cylinders);
  c->moving.assign(d->moving_spheres, d->moving_spheres + d->num_moving_spheres);
  c->meshes.clear();
  for (uint32_t i = 0; i < d->num_meshes; i++) {
    const pt_mesh_desc& m = d->meshes[i];
    pt_context::MeshHost h;
    h.v.assign(m.vertices, m.vertices + 3ull * m.num_vertices);
    if (m.normals) h.n.assign(m.normals, m.normals + 3ull * m.num_vertices);
    if (m.uvs) h.uv.assign(m.uvs, m.uvs + 2ull * m.num_vertices);
    h.idx.assign(m.indices, m.indices + 3ull * m.num_triangles);
    if (m.face_materials) h.fmat.assign(m.face_materials, m.face_materials + m.num_triangles);
    if (m.face_indices) h.fidx.assign(m.face_indices, m.face_indices + m.num_triangles);
    h.mat = m.material;
    c->meshes.push_back(std::move(h));
  }
  c->instances.assign(d->instances, d->instances + d->num_instances);
  c->materials.assign(d->materials, d->materials + d->num_materials);
  c->textures.assign(d->textures, d->textures + d->num_textures);
  c->texels.clear();
  for (auto& t : c->textures) {
    std::vector<float> tx;
    if (t.type == PT_TEX_IMAGE) tx.assign(t.texels, t.texels + 4ull * t.width * t.height);
    if (t.type == PT_TEX_NOISE) tx.assign(t.texels, t.texels + PT_NOISE_FLOATS);
    if (t.type == PT_TEX_GRADIENT) tx.assign(t.texels, t.texels + 12);
    if (t.type == PT_TEX_VECTOR) {
      tx.assign(t.texels, t.texels + t.width);
      tx.resize((tx.size() + 3) / 4 * 4, 0.f);  // whole float4 texels
    }
    c->texels.push_back(std::move(tx));
    t.texels = nullptr;
  }
  c->lights.assign(d->lights, d->lights + d->num_lights);
  c->miss = d->miss;
  {
    bool seen[PT_MAT_COUNT] = {};
    c->shading_kinds = 0;
    c->mat_types = 0;
    for (auto& m : c->materials) c->mat_types |= 1u << m.type;
    for (auto& m : c->materials)
      if (m.type != PT_MAT_DIFFUSE_LIGHT && !seen[m.type]) {
        seen[m.type] = true;
        c->shading_kinds++;
      }
  }

  // ---- original-order device data
  DevScene& s = c->dsc;
  s = DevScene{};
  std::vector<float4> sph_o(c->spheres.size());
  std::vector<int32_t> sph_mat(c->spheres.size());
  for (size_t i = 0; i < c->spheres.size(); i++) {
    auto& q = c->spheres[i];
    sph_o[i] = make_float4(q.center[0], q.center[1], q.center[2], q.radius);
    sph_mat[i] = q.material;
  }
  std::vector<float4> rect_o(2 * c->rects.size());
  std::vector<int32_t> rect_mat(c->rects.size());
  for (size_t i = 0; i < c->rects.size(); i++) {
    auto& q = c->rects[i];
    rect_o[2 * i] = make_float4(q.a0, q.a1, q.b0, q.b1);
    rect_o[2 * i + 1] = make_float4(q.k, bf((uint32_t)q.axis), bf(q.flip ? 1u : 0u), bf((uint32_t)i));
    rect_mat[i] = q.material;
  }
  std::vector<float4> box_o(2 * c->boxes.size());
  std::vector<int32_t> box_mat(c->boxes.size());
  for (size_t i = 0; i < c->boxes.size(); i++) {
    auto& q = c->boxes[i];
    box_o[2 * i] = make_float4(q.pmin[0], q.pmin[1], q.pmin[2], bf((uint32_t)i));
    box_o[2 * i + 1] = make_float4(q.pmax[0], q.pmax[1], q.pmax[2], 0.f);
    box_mat[i] = q.material;
  }
  std::vector<float4> vol_o(2 * c->volumes.size());
  std::vector<int32_t> vol_mat(c->volumes.size());
  for (size_t i = 0; i < c->volumes.size(); i++) {
    auto& q = c->volumes[i];
    const uint32_t box = q.shape == PT_VOLUME_BOX ? 1u : 0u;
    vol_o[2 * i] = make_float4(q.a[0], q.a[1], q.a[2], q.density);
    vol_o[2 * i + 1] = make_float4(q.b[0], q.b[1], q.b[2], bf((uint32_t)i | (box << 31)));
    vol_mat[i] = q.material;
  }
  std::vector<float4> cyl_o(2 * c->cylinders.size());
  std::vector<int32_t> cyl_mat(c->cylinders.size());
  for (size_t i = 0; i < c->cylinders.size(); i++) {
    auto& q = c->cylinders[i];
    cyl_o[2 * i] = make_float4(q.origin[0], q.origin[1], q.origin[2], q.length);
    cyl_o[2 * i + 1] = make_float4(q.radius, 0.f, 0.f, bf((uint32_t)i));
    cyl_mat[i] = q.material;
  }
  std::vector<float4> mov_o(3 * c->moving.size());
  std::vector<int32_t> mov_mat(c->moving.size());
  for (size_t i = 0; i < c->moving.size(); i++) {
    auto& q = c->moving[i];
    mov_o[3 * i] = make_float4(q.center0[0], q.center0[1], q.center0[2], q.radius);
    mov_o[3 * i + 1] = make_float4(q.center1[0], q.center1[1], q.center1[2], bf((uint32_t)i));
    mov_o[3 * i + 2] = make_float4(q.time0, q.time1, 0.f, 0.f);
    mov_mat[i] = q.material;
  }
  std::vector<DevMesh> dmesh;
  std::vector<float4> vtx, nrm;
  std::vector<float2> uv;
  std::vector<uint4> idx;
  std::vector<int32_t> fmat, fidx;
  for (auto& m : c->meshes) {
    DevMesh dm{};
    uint32_t nv = (uint32_t)(m.v.size() / 3);
    dm.vbase = (uint32_t)vtx.size();
    dm.ibase = (uint32_t)idx.size();
    dm.ntri = (uint32_t)(m.idx.size() / 3);
    dm.has_n = !m.n.empty();
    dm.has_uv = !m.uv.empty();
    dm.mat = m.mat;
    dm.fmat_base = m.fmat.empty() ? -1 : (int32_t)fmat.size();
    dm.fidx_base = m.fidx.empty() ? -1 : (int32_t)fidx.size();
    fidx.insert(fidx.end(), m.fidx.begin(), m.fidx.end());
    for (uint32_t k = 0; k < nv; k++) {
      vtx.push_back(make_float4(m.v[3 * k], m.v[3 * k + 1], m.v[3 * k + 2], 0.f));
      nrm.push_back(dm.has_n ? make_float4(m.n[3 * k], m.n[3 * k + 1], m.n[3 * k + 2], 0.f) : make_float4(0, 0, 0, 0));
      uv.push_back(dm.has_uv ? make_float2(m.uv[2 * k], m.uv[2 * k + 1]) : make_float2(0, 0));
    }
    for (uint32_t k = 0; k < dm.ntri; k++) idx.push_back(make_uint4(m.idx[3 * k], m.idx[3 * k + 1], m.idx[3 * k + 2], 0));
    fmat.insert(fmat.end(), m.fmat.begin(), m.fmat.end());
    dmesh.push_back(dm);
  }
  std::vector<DevMaterial> mats;
  for (auto& m : c->materials) {
    DevMaterial dm{m.type, m.tex[0], m.tex[1], m.param[0], m.param[1], m.param[2], m.param[3], 0};
    mats.push_back(dm);
  }
  std::vector<DevTexture> texs;
  std::vector<float4> texel_pool;
  for (size_t i = 0; i < c->textures.size(); i++) {
    auto& t = c->textures[i];
    DevTexture dt{t.type, t.rgb[0], t.rgb[1], t.rgb[2], t.width, t.height, (uint32_t)texel_pool.size(), 0};
    const auto& tx = c->texels[i];
    for (size_t k = 0; k + 3 < tx.size(); k += 4) texel_pool.push_back(make_float4(tx[k], tx[k + 1], tx[k + 2], tx[k + 3]));
    texs.push_back(dt);
  }
  std::vector<DevLight> lts, lts_plain;
  for (auto& l : c->lights) {
    DevLight dl{l.type, l.axis, l.a0, l.a1, l.b0, l.b1, l.k, l.center[0], l.center[1], l.center[2], l.radius,
                l.emission[0], l.emission[1], l.emission[2], {0.f, 0.f}};
    lts.push_back(dl);
    if (l.type != PT_LIGHT_ENV) lts_plain.push_back(dl);
  }
  auto& L = c->scene_bufs;
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.sph_o), sph_o.data(), sph_o.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.sph_mat), sph_mat.data(), sph_mat.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.rect_o), rect_o.data(), rect_o.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.rect_mat), rect_mat.data(), rect_mat.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.box_o), box_o.data(), box_o.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.box_mat), box_mat.data(), box_mat.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.vol_o), vol_o.data(), vol_o.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.vol_mat), vol_mat.data(), vol_mat.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.cyl_o), cyl_o.data(), cyl_o.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.cyl_mat), cyl_mat.data(), cyl_mat.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.mov_o), mov_o.data(), mov_o.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.mov_mat), mov_mat.data(), mov_mat.size() * 4));
  s.ext_instances = 0;
  s.mov_instances = 0;
  s.kinds = 0;
  for (auto& in : c->instances) {
    if (in.kind <= PT_GEOM_MESH) s.kinds |= 1u << in.kind;
    s.ext_instances += in.kind >= PT_GEOM_VOLUMES ? 1u : 0u;
    s.mov_instances += in.kind == PT_GEOM_MOVING_SPHERES ? 1u : 0u;
  }
  HIPCHK(c, upload(c, L, const_cast<DevMesh**>(&s.meshes), dmesh.data(), dmesh.size() * sizeof(DevMesh)));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.vtx), vtx.data(), vtx.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.nrm), nrm.data(), nrm.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<float2**>(&s.uv), uv.data(), uv.size() * sizeof(float2)));
  HIPCHK(c, upload(c, L, const_cast<uint4**>(&s.idx), idx.data(), idx.size() * sizeof(uint4)));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.fmat), fmat.data(), fmat.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<int32_t**>(&s.fidx), fidx.data(), fidx.size() * 4));
  HIPCHK(c, upload(c, L, const_cast<DevMaterial**>(&s.mats), mats.data(), mats.size() * sizeof(DevMaterial)));
  HIPCHK(c, upload(c, L, const_cast<DevTexture**>(&s.texs), texs.data(), texs.size() * sizeof(DevTexture)));
  HIPCHK(c, upload(c, L, const_cast<float4**>(&s.texels), texel_pool.data(), texel_pool.size() * sizeof(float4)));
  HIPCHK(c, upload(c, L, const_cast<DevLight**>(&s.lights), lts.data(), lts.size() * sizeof(DevLight)));
  HIPCHK(c, upload(c, L, const_cast<DevLight**>(&s.lights_plain), lts_plain.data(),
                   lts_plain.size() * sizeof(DevLight)));
  s.num_lights = (int32_t)lts.size();
  s.num_lights_plain = (int32_t)lts_plain.size();
  s.miss = c->miss;
  // environment importance-sampling tables, built on the GPU
  s.env_cond = s.env_marg = nullptr;
  s.env_guide = nullptr;
  s.env_total = 0;
  s.env_w = s.env_h = 0;
  if (lts.size() != lts_plain.size()) {
    const DevTexture& t = texs[c->miss.texture];
    unsigned long long *cond = nullptr, *marg = nullptr, *scratch = nullptr;
    hipError_t e;
    if ((e = dalloc(L, (void**)&cond, (size_t)t.w * t.h * 8)) != hipSuccess ||
        (e = dalloc(L, (void**)&marg, (size_t)t.h * 8)) != hipSuccess ||
        (e = dalloc(L, (void**)&scratch, (size_t)(t.h + 1) * 8)) != hipSuccess)
      return hip_fail(c, e, "hipMalloc (environment tables)");
    uint32_t* guide = nullptr;
    if ((e = dalloc(L, (void**)&guide, (size_t)(t.h + 1) * (ENV_GUIDE + 1) * 4)) != hipSuccess)
      return hip_fail(c, e, "hipMalloc (environment guide tables)");
    launch_env_tables(s.texels + t.offset, t.w, t.h, cond, marg, scratch, c->stream);
    launch_env_guides(cond, marg, t.w, t.h, guide, c->stream);
    s.env_guide = guide;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long total = 0;
    HIPCHK(c, hipMemcpy(&total, marg + (t.h - 1), 8, hipMemcpyDeviceToHost));
    s.env_cond = cond;
    s.env_marg = marg;
    s.env_total = total;
    s.env_w = t.w;
    s.env_h = t.h;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->has_scene = true;
  return PT_OK;
}

pt_status pt_build_accel(pt_context* c, pt_accel_stats* out) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!c->has_scene) return fail(c, PT_ERR_STATE, "pt_build_accel before pt_set_scene");
  (void)hipSetDevice(c->device);
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  dfree(c->accel_bufs);
  c->accel_ready = false;
  DevScene& s = c->dsc;
  auto& A = c->accel_bufs;

  // ---- bottom-level structures, one per distinct (kind, first, count)
  std::map<std::tuple<int, uint32_t, uint32_t>, int> blas_of;
  std::vector<int> inst_blas(c->instances.size());
  c->blas.clear();
  uint64_t maxn = c->instances.size();
  uint32_t pool_n[7] = {0, 0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < c->instances.size(); i++) {
    const auto& in = c->instances[i];
    uint32_t count = in.kind == PT_GEOM_MESH ? (uint32_t)(c->meshes[in.first].idx.size() / 3) : in.count;
    auto key = std::make_tuple((int)in.kind, in.first, count);
    auto it = blas_of.find(key);
    if (it == blas_of.end()) {
      BlasInfo b{in.kind, in.first, count, 0, 0, pool_n[in.kind]};
      pool_n[in.kind] += count;
      it = blas_of.emplace(key, (int)c->blas.size()).first;
      c->blas.push_back(b);
      maxn = std::max<uint64_t>(maxn, count);
    }
    inst_blas[i] = it->second;
  }
  // ---- build scratch
  BuildScratch sb{};
  std::vector<DevBuf> S;
  uint64_t cap = maxn + 1;
  uint64_t nblk = (cap + 2047) / 2048;
  auto A32 = [&](uint32_t** p, uint64_t n) { return dalloc(S, (void**)p, n * 4); };
  auto AF4 = [&](float4** p, uint64_t n) { return dalloc(S, (void**)p, n * sizeof(float4)); };
  hipError_t e = hipSuccess;
  if ((e = AF4(&sb.lo, cap)) || (e = AF4(&sb.hi, cap)) || (e = A32(&sb.code_a, cap)) || (e = A32(&sb.code_b, cap)) ||
      (e = A32(&sb.idx_a, cap)) || (e = A32(&sb.idx_b, cap)) || (e = A32(&sb.hist, 256 * nblk + 64)) ||
      (e = A32(&sb.left, cap)) || (e = A32(&sb.right, cap)) || (e = A32(&sb.first, cap)) ||
      (e = A32(&sb.last, cap)) || (e = A32(&sb.parent, 2 * cap)) || (e = A32(&sb.flags, cap)) ||
      (e = AF4(&sb.nlo, cap)) || (e = AF4(&sb.nhi, cap)) || (e = A32(&sb.frontier_a, cap)) ||
      (e = A32(&sb.frontier_b, cap)) || (e = A32(&sb.slots, 4 * cap)) || (e = A32(&sb.nchild, cap + 64)) ||
      (e = A32(&sb.scan_tmp, (256 * nblk + cap) / 256 + 1024)) || (e = A32(&sb.cbounds, 8)) ||
      (e = A32(&sb.misc, 64))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc (build scratch)");
  }
  sb.capacity = cap;

  // per-BLAS outputs into temporaries, then compacted
  std::vector<uint4*> bnodes(c->blas.size());
  std::vector<uint32_t> bcount(c->blas.size());
  float4* roots = nullptr;  // 2 per BLAS
  if ((e = dalloc(A, (void**)&roots, 2 * sizeof(float4) * (c->blas.size() + 1)))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc");
  }
  uint64_t total_sorted = 0;
  for (auto& b : c->blas) total_sorted += b.count;
  total_sorted += c->instances.size();
  if ((e = dalloc(A, (void**)&c->order_dev, total_sorted * 4)) ||
      (e = dalloc(A, (void**)&c->morton_dev, total_sorted * 4))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc");
  }
  // sorted primitive pools
  float4 *sph = nullptr, *rect = nullptr, *box = nullptr, *tri = nullptr, *vol = nullptr, *cyl = nullptr,
         *mov = nullptr;
  uint32_t* sph_idx = nullptr;
  if ((e = dalloc(A, (void**)&sph, pool_n[0] * sizeof(float4))) ||
      (e = dalloc(A, (void**)&sph_idx, pool_n[0] * 4)) ||
      (e = dalloc(A, (void**)&rect, 2ull * pool_n[1] * sizeof(float4))) ||
      (e = dalloc(A, (void**)&box, 2ull * pool_n[2] * sizeof(float4))) ||
      (e = dalloc(A, (void**)&tri, 3ull * pool_n[3] * sizeof(float4))) ||
      (e = dalloc(A, (void**)&vol, 2ull * pool_n[4] * sizeof(float4))) ||
      (e = dalloc(A, (void**)&cyl, 2ull * pool_n[5] * sizeof(float4))) ||
      (e = dalloc(A, (void**)&mov, 3ull * pool_n[6] * sizeof(float4)))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc (primitive pools)");
  }
  uint64_t order_off = 0;
  std::vector<uint64_t> order_base(c->blas.size());
  for (size_t b = 0; b < c->blas.size(); b++) {
    BlasInfo& bi = c->blas[b];
    uint32_t n = bi.count;
    switch (bi.kind) {
      case PT_GEOM_SPHERES: launch_bounds_spheres(s.sph_o + bi.first, n, sb.lo, sb.hi, c->stream); break;
      case PT_GEOM_RECTS: launch_bounds_rects(s.rect_o + 2ull * bi.first, n, sb.lo, sb.hi, c->stream); break;
      case PT_GEOM_BOXES: launch_bounds_boxes(s.box_o + 2ull * bi.first, n, sb.lo, sb.hi, c->stream); break;
      case PT_GEOM_VOLUMES: launch_bounds_volumes(s.vol_o + 2ull * bi.first, n, sb.lo, sb.hi, c->stream); break;
      case PT_GEOM_CYLINDERS: launch_bounds_cylinders(s.cyl_o + 2ull * bi.first, n, sb.lo, sb.hi, c->stream); break;
      case PT_GEOM_MOVING_SPHERES:
        launch_bounds_moving(s.mov_o + 3ull * bi.first, n, sb.lo, sb.hi, c->stream);
        break;
      default: {
        // device mesh table offsets are computed in set_scene in mesh order
        uint32_t vbase = 0, ibase = 0;
        for (uint32_t m = 0; m < bi.first; m++) {
          vbase += (uint32_t)(c->meshes[m].v.size() / 3);
          ibase += (uint32_t)(c->meshes[m].idx.size() / 3);
        }
        launch_bounds_tris(s.vtx + vbase, s.idx + ibase, n, sb.lo, sb.hi, c->stream);
      }
    }
    uint4* nodes_tmp = nullptr;
    if ((e = dalloc(S, (void**)&nodes_tmp, (size_t)std::max<uint32_t>(1, n) * 128))) {
      dfree(S);
      return hip_fail(c, e, "hipMalloc (nodes)");
    }
    uint32_t nn = build_bvh4(sb, n, BLAS_LEAF_MAX, nodes_tmp, c->order_dev + order_off, c->morton_dev + order_off,
                             roots + 2 * b, c->stream);
    if (nn == 0) {
      dfree(S);
      return fail(c, PT_ERR_DEVICE, "bottom-level build failed");
    }
    bnodes[b] = nodes_tmp;
    bcount[b] = nn;
    order_base[b] = order_off;
    // sorted primitive records of this structure
    const uint32_t* ord = c->order_dev + order_off;
    switch (bi.kind) {
      case PT_GEOM_SPHERES:
        launch_gather_records(s.sph_o + bi.first, ord, n, 1, sph + bi.prim_base, bi.first, 0, c->stream);
        launch_gather_records(s.sph_o + bi.first, ord, n, 1, (float4*)(sph_idx + bi.prim_base), bi.first, 1,
                              c->stream);
        break;
      case PT_GEOM_RECTS:
        launch_gather_records(s.rect_o + 2ull * bi.first, ord, n, 2, rect + 2ull * bi.prim_base, bi.first, 0,
                              c->stream);
        break;
      case PT_GEOM_BOXES:
        launch_gather_records(s.box_o + 2ull * bi.first, ord, n, 2, box + 2ull * bi.prim_base, bi.first, 0,
                              c->stream);
        break;
      case PT_GEOM_VOLUMES:
        launch_gather_records(s.vol_o + 2ull * bi.first, ord, n, 2, vol + 2ull * bi.prim_base, bi.first, 0,
                              c->stream);
        break;
      case PT_GEOM_CYLINDERS:
        launch_gather_records(s.cyl_o + 2ull * bi.first, ord, n, 2, cyl + 2ull * bi.prim_base, bi.first, 0,
                              c->stream);
        break;
      case PT_GEOM_MOVING_SPHERES:
        launch_gather_records(s.mov_o + 3ull * bi.first, ord, n, 3, mov + 3ull * bi.prim_base, bi.first, 0,
                              c->stream);
        break;
      default: {
        uint32_t vbase = 0, ibase = 0;
        for (uint32_t m = 0; m < bi.first; m++) {
          vbase += (uint32_t)(c->meshes[m].v.size() / 3);
          ibase += (uint32_t)(c->meshes[m].idx.size() / 3);
        }
        launch_gather_tris(s.vtx + vbase, s.idx + ibase, ord, n, tri + 3ull * bi.prim_base, c->stream);
      }
    }
    order_off += n;
  }
  uint64_t blas_nodes = 0;
  for (size_t b = 0; b < c->blas.size(); b++) {
    c->blas[b].node_base = (uint32_t)blas_nodes;
    c->blas[b].node_count = bcount[b];
    blas_nodes += bcount[b];
  }
  // ---- instances (unsorted) and the top level
  const uint32_t ni = (uint32_t)c->instances.size();
  std::vector<uint4> irec(8ull * ni);
  std::vector<uint32_t> mesh_of(ni, 0);
  for (uint32_t i = 0; i < ni; i++) {
    const auto& in = c->instances[i];
    const BlasInfo& bi = c->blas[inst_blas[i]];
    const float* w = in.world_to_obj;
    const float* m = in.obj_to_world;
    uint4* r = &irec[8ull * i];
    r[0] = make_uint4(fb(w[0]), fb(w[1]), fb(w[2]), fb(w[3]));
    r[1] = make_uint4(fb(w[4]), fb(w[5]), fb(w[6]), fb(w[7]));
    r[2] = make_uint4(fb(w[8]), fb(w[9]), fb(w[10]), fb(w[11]));
    r[3] = make_uint4(bi.node_base, bi.prim_base, (uint32_t)in.kind | ((in.has_transform ? 1u : 0u) << 8), i);
    r[4] = make_uint4(fb(m[0]), fb(m[1]), fb(m[2]), fb(m[3]));
    r[5] = make_uint4(fb(m[4]), fb(m[5]), fb(m[6]), fb(m[7]));
    r[6] = make_uint4(fb(m[8]), fb(m[9]), fb(m[10]), fb(m[11]));
    r[7] = make_uint4(in.first, bi.count, in.kind == PT_GEOM_MESH ? in.first : 0u, (uint32_t)inst_blas[i]);
  }
  uint4* irec_dev = nullptr;
  if ((e = upload(c, S, &irec_dev, irec.data(), irec.size() * sizeof(uint4)))) {
    dfree(S);
    return hip_fail(c, e, "upload instances");
  }
  launch_bounds_instances(irec_dev, roots, ni, sb.lo, sb.hi, c->stream);
  uint4* tnodes = nullptr;
  if ((e = dalloc(S, (void**)&tnodes, (size_t)std::max<uint32_t>(1, ni) * 128))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc (tlas)");
  }
  uint32_t tn = build_bvh4(sb, ni, TLAS_LEAF_MAX, tnodes, c->order_dev + order_off, c->morton_dev + order_off,
                           roots + 2 * c->blas.size(), c->stream);
  if (tn == 0) {
    dfree(S);
    return fail(c, PT_ERR_DEVICE, "top-level build failed");
  }
  uint4* insts_sorted = nullptr;
  if ((e = dalloc(A, (void**)&insts_sorted, irec.size() * sizeof(uint4)))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc");
  }
  launch_gather_instances(irec_dev, c->order_dev + order_off, ni, insts_sorted, c->stream);
  // ---- compact all nodes into one array (BLAS after BLAS, then TLAS)
  c->total_nodes = blas_nodes + tn;
  if ((e = dalloc(A, (void**)&c->nodes_dev, c->total_nodes * 128))) {
    dfree(S);
    return hip_fail(c, e, "hipMalloc (nodes)");
  }
  for (size_t b = 0; b < c->blas.size(); b++)
    HIPCHK(c, hipMemcpyAsync(c->nodes_dev + 8ull * c->blas[b].node_base, bnodes[b], (size_t)bcount[b] * 128,
                             hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->nodes_dev + 8ull * blas_nodes, tnodes, (size_t)tn * 128, hipMemcpyDeviceToDevice,
                           c->stream));
  HIPCHK(c, hipEventRecord(e1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(S);
  c->tlas_nodes = tn;
  c->total_prims_sorted = total_sorted;
  s.nodes = c->nodes_dev;
  s.tlas_root = (uint32_t)blas_nodes;
  {
    // scene box = union of the top-level root's children
    uint32_t w[NODE_WORDS];
    HIPCHK(c, hipMemcpy(w, c->nodes_dev + 8ull * blas_nodes, 128, hipMemcpyDeviceToHost));
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = 0; k < 4; k++) {
      if (w[24 + k] == EMPTY_CHILD) continue;
      for (int a = 0; a < 3; a++) {
        float l, h;
        std::memcpy(&l, &w[8 * a + k], 4);
        std::memcpy(&h, &w[8 * a + 4 + k], 4);
        lo[a] = std::min(lo[a], l);
        hi[a] = std::max(hi[a], h);
      }
    }
    for (int a = 0; a < 3; a++) {
      c->scene_lo[a] = lo[a];
      c->scene_hi[a] = hi[a] > lo[a] ? hi[a] : lo[a] + 1.f;
    }
  }
  s.num_instances = ni;
  s.insts = insts_sorted;
  s.sph = sph;
  s.sph_idx = sph_idx;
  s.rect = rect;
  s.box = box;
  s.tri = tri;
  s.vol = vol;
  s.cyl = cyl;
  s.mov = mov;
  c->accel_bytes = c->total_nodes * 128 + pool_n[0] * 20ull + pool_n[1] * 32ull + pool_n[2] * 32ull +
                   pool_n[3] * 48ull + pool_n[4] * 32ull + pool_n[5] * 32ull + pool_n[6] * 48ull + ni * 128ull;
  c->accel_ready = true;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (out) {
    out->num_blas = (uint32_t)c->blas.size();
    out->num_instances = ni;
    out->num_primitives = total_sorted - ni;
    out->num_nodes = c->total_nodes;
    out->bytes = c->accel_bytes;
    out->build_ms = ms;
  }
  return PT_OK;
}

pt_status pt_read_accel(pt_context* c, uint32_t* nodes, uint64_t* num_nodes, uint32_t* prim_order, uint32_t* morton,
                        uint64_t* num_prims, uint32_t* blas_info, uint32_t* num_blas) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!c->accel_ready) return fail(c, PT_ERR_STATE, "acceleration structure not built");
  (void)hipSetDevice(c->device);
  if (num_nodes) *num_nodes = c->total_nodes;
  if (num_prims) *num_prims = c->total_prims_sorted;
  if (num_blas) *num_blas = (uint32_t)c->blas.size();
  if (nodes) HIPCHK(c, hipMemcpy(nodes, c->nodes_dev, c->total_nodes * 128, hipMemcpyDeviceToHost));
  if (prim_order) HIPCHK(c, hipMemcpy(prim_order, c->order_dev, c->total_prims_sorted * 4, hipMemcpyDeviceToHost));
  if (morton) HIPCHK(c, hipMemcpy(morton, c->morton_dev, c->total_prims_sorted * 4, hipMemcpyDeviceToHost));
  if (blas_info) {
    uint64_t pb = 0;
    for (size_t b = 0; b < c->blas.size(); b++) {
      blas_info[6 * b + 0] = c->blas[b].node_base;
      blas_info[6 * b + 1] = c->blas[b].node_count;
      blas_info[6 * b + 2] = (uint32_t)pb;
      blas_info[6 * b + 3] = c->blas[b].count;
      blas_info[6 * b + 4] = (uint32_t)c->blas[b].kind;
      blas_info[6 * b + 5] = c->blas[b].first;
      pb += c->blas[b].count;
    }
  }
  return PT_OK;
}

pt_status pt_set_camera(pt_context* c, const pt_camera* cam) {
  if (!c || !cam) return PT_ERR_INVALID_ARG;
  c->cam = *cam;
  c->has_camera = true;
  return PT_OK;
}

// traversal-stack spill area of the grid-stride kernels, allocated by the
// first call that traces (a context that only takes part in batches run by
// another context never needs one)
static pt_status ensure_spill(pt_context* c) {
  if (c->spill) return PT_OK;
  HIPCHK(c, hipMalloc(&c->spill, (size_t)GRID_CAP * BLOCK * STK_TOTAL * sizeof(uint2)));
  return PT_OK;
}

static pt_status ensure_paths(pt_context* c, uint64_t cap) {
  if (cap <= c->path_capacity) return PT_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(c->path_bufs);
  c->path_capacity = 0;
  DevPaths& P = c->paths;
  auto& B = c->path_bufs;
  hipError_t e;
  auto F4 = [&](float4** p) { return dalloc(B, (void**)p, cap * sizeof(float4)); };
  auto U32 = [&](uint32_t** p) { return dalloc(B, (void**)p, cap * 4); };
  e = hipSuccess;
  for (int k = 0; k < 3 && e == hipSuccess; k++)
    if ((e = F4(&P.st[k].o)) == hipSuccess && (e = F4(&P.st[k].d)) == hipSuccess && (e = F4(&P.st[k].T)) == hipSuccess)
      e = F4(&P.st[k].L);
  if (e || (e = F4(&P.hit)) || (e = F4(&P.hit_sorted)) || (e = U32(&P.keys)) || (e = F4(&P.result)) ||
      (e = U32(&P.sig_final)) || (e = U32((uint32_t**)&P.ptime)) || (e = dalloc(B, (void**)&P.counters, CTR_WORDS * 4)) ||
      (e = dalloc(B, (void**)&P.seg, (size_t)3 * MAX_WAVES * 4)) ||
      (e = dalloc(B, (void**)&P.bcount, (size_t)8 * MAX_WAVES * 4)) ||
      (e = dalloc(B, (void**)&P.bbase, (size_t)8 * MAX_WAVES * 4)) ||
      (e = dalloc(B, (void**)&P.tseg, (size_t)MAX_WAVES * 4)) || (e = dalloc(B, (void**)&P.qseg, (size_t)MAX_WAVES * 4)) ||
      (e = dalloc(B, (void**)&P.tpre, (size_t)MAX_WAVES * 4)) || (e = dalloc(B, (void**)&P.qpre, (size_t)MAX_WAVES * 4))) {
    dfree(B);
    return hip_fail(c, e, "hipMalloc (path state)");
  }
  launch_fill_u32(P.counters, CTR_WORDS, 0u, c->stream);
  c->path_capacity = cap;
  return PT_OK;
}

static pt_status cache_work(pt_context* c, uint32_t n);
static pt_status cache_reset(pt_context* c);
static pt_status ensure_cache_judge(pt_context* c);

// radiance-cache record buffers for `cap` path slots: one query per path;
// training vertices for a quarter of the slots, split among the waves of a
// batch like the path state (3 % of the paths, a few vertices each, fit with
// a wide margin; a wave's vertices beyond its share are dropped)
static pt_status ensure_cache_records(pt_context* c, uint64_t cap) {
  if (cap <= c->cache_rec_cap) return PT_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(c->crec_bufs);
  c->cache_rec_cap = 0;
  DevPaths& P = c->paths;
  auto& B = c->crec_bufs;
  const uint64_t tcap = std::max<uint64_t>(65536, cap / 4);
  hipError_t e;
  if ((e = dalloc(B, (void**)&P.q_feat, cap * 12 * 16)) != hipSuccess ||
      (e = dalloc(B, (void**)&P.q_meta, cap * 16)) != hipSuccess ||
      (e = dalloc(B, (void**)&P.q_out, cap * 12)) != hipSuccess ||
      (e = dalloc(B, (void**)&P.t_feat, tcap * 12 * 16)) != hipSuccess ||
      (e = dalloc(B, (void**)&P.t_L, tcap * 16)) != hipSuccess ||
      (e = dalloc(B, (void**)&P.t_T, tcap * 16)) != hipSuccess) {
    dfree(B);
    return hip_fail(c, e, "hipMalloc (radiance-cache records)");
  }
  P.q_cap = (uint32_t)cap;
  P.t_cap = (uint32_t)tcap;
  c->cache_rec_cap = cap;
  return PT_OK;
}

// per-launch constants of a render of context c with parameters rp
static RenderConsts render_consts(const pt_context* c, const pt_render_params* rp) {
  RenderConsts rc{};
  rc.W = rp->width;
  rc.H = rp->height;
  rc.WH = rp->width * rp->height;
  rc.max_depth = rp->max_depth;
  rc.russian = rp->russian;
  rc.rr_min_depth = rp->rr_min_depth;
  rc.flags = rp->flags;
  const pt_camera& cm = c->cam;
  for (int k = 0; k < 3; k++) {
    rc.cam.llc[k] = cm.lower_left[k];
    rc.cam.hor[k] = cm.horizontal[k];
    rc.cam.ver[k] = cm.vertical[k];
    rc.cam.org[k] = cm.origin[k];
    rc.cam.u[k] = cm.u[k];
    rc.cam.v[k] = cm.v[k];
  }
  for (int k = 0; k < 3; k++) {
    rc.cache.lo[k] = c->scene_lo[k];
    rc.cache.inv_ext[k] = 1.f / (c->scene_hi[k] - c->scene_lo[k]);
  }
  rc.cache.query_depth = c->cache_query_depth;
  rc.cache.train_threshold = (uint32_t)std::min(16777216.0, std::max(0.0, (double)c->cache_train_fraction * 16777216.0));
  rc.cam.lens_radius = cm.lens_radius;
  rc.store_time = c->dsc.mov_instances != 0 ? 1u : 0u;
  rc.cam.time0 = cm.time0;
  rc.cam.time1 = cm.time1;
  return rc;
}

pt_status pt_render(pt_context* c, const pt_render_params* rp, uint32_t frame_begin, uint32_t frame_count,
                    pt_render_stats* out) {
  if (!c || !rp) return PT_ERR_INVALID_ARG;
  if (!c->accel_ready) return fail(c, PT_ERR_STATE, "pt_render before pt_build_accel");
  if (!c->has_camera) return fail(c, PT_ERR_STATE, "pt_render before pt_set_camera");
  if (rp->width == 0 || rp->height == 0) return fail(c, PT_ERR_INVALID_ARG, "width and height must be positive");
  if ((uint64_t)rp->width * rp->height > (1ull << 28)) return fail(c, PT_ERR_INVALID_ARG, "image too large");
  if (rp->max_depth == 0) return fail(c, PT_ERR_INVALID_ARG, "max_depth must be positive");
  (void)hipSetDevice(c->device);
  const uint32_t WH = rp->width * rp->height;
  // accumulation buffers
  if (frame_begin == 0 || rp->width != c->W || rp->height != c->H || !c->accum) {
    if (frame_begin != 0 && (rp->width != c->W || rp->height != c->H))
      return fail(c, PT_ERR_STATE, "resolution changed without restarting at frame 0");
    if (rp->width != c->W || rp->height != c->H || !c->accum) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (c->accum) (void)hipFree(c->accum);
      if (c->rgba) (void)hipFree(c->rgba);
      c->accum = nullptr;
      c->rgba = nullptr;
      HIPCHK(c, hipMalloc(&c->accum, (size_t)WH * 3 * sizeof(double)));
      HIPCHK(c, hipMalloc(&c->rgba, (size_t)WH * sizeof(uchar4)));
      c->W = rp->width;
      c->H = rp->height;
    }
    launch_fill_f64(c->accum, (uint64_t)WH * 3, 0.0, c->stream);
    c->frames = 0;
  }
  if (frame_count == 0) return PT_OK;
  // launch shape: the autotuned one for this scene and resolution when asked
  // for and available, else frames_per_batch or the automatic batch size
  const uint32_t tune_bits =
      PT_FLAG_TUNED | PT_FLAG_GRAPH | PT_FLAG_MATERIAL_SORT | PT_FLAG_DIRECT | PT_FLAG_COUNT_TRAVERSAL;
  const pt_context::Tuned& tu = c->tuned;
  const bool tuned = (rp->flags & PT_FLAG_TUNED) && tu.valid && tu.W == rp->width && tu.H == rp->height &&
                     tu.max_depth == rp->max_depth && (tu.flags & ~tune_bits) == (rp->flags & ~tune_bits);
  uint32_t fpb = tuned ? tu.frames_per_batch
                       : (rp->frames_per_batch ? rp->frames_per_batch
                                               : (uint32_t)std::max<uint64_t>(1, c->paths_target / WH));
  fpb = std::min(fpb, frame_count);
  if ((uint64_t)fpb * WH > SLOT_MASK) return fail(c, PT_ERR_INVALID_ARG, "frames_per_batch x W x H exceeds 2^30 - 1");
  const bool use_cache = (rp->flags & PT_FLAG_RADIANCE_CACHE) != 0;
  if (use_cache && !c->cache_ready) return fail(c, PT_ERR_STATE, "PT_FLAG_RADIANCE_CACHE needs pt_cache_init");
  if (use_cache && (c->dsc.ext_instances != 0 || (c->mat_types >> PT_MAT_ISOTROPIC) != 0))
    return fail(c, PT_ERR_UNSUPPORTED, "the radiance cache does not take volumes, isotropic or normal-shader materials");
  pt_status ps = ensure_paths(c, (uint64_t)fpb * WH);
  if (ps == PT_OK) ps = ensure_spill(c);
  if (ps != PT_OK) return ps;
  if (use_cache) {
    if ((ps = ensure_cache_records(c, (uint64_t)fpb * WH)) != PT_OK) return ps;
    if ((ps = ensure_cache_judge(c)) != PT_OK) return ps;
    c->cache_last_judge = -1.0;
    // dense training rows of a batch: steps x batch, at most every record
    const uint64_t rows = std::min<uint64_t>((uint64_t)c->cache_steps * c->cache_batch, c->paths.t_cap);
    if ((ps = cache_work(c, (uint32_t)std::max<uint64_t>(rows, c->cache_batch))) != PT_OK) return ps;
  }

  RenderConsts rc = render_consts(c, rp);

  const bool want_stats = out != nullptr;
  const bool count_trav = (rp->flags & PT_FLAG_COUNT_TRAVERSAL) != 0;
  // sorting pays only when more than one shading program can meet in a wave
  const bool sort = (tuned ? tu.sort : (rp->flags & PT_FLAG_MATERIAL_SORT) != 0) && c->shading_kinds > 1;
  // whole batches replayed from one captured graph (no per-kernel timing)
  const bool use_graph =
      (tuned ? tu.graph : (rp->flags & PT_FLAG_GRAPH) != 0) && !use_cache && !(rp->flags & PT_FLAG_DIRECT);
  const bool kernel_events = want_stats && !use_graph;
  std::vector<hipEvent_t> ev;
  auto mark = [&]() -> hipEvent_t {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, c->stream);
    ev.push_back(e);
    return e;
  };
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ext_ev, sh_ev, train_ev, infer_ev;
  uint64_t train_samples = 0, infer_samples = 0;
  launch_fill_u32((uint32_t*)c->stats, (uint64_t)MAX_WAVES * STAT_WORDS * 2, 0u, c->stream);
  hipEvent_t t_begin = want_stats ? mark() : nullptr;
  DevPaths& P = c->paths;
  const int matset = shade_material_set(c->mat_types);
  // one batch: camera rays, the bounce loop, the paths that outlived it
  auto enqueue_batch = [&](uint32_t nslots, int grid, bool events) {
    int in = 0, out = 1;
    for (uint32_t depth = 0; depth < rp->max_depth; depth++) {
      hipEvent_t a = events ? mark() : nullptr;
      // depth 0 starts the camera paths in the same launch
      // material sort from the first bounce on: a wave's camera hits are
      // runs of adjacent pixels, already coherent (sorting them as well
      // measured 169 against 163 ms per mesh-scene step)
      const bool sd = sort && depth > 0;
      launch_extend(c->dsc, P, in, depth == 0 ? 1e-6f : 1e-3f, c->spill, c->stats, count_trav, sd, grid, c->stream,
                    nullptr, depth == 0 ? &rc : nullptr, depth == 0 ? nslots : 0u);
      hipEvent_t b = events ? mark() : nullptr;
      if (events) ext_ev.push_back({a, b});
      if (sd) launch_sort(P, in, grid, c->stream);
      hipEvent_t a2 = events ? mark() : nullptr;
      launch_shade(c->dsc, P, in, sd, out, depth, rc, matset, count_trav, c->spill, c->stats, grid, c->stream);
      hipEvent_t b2 = events ? mark() : nullptr;
      if (events) sh_ev.push_back({a2, b2});
      std::swap(in, out);
    }
    launch_exhausted(P, in, c->stats, grid, c->stream);
  };
  if (use_graph) {
    if (!c->frame_dev) HIPCHK(c, hipMalloc(&c->frame_dev, 64));
    launch_set_frame(c->frame_dev, frame_begin, c->stream);
  }
  uint32_t fallbacks = 0, plain_batches = 0;
  for (uint32_t f0 = 0; f0 < frame_count; f0 += fpb) {
    uint32_t nf = std::min(fpb, frame_count - f0);
    uint32_t nslots = nf * WH;
    rc.frame0 = frame_begin + f0;
    // a held batch asks nothing (no path reaches this query depth)
    if (use_cache) rc.cache.query_depth = c->cache_hold ? 0xFFFFFFFFu : c->cache_query_depth;
    const int grid = grid_for(nslots);
    if (use_graph && nf == fpb) {
      // full batches replay a graph captured once for this launch state;
      // the batch's first frame comes from frame_dev, advanced in-graph
      RenderConsts grc = rc;
      grc.frame0 = 0;
      grc.frame_dev = c->frame_dev;
      std::vector<uint8_t> key;
      auto add = [&](const void* p, size_t n) { key.insert(key.end(), (const uint8_t*)p, (const uint8_t*)p + n); };
      add(&grc, sizeof(grc));
      add(&c->dsc, sizeof(c->dsc));
      add(&P, sizeof(P));
      uint32_t shape[6] = {nslots, (uint32_t)grid, rp->max_depth, (uint32_t)sort, (uint32_t)matset, (uint32_t)count_trav};
      add(shape, sizeof(shape));
      add(&c->accum, sizeof(c->accum));
      add(&c->stats, sizeof(c->stats));
      if (!c->graph_exec || key != c->graph_key) {
        if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
        if (c->graph) (void)hipGraphDestroy(c->graph);
        c->graph_exec = nullptr;
        c->graph = nullptr;
        c->graph_key.clear();
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        RenderConsts saved = rc;
        rc = grc;
        enqueue_batch(nslots, grid, false);
        launch_accumulate(P, rc, nf, c->accum, c->stats, grid_for(WH), c->stream);
        launch_next_frames(c->frame_dev, nf, c->stream);
        rc = saved;
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamEndCapture(c->stream, &g));
        c->graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
        c->graph_key = key;
      }
      HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
      c->last_fpb = nf;
      continue;
    }
    if (use_graph) launch_set_frame(c->frame_dev, frame_begin + f0 + nf, c->stream);
    enqueue_batch(nslots, grid, kernel_events);
    if (use_cache) {
      // train on this batch's training vertices (an evenly spread dense
      // copy of steps x batch of them, with targets), then answer its queries
      launch_cache_scan(P, grid, c->stream);
      uint32_t cnt[2] = {0, 0};  // CTR_QUERY, CTR_TRAIN
      HIPCHK(c, hipMemcpyAsync(cnt, P.counters + CTR_QUERY, 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      const uint32_t nq = cnt[0], nt = cnt[1];
      const uint32_t rows = (uint32_t)std::min<uint64_t>((uint64_t)c->cache_steps * c->cache_batch, nt);
      uint32_t steps = 0;
      if (rows >= 32) {
        launch_cache_gather(P, 0, rows, golden_stride(rows), grid, (uint4*)c->cfeat, c->ctgt, nullptr, nullptr,
                            c->stream);
        hipEvent_t ta = want_stats ? mark() : nullptr;
        // steps x batch rows when there are that many; fewer rows are
        // cycled through, so every batch takes its cache_steps steps
        for (uint32_t k = 0; k < c->cache_steps; k++) {
          const uint32_t off = (uint32_t)(((uint64_t)k * c->cache_batch) % rows);
          if (rows - off < 32) continue;
          const uint32_t n = std::min(c->cache_batch, rows - off);
          c->cache_step++;
          launch_mlp_train(c->cfeat + (size_t)off * MLP_IN, c->ctgt + (size_t)off * 3, n, c->cfw, c->cbw, c->cbias,
                           c->cX, c->cdZ, c->cdb, c->closs, c->cdw, c->cp, c->cm1, c->cm2, c->cgrad, c->cache_lr,
                           c->cache_step, c->cfw, c->cbw, c->cbias, c->stream);
          steps++;
          train_samples += n;
          c->cache_last_batch = n;
        }
        if (want_stats) train_ev.push_back({ta, mark()});
      }
      // the judge: training vertices halfway between the ones trained on
      // (the same ones when there are few), finished by the plain tracer,
      // against the network that is about to answer the queries (or, for a
      // held batch, that would answer the next batch's)
      bool use_answers = true;
      const bool held = c->cache_hold;
      if (held) plain_batches++;
      const uint32_t nj = std::min(nt, c->cache_judge_rows);
      if (nj >= 32) {
        launch_cache_gather(P, 0, nj, golden_stride(nj), grid, c->cjfeat, nullptr, c->cjvals, nullptr, c->stream,
                            std::max<uint32_t>(rows, 1u));
        hipEvent_t ia = want_stats ? mark() : nullptr;
        launch_mlp_infer((const uint16_t*)c->cjfeat, nj, c->cfw, c->cbias, c->cjout,
                         (int)std::min<uint32_t>(1024, (nj + 63) / 64), c->stream);
        if (want_stats) infer_ev.push_back({ia, mark()});
        infer_samples += nj;
        launch_cache_judge(c->cjvals, c->cjout, nj, c->cjres, c->stream);
        double jr[3] = {0.0, 0.0, 0.0};
        HIPCHK(c, hipMemcpyAsync(jr, c->cjres, sizeof(jr), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double ratio = jr[1] > 0.0 ? jr[0] / jr[1] : (jr[0] > 0.0 ? INFINITY : 0.0);
        const bool finite = jr[2] == 0.0 && std::isfinite(jr[0]) && std::isfinite(jr[1]);
        c->cache_last_judge = finite ? ratio : NAN;
        use_answers = finite && ratio < (double)c->cache_judge_max;
        c->cache_hold = !use_answers;
        // a network whose outputs are no longer finite starts again
        if (!finite && (ps = cache_reset(c)) != PT_OK) return ps;
      }
      if (nq && use_answers) {
        hipEvent_t ia = want_stats ? mark() : nullptr;
        launch_mlp_infer((const uint16_t*)P.q_feat, nslots, c->cfw, c->cbias, P.q_out, 2048, c->stream, P.qseg,
                         (uint32_t)grid * BLOCK);
        if (want_stats) infer_ev.push_back({ia, mark()});
        infer_samples += nq;
        launch_cache_apply(P, grid, c->stream);
      } else if (nq) {
        // fallback: the batch again with the plain tracer, whose results
        // replace the cut paths' (the same samples: seeds depend on pixel
        // and frame only); the batches after it render plain until the
        // network passes the judge again
        RenderConsts saved = rc;
        rc.flags &= ~(uint32_t)PT_FLAG_RADIANCE_CACHE;
        enqueue_batch(nslots, grid, kernel_events);
        rc = saved;
        fallbacks++;
      }
      c->cache_last_train = nt;
      c->cache_last_query = nq;
      c->cache_last_steps = steps;
      c->cache_last_grid = grid;
    }
    launch_accumulate(P, rc, nf, c->accum, c->stats, grid_for(WH), c->stream);
    c->last_fpb = nf;
  }
  if (use_cache) {
    c->cache_last_fallbacks = fallbacks;
    c->cache_last_plain_batches = plain_batches;
  }
  hipEvent_t t_end = want_stats ? mark() : nullptr;
  if (want_stats) launch_stats_reduce(c->stats, c->stats_total, c->stream);
  c->frames = frame_begin + frame_count;
  c->paths_valid = true;
  if (want_stats) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    uint64_t st[STAT_WORDS];
    uint32_t ctr[CTR_WORDS];
    HIPCHK(c, hipMemcpy(st, c->stats_total, sizeof(st), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ctr, P.counters, sizeof(ctr), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->paths = (uint64_t)frame_count * WH;
    out->extension_rays = st[STAT_EXT];
    out->shadow_rays = st[STAT_SHADOW];
    out->node_fetches = st[STAT_NODES];
    out->prim_fetches = st[STAT_PRIMS];
    out->instance_fetches = st[STAT_INSTS];
    out->shadow_node_fetches = st[STAT_SH_NODES];
    out->shadow_prim_fetches = st[STAT_SH_PRIMS];
    out->shadow_instance_fetches = st[STAT_SH_INSTS];
    out->exhausted_paths = st[STAT_EXHAUSTED];
    out->nan_paths = st[STAT_NAN];
    out->cache_queries = c->cache_last_query;
    out->cache_train_vertices = c->cache_last_train;
    out->cache_steps = c->cache_last_steps;
    out->cache_loss = 0.0;
    out->cache_judge = use_cache ? c->cache_last_judge : -1.0;
    out->cache_fallbacks = use_cache ? c->cache_last_fallbacks : 0u;
    out->cache_plain_batches = use_cache ? c->cache_last_plain_batches : 0u;
    if (use_cache && c->cache_last_steps) {
      const uint32_t n = c->cache_last_batch;
      const uint32_t waves = mlp_fwd_bwd_waves(n);
      std::vector<float> lp(waves);
      HIPCHK(c, hipMemcpy(lp.data(), c->closs, waves * 4, hipMemcpyDeviceToHost));
      double sum = 0.0;
      for (float v : lp) sum += v;
      out->cache_loss = sum;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t_begin, t_end);
    out->gpu_ms = ms;
    for (auto& p : ext_ev) {
      float m = 0.f;
      (void)hipEventElapsedTime(&m, p.first, p.second);
      out->extend_ms += m;
    }
    for (auto& p : sh_ev) {
      float m = 0.f;
      (void)hipEventElapsedTime(&m, p.first, p.second);
      out->shade_ms += m;
    }
    for (auto& p : train_ev) {
      float m = 0.f;
      (void)hipEventElapsedTime(&m, p.first, p.second);
      out->cache_train_ms += m;
    }
    for (auto& p : infer_ev) {
      float m = 0.f;
      (void)hipEventElapsedTime(&m, p.first, p.second);
      out->cache_infer_ms += m;
    }
    out->cache_train_samples = train_samples;
    out->cache_infer_samples = infer_samples;
    out->extend_launches = (uint32_t)ext_ev.size();
    out->shade_launches = (uint32_t)sh_ev.size();
    for (auto e : ev) (void)hipEventDestroy(e);
    if (ctr[CTR_STACK_OVERFLOW]) return fail(c, PT_ERR_DEVICE, "traversal stack overflow");
  }
  return PT_OK;
}

pt_status pt_sync(pt_context* c) {
  if (!c) return PT_ERR_INVALID_ARG;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PT_OK;
}

pt_status pt_read_accum(pt_context* c, double* rgb, uint32_t* frames) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!c->accum) return fail(c, PT_ERR_STATE, "nothing rendered");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (rgb) HIPCHK(c, hipMemcpy(rgb, c->accum, (size_t)c->W * c->H * 3 * sizeof(double), hipMemcpyDeviceToHost));
  if (frames) *frames = c->frames;
  return PT_OK;
}

// image_save.hpp:7-73 (Save_PNG / Save_HDR) on the accumulation average:
// PNG: 255.99 * clamp(sqrt(c), 0, 1) truncated to 8 bits; HDR: c / (1 + c)
// as flat RGBE; rows top first.
pt_status pt_save_image(pt_context* c, const char* path, int32_t format) {
  if (!c || !path || (format != PT_IMAGE_PNG && format != PT_IMAGE_HDR)) return PT_ERR_INVALID_ARG;
  if (!c->accum || c->frames == 0) return fail(c, PT_ERR_STATE, "nothing rendered");
  const size_t n = (size_t)c->W * c->H;
  std::vector<double> acc(n * 3);
  uint32_t frames = 0;
  pt_status ps = pt_read_accum(c, acc.data(), &frames);
  if (ps != PT_OK) return ps;
  const float fr = (float)frames;
  bool ok;
  if (format == PT_IMAGE_PNG) {
    std::vector<unsigned char> rgb(n * 3);
    for (size_t i = 0; i < n * 3; i++) {
      const float v = std::sqrt((float)acc[i] / fr);
      rgb[i] = (unsigned char)(int)(255.99f * std::fmin(std::fmax(v, 0.f), 1.f));
    }
    ok = write_png_rgb(path, (int)c->W, (int)c->H, rgb.data()) != 0;
  } else {
    std::vector<float> rgb(n * 3);
    for (size_t i = 0; i < n * 3; i++) {
      const float v = (float)acc[i] / fr;
      rgb[i] = v / (1.f + v);
    }
    ok = rgbe_save(path, c->W, c->H, rgb.data());
  }
  if (!ok) return fail(c, PT_ERR_INVALID_ARG, std::string("cannot write ") + path);
  return PT_OK;
}

pt_status pt_read_display(pt_context* c, uint8_t* rgba) {
  if (!c || !rgba) return PT_ERR_INVALID_ARG;
  if (!c->accum || c->frames == 0) return fail(c, PT_ERR_STATE, "nothing rendered");
  (void)hipSetDevice(c->device);
  uint32_t n = c->W * c->H;
  launch_display(c->accum, n, c->frames, c->rgba, grid_for(n), c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(rgba, c->rgba, (size_t)n * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

// ------------------------------------------------------- radiance cache
static pt_status cache_work(pt_context* c, uint32_t n) {
  if (n <= c->cwork_cap) return PT_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(c->cwork_bufs);
  c->cwork_cap = 0;
  const uint32_t cap = std::max<uint32_t>(n, 1024);
  const size_t ld = mlp_padded(cap), waves = mlp_fwd_bwd_waves(cap), chunks = mlp_chunks(cap);
  auto& B = c->cwork_bufs;
  hipError_t e;
  if ((e = dalloc(B, (void**)&c->cfeat, (size_t)cap * MLP_IN * 2)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->ctgt, (size_t)cap * MLP_OUT * 4)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cX, (size_t)MLP_X_ROWS * ld * 2)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cdZ, (size_t)MLP_DZ_ROWS * ld * 2)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cdb, (waves + 1) * MLP_DB_WORDS * 4)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->closs, waves * 4)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cdw, chunks * MLP_PARAMS * 4)) != hipSuccess) {
    dfree(B);
    return hip_fail(c, e, "hipMalloc (cache work buffers)");
  }
  c->cwork_cap = cap;
  return PT_OK;
}

pt_status pt_cache_init(pt_context* c, uint32_t seed) {
  if (!c) return PT_ERR_INVALID_ARG;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!c->cp) {
    auto& B = c->cache_bufs;
    hipError_t e;
    if ((e = dalloc(B, (void**)&c->cp, MLP_PARAMS * 4)) != hipSuccess ||
        (e = dalloc(B, (void**)&c->cm1, MLP_PARAMS * 4)) != hipSuccess ||
        (e = dalloc(B, (void**)&c->cm2, MLP_PARAMS * 4)) != hipSuccess ||
        (e = dalloc(B, (void**)&c->cgrad, MLP_PARAMS * 4)) != hipSuccess ||
        (e = dalloc(B, (void**)&c->cfw, MLP_FW_UINT4 * 16)) != hipSuccess ||
        (e = dalloc(B, (void**)&c->cbw, MLP_BW_UINT4 * 16)) != hipSuccess ||
        (e = dalloc(B, (void**)&c->cbias, MLP_BIAS_WORDS * 4)) != hipSuccess) {
      dfree(B);
      return hip_fail(c, e, "hipMalloc (cache parameters)");
    }
  }
  c->cache_seed = seed;
  pt_status ps = cache_reset(c);
  if (ps != PT_OK) return ps;
  c->cache_ready = true;
  return PT_OK;
}

// He-uniform weights U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)) from
// mt19937(cache_seed) (24-bit uniforms), zero biases, zero Adam state
static pt_status cache_reset(pt_context* c) {
  const uint32_t seed = c->cache_seed;
  std::vector<float> p(MLP_PARAMS, 0.f);
  std::mt19937 g(seed);
  const int in[5] = {MLP_IN, MLP_H, MLP_H, MLP_H, MLP_H}, out[5] = {MLP_H, MLP_H, MLP_H, MLP_H, MLP_OUT};
  size_t off = 0;
  for (int L = 0; L < 5; L++) {
    const float a = std::sqrt(6.f / (float)in[L]);
    for (int k = 0; k < in[L] * out[L]; k++) {
      float u = (float)(g() >> 8) * (1.f / 16777216.f);
      p[off + k] = (2.f * u - 1.f) * a;
    }
    off += (size_t)in[L] * out[L] + out[L];
  }
  HIPCHK(c, hipMemcpy(c->cp, p.data(), MLP_PARAMS * 4, hipMemcpyHostToDevice));
  launch_fill_u32((uint32_t*)c->cm1, MLP_PARAMS, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->cm2, MLP_PARAMS, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->cgrad, MLP_PARAMS, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->cfw, MLP_FW_UINT4 * 4, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->cbw, MLP_BW_UINT4 * 4, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->cbias, MLP_BIAS_WORDS, 0u, c->stream);
  launch_mlp_pack(c->cp, c->cfw, c->cbw, c->cbias, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cache_step = 0;
  c->cache_hold = true;
  return PT_OK;
}

static pt_status ensure_cache_judge(pt_context* c) {
  const uint32_t n = c->cache_judge_rows;
  if (n <= c->cjudge_cap && c->cjres) return PT_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(c->cjudge_bufs);
  c->cjudge_cap = 0;
  auto& B = c->cjudge_bufs;
  hipError_t e;
  if ((e = dalloc(B, (void**)&c->cjfeat, (size_t)n * 12 * 16)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cjvals, (size_t)n * 9 * 4)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cjout, (size_t)n * 3 * 4)) != hipSuccess ||
      (e = dalloc(B, (void**)&c->cjres, 4 * sizeof(double))) != hipSuccess) {
    dfree(B);
    return hip_fail(c, e, "hipMalloc (cache judge)");
  }
  c->cjudge_cap = n;
  return PT_OK;
}

pt_status pt_cache_params(pt_context* c, float* out, const float* in) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!c->cache_ready) return fail(c, PT_ERR_STATE, "pt_cache_init first");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) HIPCHK(c, hipMemcpy(out, c->cp, MLP_PARAMS * 4, hipMemcpyDeviceToHost));
  if (in) {
    HIPCHK(c, hipMemcpy(c->cp, in, MLP_PARAMS * 4, hipMemcpyHostToDevice));
    launch_mlp_pack(c->cp, c->cfw, c->cbw, c->cbias, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return PT_OK;
}

pt_status pt_cache_infer(pt_context* c, const uint16_t* features, uint64_t n, float* out) {
  if (!c || (n && (!features || !out))) return PT_ERR_INVALID_ARG;
  if (!c->cache_ready) return fail(c, PT_ERR_STATE, "pt_cache_init first");
  if (n == 0) return PT_OK;
  if (n > (1u << 26)) return fail(c, PT_ERR_INVALID_ARG, "too many samples");
  (void)hipSetDevice(c->device);
  pt_status ps = cache_work(c, (uint32_t)n);
  if (ps != PT_OK) return ps;
  HIPCHK(c, hipMemcpy(c->cfeat, features, n * MLP_IN * 2, hipMemcpyHostToDevice));
  const int grid = (int)std::min<uint64_t>(1024, (n + 63) / 64);
  launch_mlp_infer(c->cfeat, (uint32_t)n, c->cfw, c->cbias, c->ctgt, grid, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(out, c->ctgt, n * MLP_OUT * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

pt_status pt_cache_train(pt_context* c, const uint16_t* features, const float* targets, uint64_t n, float lr,
                         float* loss, float* grad) {
  if (!c || !n || !features || !targets) return PT_ERR_INVALID_ARG;
  if (!c->cache_ready) return fail(c, PT_ERR_STATE, "pt_cache_init first");
  if (n > (1u << 22)) return fail(c, PT_ERR_INVALID_ARG, "training batch larger than 2^22 samples");
  (void)hipSetDevice(c->device);
  pt_status ps = cache_work(c, (uint32_t)n);
  if (ps != PT_OK) return ps;
  HIPCHK(c, hipMemcpy(c->cfeat, features, n * MLP_IN * 2, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->ctgt, targets, n * MLP_OUT * 4, hipMemcpyHostToDevice));
  c->cache_step++;
  launch_mlp_train(c->cfeat, c->ctgt, (uint32_t)n, c->cfw, c->cbw, c->cbias, c->cX, c->cdZ, c->cdb, c->closs, c->cdw,
                   c->cp, c->cm1, c->cm2, c->cgrad, lr, c->cache_step, c->cfw, c->cbw, c->cbias, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) {
    const uint32_t waves = mlp_fwd_bwd_waves((uint32_t)n);
    std::vector<float> lp(waves);
    HIPCHK(c, hipMemcpy(lp.data(), c->closs, waves * 4, hipMemcpyDeviceToHost));
    double sum = 0.0;
    for (float v : lp) sum += v;
    *loss = (float)sum;
  }
  if (grad) HIPCHK(c, hipMemcpy(grad, c->cgrad, MLP_PARAMS * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

// A batch of scenes in one wavefront: the frames of every context form one
// work list of (scene, frame) items, cut into batches of about the
// context's paths_target
// slots; each batch's paths share one path pool and one launch per kernel
// and bounce, each path finding its scene through the index kept in its RNG
// state.  The results of every context are those of its own pt_render.
}  // extern "C"

namespace {

// launch shape of a batch of scenes
struct BatchShape {
  uint64_t paths = PATHS_MAX;     // slots per sub-batch (a scene's frame is never split)
  bool sort = true;               // material sort (when the batch has more than one shading program)
  bool graph = false;             // the whole batch replayed from one captured graph
};

// the batch's composition: what a tuned shape is valid for (contexts,
// resolutions, depths and flags other than the shape's own)
std::vector<uint8_t> batch_composition(pt_context* const* ctxs, const pt_render_params* params, uint32_t n) {
  std::vector<uint8_t> k;
  auto add = [&](const void* p, size_t b) { k.insert(k.end(), (const uint8_t*)p, (const uint8_t*)p + b); };
  add(&n, sizeof(n));
  for (uint32_t i = 0; i < n; i++) {
    add(&ctxs[i], sizeof(ctxs[i]));
    pt_render_params q = params[i];
    q.flags &= ~(uint32_t)(PT_FLAG_TUNED | PT_FLAG_GRAPH | PT_FLAG_MATERIAL_SORT | PT_FLAG_DIRECT |
                           PT_FLAG_COUNT_TRAVERSAL);
    q.frames_per_batch = 0;
    add(&q, sizeof(q));
  }
  return k;
}

pt_status render_batch_shaped(pt_context* const* ctxs, const pt_render_params* params, const uint32_t* frame_counts,
                              uint32_t n, pt_render_stats* out, const BatchShape& shape) {
  pt_context* B = ctxs[0];
  uint32_t union_types = 0, max_depth = 0;
  bool env = false, ext = false;  // environment MIS; media, cylinders, moving spheres or the normal shader
  for (uint32_t i = 0; i < n; i++) {
    pt_context* c = ctxs[i];
    const pt_render_params* rp = params + i;
    if (!c) return fail(B, PT_ERR_INVALID_ARG, "null context in the batch");
    for (uint32_t j = 0; j < i; j++)
      if (ctxs[j] == c) return fail(B, PT_ERR_INVALID_ARG, "a context appears twice in the batch");
    if (c->device != B->device) return fail(B, PT_ERR_INVALID_ARG, "the batch's contexts are on different devices");
    if (!c->accel_ready || !c->has_camera) return fail(B, PT_ERR_STATE, "a batch context has no built scene or camera");
    if (rp->width == 0 || rp->height == 0 || (uint64_t)rp->width * rp->height > (1ull << 28) || rp->max_depth == 0)
      return fail(B, PT_ERR_INVALID_ARG, "invalid render parameters in the batch");
    if ((uint64_t)rp->width * rp->height > SLOT_MASK)
      return fail(B, PT_ERR_INVALID_ARG, "image too large for one batch");
    if (rp->flags & PT_FLAG_RADIANCE_CACHE) return fail(B, PT_ERR_UNSUPPORTED, "no radiance cache in a batch");
    ext = ext || c->dsc.ext_instances != 0 || (c->mat_types >> PT_MAT_ISOTROPIC) != 0;
    union_types |= c->mat_types;
    max_depth = std::max(max_depth, rp->max_depth);
    env = env || ((rp->flags & PT_FLAG_ENV_MIS) && c->dsc.num_lights != c->dsc.num_lights_plain);
  }
  (void)hipSetDevice(B->device);
  for (uint32_t i = 0; i < n; i++) HIPCHK(B, hipStreamSynchronize(ctxs[i]->stream));
  // accumulation buffers (cleared on the batch's stream, below)
  for (uint32_t i = 0; i < n; i++) {
    pt_context* c = ctxs[i];
    const pt_render_params* rp = params + i;
    const uint32_t WH = rp->width * rp->height;
    if (rp->width != c->W || rp->height != c->H || !c->accum) {
      if (c->accum) (void)hipFree(c->accum);
      if (c->rgba) (void)hipFree(c->rgba);
      c->accum = nullptr;
      c->rgba = nullptr;
      hipError_t e;
      if ((e = hipMalloc(&c->accum, (size_t)WH * 3 * sizeof(double))) != hipSuccess ||
          (e = hipMalloc(&c->rgba, (size_t)WH * sizeof(uchar4))) != hipSuccess)
        return hip_fail(B, e, "hipMalloc (accumulation)");
      c->W = rp->width;
      c->H = rp->height;
    }
    c->frames = 0;
    c->paths_valid = false;
  }
  // work items (scene, frame) in scene-then-frame order, cut into batches
  // of at most shape.paths slots (or one frame, if larger)
  std::vector<uint32_t> off, scene, frame;
  std::vector<size_t> first;      // first item of each batch
  std::vector<uint32_t> bslots;   // slots of each batch
  uint64_t cur = 0;
  const uint64_t cap = std::min<uint64_t>(shape.paths, SLOT_MASK);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t WH = params[i].width * params[i].height;
    for (uint32_t f = 0; f < frame_counts[i]; f++) {
      if (first.empty() || (cur > 0 && cur + WH > std::max<uint64_t>(cap, WH))) {
        if (!first.empty()) bslots.push_back((uint32_t)cur);
        first.push_back(off.size());
        cur = 0;
      }
      off.push_back((uint32_t)cur);
      scene.push_back(i);
      frame.push_back(f);
      cur += WH;
    }
  }
  if (!first.empty()) bslots.push_back((uint32_t)cur);
  first.push_back(off.size());
  const size_t nitems = off.size();
  uint32_t max_slots = 0;
  for (uint32_t b : bslots) max_slots = std::max(max_slots, b);
  // tables
  if (n > B->batch_scene_cap || nitems > B->batch_item_cap) {
    dfree(B->batch_bufs);
    B->batch_scene_cap = B->batch_item_cap = 0;
    const size_t sc = std::max<size_t>(n, 64), ic = std::max<size_t>(nitems, 1024);
    hipError_t e;
    if ((e = dalloc(B->batch_bufs, (void**)&B->d_scenes, sc * sizeof(DevScene))) != hipSuccess ||
        (e = dalloc(B->batch_bufs, (void**)&B->d_rcs, sc * sizeof(RenderConsts))) != hipSuccess ||
        (e = dalloc(B->batch_bufs, (void**)&B->d_item_off, ic * 4)) != hipSuccess ||
        (e = dalloc(B->batch_bufs, (void**)&B->d_item_scene, ic * 4)) != hipSuccess ||
        (e = dalloc(B->batch_bufs, (void**)&B->d_item_frame, ic * 4)) != hipSuccess) {
      dfree(B->batch_bufs);
      return hip_fail(B, e, "hipMalloc (batch tables)");
    }
    B->batch_scene_cap = sc;
    B->batch_item_cap = ic;
  }
  std::vector<DevScene> scenes(n);
  std::vector<RenderConsts> rcs(n);
  for (uint32_t i = 0; i < n; i++) {
    scenes[i] = ctxs[i]->dsc;
    rcs[i] = render_consts(ctxs[i], params + i);
  }
  HIPCHK(B, hipMemcpy(B->d_scenes, scenes.data(), n * sizeof(DevScene), hipMemcpyHostToDevice));
  HIPCHK(B, hipMemcpy(B->d_rcs, rcs.data(), n * sizeof(RenderConsts), hipMemcpyHostToDevice));
  if (nitems) {
    HIPCHK(B, hipMemcpy(B->d_item_off, off.data(), nitems * 4, hipMemcpyHostToDevice));
    HIPCHK(B, hipMemcpy(B->d_item_scene, scene.data(), nitems * 4, hipMemcpyHostToDevice));
    HIPCHK(B, hipMemcpy(B->d_item_frame, frame.data(), nitems * 4, hipMemcpyHostToDevice));
  }
  if (max_slots) {
    pt_status ps = ensure_paths(B, max_slots);
    if (ps == PT_OK) ps = ensure_spill(B);
    if (ps != PT_OK) return ps;
  }
  const bool want_stats = out != nullptr;
  const bool count_trav = (params[0].flags & PT_FLAG_COUNT_TRAVERSAL) != 0;
  const bool sort = shape.sort && __builtin_popcount(union_types & ~(1u << PT_MAT_DIFFUSE_LIGHT)) > 1;
  const int matset = shade_material_set(union_types);
  DevPaths& P = B->paths;
  launch_fill_u32((uint32_t*)B->stats, (uint64_t)MAX_WAVES * STAT_WORDS * 2, 0u, B->stream);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (want_stats) {
    HIPCHK(B, hipEventCreate(&t0));
    HIPCHK(B, hipEventCreate(&t1));
    HIPCHK(B, hipEventRecord(t0, B->stream));
  }
  RenderConsts rc0 = rcs[0];
  rc0.flags &= ~(uint32_t)PT_FLAG_RADIANCE_CACHE;
  // per-kernel HIP events around the closest-hit and shading launches of
  // direct (not replayed) batches with statistics
  const bool kernel_events = want_stats && !shape.graph;
  std::vector<hipEvent_t> kev;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ext_ev, sh_ev;
  auto mark = [&]() -> hipEvent_t {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, B->stream);
    kev.push_back(e);
    return e;
  };
  // the whole batch: clear the accumulations, then per sub-batch the bounce
  // loop over every live path (its depth-0 launch starts the paths), the
  // paths that outlived it, and each scene's frames of the sub-batch
  // accumulated in frame order
  auto enqueue = [&]() {
    for (uint32_t i = 0; i < n; i++) launch_fill_f64(ctxs[i]->accum, (uint64_t)rcs[i].WH * 3, 0.0, B->stream);
    for (size_t b = 0; b + 1 < first.size(); b++) {
      const uint32_t nslots = bslots[b];
      const size_t k0 = first[b], k1 = first[b + 1];
      if (nslots == 0) continue;
      const int grid = grid_for(nslots);
      // the depth-0 closest-hit launch starts the sub-batch's camera paths
      const BatchStart bs{B->d_rcs, B->d_item_off + k0, B->d_item_scene + k0, B->d_item_frame + k0,
                          (uint32_t)(k1 - k0)};
      int in = 0, o = 1;
      for (uint32_t depth = 0; depth < max_depth; depth++) {
        const bool sd = sort && depth > 0;  // from the first bounce on, as in pt_render
        hipEvent_t a = kernel_events ? mark() : nullptr;
        launch_extend(B->dsc, P, in, depth == 0 ? 1e-6f : 1e-3f, B->spill, B->stats, count_trav, sd, grid,
                      B->stream, B->d_scenes, nullptr, depth == 0 ? nslots : 0u, ext, depth == 0 ? &bs : nullptr);
        if (kernel_events) ext_ev.push_back({a, mark()});
        if (sd) launch_sort(P, in, grid, B->stream);
        hipEvent_t a2 = kernel_events ? mark() : nullptr;
        launch_shade(B->dsc, P, in, sd, o, depth, rc0, matset, count_trav, B->spill, B->stats, grid, B->stream,
                     B->d_scenes, B->d_rcs, env, ext);
        if (kernel_events) sh_ev.push_back({a2, mark()});
        std::swap(in, o);
      }
      launch_exhausted(P, in, B->stats, grid, B->stream);
      for (size_t k = k0; k < k1;) {
        size_t e = k;
        while (e < k1 && scene[e] == scene[k]) e++;
        DevPaths Pk = P;
        Pk.result = P.result + off[k];
        const uint32_t i = scene[k];
        launch_accumulate(Pk, rcs[i], (uint32_t)(e - k), ctxs[i]->accum, B->stats, grid_for(rcs[i].WH), B->stream);
        k = e;
      }
    }
  };
  if (shape.graph) {
    // replay when everything the captured launches take by value or by
    // address is unchanged: contexts, parameters, frame counts, the scene
    // and constant tables, the path pool, accumulations and the shape
    std::vector<uint8_t> key;
    auto add = [&](const void* p, size_t b) { key.insert(key.end(), (const uint8_t*)p, (const uint8_t*)p + b); };
    add(&n, sizeof(n));
    add(ctxs, n * sizeof(ctxs[0]));
    add(params, n * sizeof(params[0]));
    add(frame_counts, n * sizeof(frame_counts[0]));
    add(scenes.data(), n * sizeof(DevScene));
    add(rcs.data(), n * sizeof(RenderConsts));
    for (uint32_t i = 0; i < n; i++) add(&ctxs[i]->accum, sizeof(double*));
    add(&B->dsc, sizeof(B->dsc));
    add(&P, sizeof(P));
    const void* ptrs[7] = {B->spill, B->stats, B->d_scenes, B->d_rcs, B->d_item_off, B->d_item_scene, B->d_item_frame};
    add(ptrs, sizeof(ptrs));
    const uint64_t sh[4] = {cap, (uint64_t)sort, (uint64_t)count_trav, (uint64_t)max_depth};
    add(sh, sizeof(sh));
    if (!B->batch_graph_exec || key != B->batch_graph_key) {
      if (B->batch_graph_exec) (void)hipGraphExecDestroy(B->batch_graph_exec);
      if (B->batch_graph) (void)hipGraphDestroy(B->batch_graph);
      B->batch_graph_exec = nullptr;
      B->batch_graph = nullptr;
      B->batch_graph_key.clear();
      HIPCHK(B, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      enqueue();
      hipGraph_t g = nullptr;
      HIPCHK(B, hipStreamEndCapture(B->stream, &g));
      B->batch_graph = g;
      HIPCHK(B, hipGraphInstantiate(&B->batch_graph_exec, g, nullptr, nullptr, 0));
      B->batch_graph_key = key;
    }
    HIPCHK(B, hipGraphLaunch(B->batch_graph_exec, B->stream));
  } else {
    enqueue();
  }
  for (uint32_t i = 0; i < n; i++) ctxs[i]->frames = frame_counts[i];
  // later work on any batch context (reads, renders) comes after the batch
  hipEvent_t done = nullptr;
  HIPCHK(B, hipEventCreateWithFlags(&done, hipEventDisableTiming));
  HIPCHK(B, hipEventRecord(done, B->stream));
  for (uint32_t i = 1; i < n; i++) HIPCHK(B, hipStreamWaitEvent(ctxs[i]->stream, done, 0));
  (void)hipEventDestroy(done);
  if (want_stats) {
    HIPCHK(B, hipEventRecord(t1, B->stream));
    launch_stats_reduce(B->stats, B->stats_total, B->stream);
    HIPCHK(B, hipStreamSynchronize(B->stream));
    uint64_t st[STAT_WORDS];
    HIPCHK(B, hipMemcpy(st, B->stats_total, sizeof(st), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    for (uint32_t i = 0; i < n; i++) out->paths += (uint64_t)frame_counts[i] * params[i].width * params[i].height;
    out->extension_rays = st[STAT_EXT];
    out->shadow_rays = st[STAT_SHADOW];
    out->node_fetches = st[STAT_NODES];
    out->prim_fetches = st[STAT_PRIMS];
    out->instance_fetches = st[STAT_INSTS];
    out->shadow_node_fetches = st[STAT_SH_NODES];
    out->shadow_prim_fetches = st[STAT_SH_PRIMS];
    out->shadow_instance_fetches = st[STAT_SH_INSTS];
    out->exhausted_paths = st[STAT_EXHAUSTED];
    out->nan_paths = st[STAT_NAN];
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    out->gpu_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    for (auto& e : ext_ev) {
      float m = 0.f;
      if (e.first && e.second && hipEventElapsedTime(&m, e.first, e.second) == hipSuccess) out->extend_ms += m;
    }
    for (auto& e : sh_ev) {
      float m = 0.f;
      if (e.first && e.second && hipEventElapsedTime(&m, e.first, e.second) == hipSuccess) out->shade_ms += m;
    }
    out->extend_launches = (uint32_t)ext_ev.size();
    out->shade_launches = (uint32_t)sh_ev.size();
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
    uint32_t ovf = 0;
    HIPCHK(B, hipMemcpy(&ovf, P.counters + CTR_STACK_OVERFLOW, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(B, PT_ERR_DEVICE, "traversal stack overflow");
  }
  return PT_OK;
}

}  // namespace

extern "C" {

pt_status pt_render_batch(pt_context* const* ctxs, const pt_render_params* params, const uint32_t* frame_counts,
                          uint32_t n, pt_render_stats* out) {
  if (!ctxs || !params || !frame_counts || n == 0 || !ctxs[0]) return PT_ERR_INVALID_ARG;
  pt_context* B = ctxs[0];
  if (n > 256) return fail(B, PT_ERR_INVALID_ARG, "at most 256 scenes per batch");
  // the shape: the one pt_autotune_batch picked for this composition when
  // params[0] asks for it, else ~16 Mi paths per sub-batch and
  // params[0]'s material-sort and graph flags
  BatchShape shape;
  shape.paths = B->paths_target;
  const uint32_t f0 = params[0].flags;
  shape.sort = (f0 & PT_FLAG_MATERIAL_SORT) != 0;
  shape.graph = (f0 & PT_FLAG_GRAPH) != 0;
  if ((f0 & PT_FLAG_TUNED) && B->batch_tuned.valid && B->batch_tuned.key == batch_composition(ctxs, params, n)) {
    shape.paths = B->batch_tuned.paths;
    shape.sort = B->batch_tuned.sort;
    shape.graph = B->batch_tuned.graph;
  }
  if (f0 & PT_FLAG_DIRECT) shape.graph = false;
  return render_batch_shaped(ctxs, params, frame_counts, n, out, shape);
}

// Launch-shape autotuner of a batch of scenes: renders the batch (with the
// given frame counts, which may be fewer than the production render's)
// under every candidate shape, twice each, and keeps the fastest second
// run.  Candidates: 8, 16, 32, 64, 128 and 256 Mi paths per sub-batch (at
// most the context's paths_target), material sort on and off (when the
// batch has more than one shading program), graph replay on and off.
// Results are identical under every shape.
pt_status pt_autotune_batch(pt_context* const* ctxs, const pt_render_params* params, const uint32_t* frame_counts,
                            uint32_t n, pt_tune_result* out) {
  if (!ctxs || !params || !frame_counts || n == 0 || !ctxs[0]) return PT_ERR_INVALID_ARG;
  pt_context* B = ctxs[0];
  if (n > 256) return fail(B, PT_ERR_INVALID_ARG, "at most 256 scenes per batch");
  uint32_t union_types = 0;
  uint64_t frames = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (!ctxs[i]) return fail(B, PT_ERR_INVALID_ARG, "null context in the batch");
    union_types |= ctxs[i]->mat_types;
    frames += frame_counts[i];
  }
  if (frames == 0) return fail(B, PT_ERR_INVALID_ARG, "no frames to tune on");
  const bool can_sort = __builtin_popcount(union_types & ~(1u << PT_MAT_DIFFUSE_LIGHT)) > 1;
  B->batch_tuned.valid = false;
  BatchShape best;
  double best_ms = 1e300;
  uint32_t cands = 0;
  std::vector<uint64_t> sizes;
  for (uint64_t paths : {1ull << 23, 1ull << 24, 1ull << 25, 1ull << 26, 1ull << 27, 1ull << 28}) {
    paths = std::min(paths, B->paths_target);
    if (std::find(sizes.begin(), sizes.end(), paths) == sizes.end()) sizes.push_back(paths);
  }
  for (uint64_t paths : sizes)
    for (int so = 0; so < (can_sort ? 2 : 1); so++)
      for (int gr = 0; gr < 2; gr++) {
        BatchShape s;
        s.paths = paths;
        s.sort = can_sort && so == 0;
        s.graph = gr != 0;
        pt_render_stats st{};
        for (int rep = 0; rep < 2; rep++) {
          pt_status ps = render_batch_shaped(ctxs, params, frame_counts, n, &st, s);
          if (ps != PT_OK) return ps;
        }
        cands++;
        if (st.gpu_ms < best_ms) {
          best_ms = st.gpu_ms;
          best = s;
        }
      }
  // the candidates' graphs are not kept (the production render's frame
  // counts usually differ)
  if (B->batch_graph_exec) (void)hipGraphExecDestroy(B->batch_graph_exec);
  if (B->batch_graph) (void)hipGraphDestroy(B->batch_graph);
  B->batch_graph_exec = nullptr;
  B->batch_graph = nullptr;
  B->batch_graph_key.clear();
  B->batch_tuned.valid = true;
  B->batch_tuned.paths = best.paths;
  B->batch_tuned.sort = best.sort;
  B->batch_tuned.graph = best.graph;
  B->batch_tuned.key = batch_composition(ctxs, params, n);
  if (out) {
    out->frames_per_batch = 0;
    out->material_sort = best.sort ? 1 : 0;
    out->graph = best.graph ? 1 : 0;
    out->candidates = cands;
    out->ms_per_frame = best_ms / (double)frames;
    out->paths_per_batch = best.paths;
  }
  return PT_OK;
}

// Launch-shape autotuner: renders `frames` frames (frames 0.. of the current
// scene, camera and parameters) under every candidate shape, twice each,
// and keeps the fastest second run.  Candidates: paths in flight 1, 4, 16,
// 64 and 256 Mi, at most the context's paths_target (frames per batch =
// paths / (W H), at least 1), material sort on and off (when the scene has
// more than one shading program), batch graphs on and off.
pt_status pt_autotune(pt_context* c, const pt_render_params* rp, uint32_t frames, pt_tune_result* out) {
  if (!c || !rp || frames == 0) return PT_ERR_INVALID_ARG;
  if (rp->flags & PT_FLAG_RADIANCE_CACHE) return fail(c, PT_ERR_UNSUPPORTED, "autotune the plain path only");
  const uint64_t WH = (uint64_t)rp->width * rp->height;
  if (WH == 0) return fail(c, PT_ERR_INVALID_ARG, "empty image");
  std::vector<uint32_t> fpbs;
  for (uint64_t paths : {1ull << 20, 1ull << 22, 1ull << 24, 1ull << 26, 1ull << 28, 1ull << 29}) {
    paths = std::min(paths, c->paths_target);
    uint32_t f = (uint32_t)std::max<uint64_t>(1, paths / WH);
    f = std::min(f, frames);
    if (f * WH > SLOT_MASK) continue;
    if (std::find(fpbs.begin(), fpbs.end(), f) == fpbs.end()) fpbs.push_back(f);
  }
  const bool can_sort = c->shading_kinds > 1;
  pt_context::Tuned best;
  best.ms_per_frame = 1e300;
  c->tuned.valid = false;
  for (uint32_t f : fpbs)
    for (int so = 0; so < (can_sort ? 2 : 1); so++)
      for (int gr = 0; gr < 2; gr++) {
        pt_render_params p = *rp;
        p.frames_per_batch = f;
        p.flags = (rp->flags & ~(PT_FLAG_TUNED | PT_FLAG_GRAPH | PT_FLAG_MATERIAL_SORT)) |
                  (so ? PT_FLAG_MATERIAL_SORT : 0u) | (gr ? PT_FLAG_GRAPH : 0u);
        // a first render sizes the pools (and captures the graph), the
        // faster of two more is the candidate's time
        pt_render_stats st{};
        double ms = 1e300;
        for (int rep = 0; rep < 3; rep++) {
          pt_status ps = pt_render(c, &p, 0, frames, &st);
          if (ps != PT_OK) return ps;
          if (rep > 0) ms = std::min(ms, st.gpu_ms);
        }
        const double per = ms / frames;
        if (per < best.ms_per_frame) {
          best.ms_per_frame = per;
          best.frames_per_batch = f;
          best.sort = so != 0;
          best.graph = gr != 0;
        }
      }
  best.valid = true;
  best.W = rp->width;
  best.H = rp->height;
  best.max_depth = rp->max_depth;
  best.flags = rp->flags;
  c->tuned = best;
  if (out) {
    out->frames_per_batch = best.frames_per_batch;
    out->material_sort = best.sort ? 1 : 0;
    out->graph = best.graph ? 1 : 0;
    out->candidates = (uint32_t)(fpbs.size() * (can_sort ? 2 : 1) * 2);
    out->ms_per_frame = best.ms_per_frame;
    out->paths_per_batch = (uint64_t)best.frames_per_batch * WH;
  }
  return PT_OK;
}

pt_status pt_cache_configure(pt_context* c, uint32_t query_depth, float train_fraction, uint32_t batch,
                             uint32_t steps, float lr) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!(train_fraction >= 0.f && train_fraction <= 1.f)) return fail(c, PT_ERR_INVALID_ARG, "train_fraction not in [0, 1]");
  if (batch < 32 || batch > (1u << 22)) return fail(c, PT_ERR_INVALID_ARG, "batch not in [32, 2^22]");
  c->cache_query_depth = query_depth;
  c->cache_train_fraction = train_fraction;
  c->cache_batch = batch;
  c->cache_steps = steps;
  c->cache_lr = lr;
  return PT_OK;
}

pt_status pt_cache_judge_configure(pt_context* c, float max_ratio, uint32_t rows) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!(max_ratio >= 0.f)) return fail(c, PT_ERR_INVALID_ARG, "max_ratio must be >= 0");
  if (rows < 32 || rows > (1u << 20)) return fail(c, PT_ERR_INVALID_ARG, "rows not in [32, 2^20]");
  c->cache_judge_max = max_ratio;
  c->cache_judge_rows = rows;
  return PT_OK;
}

pt_status pt_cache_read_records(pt_context* c, int32_t which, uint32_t max, uint16_t* features, float* values,
                                uint32_t* slots, uint32_t* count) {
  if (!c || (which != 0 && which != 1)) return PT_ERR_INVALID_ARG;
  if (!c->cache_rec_cap || !c->cache_last_grid) return fail(c, PT_ERR_STATE, "no radiance-cache render yet");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevPaths& P = c->paths;
  const uint32_t total = which == 0 ? c->cache_last_train : c->cache_last_query;
  const uint32_t n = std::min(max, total);
  if (count) *count = total;
  if (n == 0) return PT_OK;
  // every record in order (rows == total), copied densely, then the first n
  std::vector<DevBuf> tmp;
  uint4* dfeat = nullptr;
  float* dvals = nullptr;
  uint32_t* dslots = nullptr;
  hipError_t e;
  if ((e = dalloc(tmp, (void**)&dfeat, (size_t)total * 192)) != hipSuccess ||
      (e = dalloc(tmp, (void**)&dvals, (size_t)total * 36)) != hipSuccess ||
      (e = dalloc(tmp, (void**)&dslots, (size_t)total * 4)) != hipSuccess) {
    dfree(tmp);
    return hip_fail(c, e, "hipMalloc (record copies)");
  }
  launch_cache_gather(P, which, total, 1u, c->cache_last_grid, dfeat, nullptr, dvals, dslots, c->stream);
  e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && features) e = hipMemcpy(features, dfeat, (size_t)n * 192, hipMemcpyDeviceToHost);
  if (e == hipSuccess && values) e = hipMemcpy(values, dvals, (size_t)n * 36, hipMemcpyDeviceToHost);
  if (e == hipSuccess && slots) e = hipMemcpy(slots, dslots, (size_t)n * 4, hipMemcpyDeviceToHost);
  dfree(tmp);
  if (e != hipSuccess) return hip_fail(c, e, "reading radiance-cache records");
  return PT_OK;
}

pt_status pt_read_env_tables(pt_context* c, uint64_t* cond, uint64_t* marg, uint32_t* width, uint32_t* height,
                             uint64_t* total) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!c->has_scene) return fail(c, PT_ERR_STATE, "no scene");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevScene& s = c->dsc;
  if (width) *width = s.env_w;
  if (height) *height = s.env_h;
  if (total) *total = s.env_total;
  if (cond && s.env_cond) HIPCHK(c, hipMemcpy(cond, s.env_cond, (size_t)s.env_w * s.env_h * 8, hipMemcpyDeviceToHost));
  if (marg && s.env_marg) HIPCHK(c, hipMemcpy(marg, s.env_marg, (size_t)s.env_h * 8, hipMemcpyDeviceToHost));
  return PT_OK;
}

pt_status pt_read_paths(pt_context* c, float* rgb, uint32_t* signature, uint32_t* bounces) {
  if (!c) return PT_ERR_INVALID_ARG;
  if (!c->paths_valid) return fail(c, PT_ERR_STATE, "no frame rendered");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint64_t WH = (uint64_t)c->W * c->H;
  const uint64_t base = (uint64_t)(c->last_fpb - 1) * WH;
  std::vector<float4> res(WH);
  HIPCHK(c, hipMemcpy(res.data(), c->paths.result + base, WH * sizeof(float4), hipMemcpyDeviceToHost));
  if (rgb)
    for (uint64_t i = 0; i < WH; i++) {
      rgb[3 * i] = res[i].x;
      rgb[3 * i + 1] = res[i].y;
      rgb[3 * i + 2] = res[i].z;
    }
  if (bounces)
    for (uint64_t i = 0; i < WH; i++) bounces[i] = fb(res[i].w);
  if (signature) HIPCHK(c, hipMemcpy(signature, c->paths.sig_final + base, WH * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

static pt_status trace_common(pt_context* c, const pt_ray* rays, uint64_t n, void* out, uint32_t* counts, int mode) {
  if (!c || (n && (!rays || !out))) return PT_ERR_INVALID_ARG;
  if (!c->accel_ready) return fail(c, PT_ERR_STATE, "acceleration structure not built");
  if (n == 0) return PT_OK;
  (void)hipSetDevice(c->device);
  if (pt_status ps = ensure_spill(c); ps != PT_OK) return ps;
  std::vector<DevBuf> T;
  pt_ray* dr = nullptr;
  void* dout = nullptr;
  uint32_t* dcnt = nullptr;
  hipError_t e;
  size_t out_bytes = mode == 2 ? n : n * sizeof(pt_hit);
  if ((e = dalloc(T, (void**)&dr, n * sizeof(pt_ray))) || (e = dalloc(T, &dout, out_bytes)) ||
      (counts && (e = dalloc(T, (void**)&dcnt, n * 12)))) {
    dfree(T);
    return hip_fail(c, e, "hipMalloc (rays)");
  }
  e = hipMemcpyAsync(dr, rays, n * sizeof(pt_ray), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_fill_u32(c->qcounters, CTR_WORDS, 0u, c->stream);
    if (mode == 2)
      launch_trace_occluded(c->dsc, dr, n, (uint8_t*)dout, c->spill, c->qcounters + CTR_STACK_OVERFLOW, grid_for(n),
                            c->stream);
    else
      launch_trace_closest(c->dsc, dr, n, (pt_hit*)dout, dcnt, c->spill, c->qcounters + CTR_STACK_OVERFLOW,
                           grid_for(n), c->stream);
    e = hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess && counts) e = hipMemcpyAsync(counts, dcnt, n * 12, hipMemcpyDeviceToHost, c->stream);
  uint32_t ovf = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&ovf, c->qcounters + CTR_STACK_OVERFLOW, 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(T);
  if (e != hipSuccess) return hip_fail(c, e, "trace");
  if (ovf) return fail(c, PT_ERR_DEVICE, "traversal stack overflow");
  return PT_OK;
}

pt_status pt_trace_closest(pt_context* c, const pt_ray* rays, uint64_t n, pt_hit* hits) {
  return trace_common(c, rays, n, hits, nullptr, 0);
}
pt_status pt_trace_closest_counted(pt_context* c, const pt_ray* rays, uint64_t n, pt_hit* hits, uint32_t* counts) {
  if (!counts) return PT_ERR_INVALID_ARG;
  return trace_common(c, rays, n, hits, counts, 1);
}
pt_status pt_trace_occluded(pt_context* c, const pt_ray* rays, uint64_t n, uint8_t* occ) {
  return trace_common(c, rays, n, occ, nullptr, 2);
}
pt_status pt_trace_closest_device(pt_context* c, const pt_ray* d_rays, uint64_t n, pt_hit* d_hits) {
  if (!c || (n && (!d_rays || !d_hits))) return PT_ERR_INVALID_ARG;
  if (!c->accel_ready) return fail(c, PT_ERR_STATE, "acceleration structure not built");
  if (n == 0) return PT_OK;
  (void)hipSetDevice(c->device);
  if (pt_status ps = ensure_spill(c); ps != PT_OK) return ps;
  launch_trace_closest(c->dsc, d_rays, n, d_hits, nullptr, c->spill, c->qcounters + CTR_STACK_OVERFLOW, grid_for(n),
                       c->stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, "trace launch");
  return PT_OK;
}

pt_status pt_trace_closest_device_noalloc(pt_context* c, const pt_ray* d_rays, uint64_t n, pt_hit* d_hits,
                                          uint32_t* d_overflow) {
  if (!c || (n && (!d_rays || !d_hits))) return PT_ERR_INVALID_ARG;
  if (!c->accel_ready) return fail(c, PT_ERR_STATE, "acceleration structure not built");
  if (n > PT_NOALLOC_MAX_RAYS)
    return fail(c, PT_ERR_INVALID_ARG, "pt_trace_closest_device_noalloc: more than PT_NOALLOC_MAX_RAYS (2^31) rays");
  if (n == 0) return PT_OK;
  (void)hipSetDevice(c->device);
  unsigned char* base = static_cast<unsigned char*>(g_scratch[c->device]);
  if (!base) return fail(c, PT_ERR_STATE, "scratch of this device was not resolved");
  uint2* spill = reinterpret_cast<uint2*>(base);
  uint32_t* overflow = reinterpret_cast<uint32_t*>(base + NOALLOC_SPILL_BYTES);
  // the overflow counter starts at zero (a fill kernel, not a memset node)
  launch_fill_u32(overflow, 1, 0u, c->stream);
  launch_trace_closest(c->dsc, d_rays, n, d_hits, nullptr, spill, overflow,
                       std::min<int>(grid_for(n), (int)NOALLOC_GRID), c->stream);
  if (d_overflow) {
    hipError_t e = hipMemcpyAsync(d_overflow, overflow, 4, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "overflow copy");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, "trace launch");
  return PT_OK;
}

}  // extern "C"
