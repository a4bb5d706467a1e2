// capi.hip -- the C ABI (include/meshhashing.h): volume lifecycle, frame
// upload, stage launchers, canonical dumps and restore.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "mh_kernels.h"
#include "mh_tune.h"
#include "mh_mc_tables.h"

using namespace mh;

static constexpr McTables kHostMc = build_mc_tables();

namespace {

thread_local std::string g_err;
int scratch_address(int dev, unsigned char** out);

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(MH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

// 4x4 inverse by 2x2 minors (Laplace expansion over rows 0-1 / 2-3), the
// general inverse the reference uses for poses (matrix.h:964-1089 computes
// the same adjugate / determinant).  T = float gives the FP32 pose used by
// every decision, T = double the FP64 pose used for SDF values.
template <typename T>
void invert4(const T* a, T* r) {
  const T s0 = a[0] * a[5] - a[4] * a[1];
  const T s1 = a[0] * a[6] - a[4] * a[2];
  const T s2 = a[0] * a[7] - a[4] * a[3];
  const T s3 = a[1] * a[6] - a[5] * a[2];
  const T s4 = a[1] * a[7] - a[5] * a[3];
  const T s5 = a[2] * a[7] - a[6] * a[3];
  const T c5 = a[10] * a[15] - a[14] * a[11];
  const T c4 = a[9] * a[15] - a[13] * a[11];
  const T c3 = a[9] * a[14] - a[13] * a[10];
  const T c2 = a[8] * a[15] - a[12] * a[11];
  const T c1 = a[8] * a[14] - a[12] * a[10];
  const T c0 = a[8] * a[13] - a[12] * a[9];
  const T det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  const T id = T(1) / det;
  r[0] = (a[5] * c5 - a[6] * c4 + a[7] * c3) * id;
  r[1] = (-a[1] * c5 + a[2] * c4 - a[3] * c3) * id;
  r[2] = (a[13] * s5 - a[14] * s4 + a[15] * s3) * id;
  r[3] = (-a[9] * s5 + a[10] * s4 - a[11] * s3) * id;
  r[4] = (-a[4] * c5 + a[6] * c2 - a[7] * c1) * id;
  r[5] = (a[0] * c5 - a[2] * c2 + a[3] * c1) * id;
  r[6] = (-a[12] * s5 + a[14] * s2 - a[15] * s1) * id;
  r[7] = (a[8] * s5 - a[10] * s2 + a[11] * s1) * id;
  r[8] = (a[4] * c4 - a[5] * c2 + a[7] * c0) * id;
  r[9] = (-a[0] * c4 + a[1] * c2 - a[3] * c0) * id;
  r[10] = (a[12] * s4 - a[13] * s2 + a[15] * s0) * id;
  r[11] = (-a[8] * s4 + a[9] * s2 - a[11] * s0) * id;
  r[12] = (-a[4] * c3 + a[5] * c1 - a[6] * c0) * id;
  r[13] = (a[0] * c3 - a[1] * c1 + a[2] * c0) * id;
  r[14] = (-a[12] * s3 + a[13] * s1 - a[14] * s0) * id;
  r[15] = (a[8] * s3 - a[9] * s1 + a[10] * s0) * id;
}

template <typename T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, n * sizeof(T) + 16);
}

// Device allocations freed together when the holder goes out of scope,
// unless handed over to a long-lived owner list first: an entry point that
// needs several buffers publishes them only once every one of them exists.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t get(T** p, size_t n) {
    const hipError_t e = dalloc(p, n);
    if (e == hipSuccess) ptrs.push_back((void*)*p);
    return e;
  }
  void release_to(std::vector<void*>& owner) {
    owner.insert(owner.end(), ptrs.begin(), ptrs.end());
    ptrs.clear();
  }
};

}  // namespace

// Online marching-cubes tuning (of a volume's meshing calls or a batch's
// graph replays): trial p of n_trials runs mc_trial_shape(p, kNumMcShapes);
// its time is read at the next call, per worklist block for a volume (n_host
// set: frames differ in size), per frame for a batch.
struct McTune {
  bool active = false;
  uint32_t trial = 0, n_trials = 0;
  int pending = -1;  // shape whose events are recorded but not yet read
  hipEvent_t a = nullptr, b = nullptr;
  uint32_t* n_host = nullptr;  // pinned: the pending trial's worklist length (volumes)
  std::vector<double> ms;      // each trial's time, in order
  mh_tune_mc_result result{};
};

static void mc_tune_free(McTune& t) {
  if (t.a) (void)hipEventDestroy(t.a);
  if (t.b) (void)hipEventDestroy(t.b);
  if (t.n_host) (void)hipHostFree(t.n_host);
  t = McTune{};
}

struct mh_volume {
  mh_hash_params hp;
  mh_volume_params vp;
  mh_mesh_params mp;
  mh_sensor_params sp;
  bool psdf = false;
  View V{};
  hipStream_t stream = nullptr;
  int device = 0;
  int cus = 256;
  uint16_t* raw = nullptr;  // staging for host uploads
  uint8_t* raw_color = nullptr;  // staging for host colour uploads (BGRA)
  // Frame input not converted yet: mh_alloc_blocks converts it inside the
  // allocation kernel (one pass over the image); any other stage that reads
  // the frame first runs the standalone conversion (flush_frame).
  const uint16_t* pending_raw = nullptr;
  const uint8_t* pending_color = nullptr;
  bool have_frame = false;
  int32_t* pd_nbr = nullptr;
  uint32_t pd_nbr_cap = 0;
  // Bumped by every call that rebuilds or clears the worklist; the
  // primal-dual neighbour table records the value it was built at.
  uint64_t work_gen = 0;
  bool cand_cleared = false;  // the worklist counter is known to be 0 (an allocation pass ran)
  uint64_t pd_work_gen = ~0ull;
  double* pd_partials = nullptr;
  double* pd_energy = nullptr;
  uint32_t pd_npartials = 0;
  uint32_t mc_pass = 0;
  // mesh compaction (allocated on first use)
  uint32_t* cm_used = nullptr;    // per vertex slot: used by a live triangle
  uint32_t* cm_remap = nullptr;   // per vertex slot: compact index
  uint32_t* cm_blk = nullptr;     // per worklist block: triangle count, then offset
  uint32_t* cm_blk_off = nullptr;
  uint32_t* cm_sums = nullptr;    // scan chunk sums
  uint32_t* cm_total = nullptr;   // 2 totals: triangles, vertices
  float* cm_pos = nullptr;
  float* cm_nrm = nullptr;
  float* cm_col = nullptr;
  int32_t* cm_tri = nullptr;
  uint32_t cm_vertices = 0, cm_triangles = 0;
  double* rgrad_partials = nullptr;     // raycast adjoint: per-workgroup sums
  uint32_t rgrad_cap = 0;
  double* rgrad_out = nullptr;
  uint32_t* ray_range = nullptr;        // raycast: camera-z range per 8x8 tile
  uint32_t ray_range_words = 0;
  int tiling = kDefaultIntegrateTiling;  // integration tiling (kIntegrateTilings)
  int grid_per_cu = 8;                   // integration workgroups per CU
  int ray_shape = 0;                     // raycast tile shape (kRayTileWidths)
  int mc_shape = 0;                      // marching-cubes launch shape (kMcShapes)
  int psdf_tiling = kDefaultPsdfTiling;  // PSDF fusion workgroup size (kPsdfLanes)
  int psdf_grid_per_cu = 8;
  McTune mct;  // online marching-cubes tuning (mh_tune_marching_cubes)
  std::vector<void*> owned;
  // diagnostics (mh_volume_set_diagnostics)
  uint32_t diag = 0;
  struct Interval {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> ev_pool, ev_used;
  std::vector<Interval> intervals;
  const char* open_name = nullptr;
  hipEvent_t open_ev = nullptr;
  std::vector<std::string> kt_names;
  std::vector<uint64_t> kt_launches;
  std::vector<double> kt_ms;
  KernelClock clock{};
  uint32_t* ray_mark = nullptr;  // per block (MH_DIAG_RAY_BLOCKS)
  int32_t* mc_nbt = nullptr;     // marching cubes: 27 neighbours per worklist item
  uint32_t ray_stamp = 0;

  int grid_wg256() const { return cus * 8; }
  int grid_wg512() const { return cus * 4; }
  const KernelClock* kc() const { return (diag & MH_DIAG_KERNEL_TIMES) ? &clock : nullptr; }
};

namespace {

int check_vol(mh_volume* v) {
  if (!v) return fail(MH_ERR_ARG, "null volume");
  return MH_OK;
}

// Per-kernel timing: an event at every tick; the interval from one tick to
// the next is the named kernel's.  Event creation failures drop the
// interval (timing is best effort and changes no result).
void kernel_tick(void* ctx, hipStream_t s, const char* next) {
  mh_volume* v = static_cast<mh_volume*>(ctx);
  hipEvent_t e = nullptr;
  if (!v->ev_pool.empty()) {
    e = v->ev_pool.back();
    v->ev_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  if (e && hipEventRecord(e, s) == hipSuccess) {
    v->ev_used.push_back(e);
  } else {
    if (e) v->ev_pool.push_back(e);
    e = nullptr;
  }
  if (v->open_name && v->open_ev && e) v->intervals.push_back({v->open_name, v->open_ev, e});
  v->open_name = next;
  v->open_ev = e;
}

// Waits for the stream and adds the recorded intervals to the totals.
int harvest_times(mh_volume* v) {
  HIP_TRY(hipStreamSynchronize(v->stream));
  for (const auto& iv : v->intervals) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, iv.a, iv.b) != hipSuccess) continue;
    size_t k = 0;
    while (k < v->kt_names.size() && v->kt_names[k] != iv.name) ++k;
    if (k == v->kt_names.size()) {
      v->kt_names.push_back(iv.name);
      v->kt_launches.push_back(0);
      v->kt_ms.push_back(0.0);
    }
    v->kt_launches[k] += 1;
    v->kt_ms[k] += ms;
  }
  v->intervals.clear();
  if (!v->open_name) {  // no interval open: every used event can be reused
    v->ev_pool.insert(v->ev_pool.end(), v->ev_used.begin(), v->ev_used.end());
    v->ev_used.clear();
  }
  return MH_OK;
}

void flush_frame(mh_volume* v) {
  if (v->pending_raw) launch_convert_depth(v->V, v->pending_raw, v->stream);
  if (v->pending_color) launch_convert_color(v->V, v->pending_color, v->stream);
  v->pending_raw = nullptr;
  v->pending_color = nullptr;
}

Counters read_counters(mh_volume* v) {
  Counters c{};
  (void)hipMemcpyAsync(&c, v->V.ctr, sizeof(Counters), hipMemcpyDeviceToHost, v->stream);
  (void)hipStreamSynchronize(v->stream);
  return c;
}

void set_pose(mh_volume* v, const float wTc[16]) {
  float c32[16];
  invert4<float>(wTc, c32);
  double w64[16], c64[16];
  for (int i = 0; i < 16; ++i) w64[i] = (double)wTc[i];
  invert4<double>(w64, c64);
  std::memcpy(v->V.wTc.m, wTc, sizeof(float) * 16);
  std::memcpy(v->V.cTw.m, c32, sizeof(float) * 16);
  for (int i = 0; i < 12; ++i) v->V.cTw64.m[i] = c64[i];
}

// MarchingCubes constants of one call (thresholds, then the next pass stamp).
int build_mc(mh_volume* v, const mh_mc_params* p, McConst* c) {
  mh_mc_params d;
  mh_mc_params_default(&d);
  if (!p) p = &d;
  if (!p->enable_sdf_gradient)
    return fail(MH_ERR_ARG, "flat per-triangle normals are not supported");
  c->sdf_threshold = p->sdf_threshold;
  c->rho_min = p->rho_min;
  const float inv = 1.0f / v->vp.voxel_size;
  c->weight_min = p->weight_min < 0.0f ? inv * inv : p->weight_min;
  c->snap = p->snap_distance;
  c->flat_normals = 0;
  c->pass = ++v->mc_pass;
  return MH_OK;
}

// RayCaster constants (ray_caster.cu:152-224): wTc is the FP32 inverse of cTw.
int build_ray(const mh_volume* v, const mh_raycaster_params* p, const float cTw[16], RayConst* R) {
  if (p->width == 0 || p->height == 0 || p->width > 16384 || p->height > 16384 ||
      !(p->raycast_step > 0.0f) || !(p->min_raycast_depth > 0.0f) ||
      !(p->max_raycast_depth >= p->min_raycast_depth) ||
      (p->max_raycast_depth - p->min_raycast_depth) / p->raycast_step > 65536.0f)
    return fail(MH_ERR_ARG, "bad raycaster params (sizes, depth range, or > 65536 steps per ray)");
  R->cam = Cam{p->fx, p->fy, p->cx, p->cy, p->width, p->height, p->min_raycast_depth,
               p->max_raycast_depth};
  R->boost = 0;
  R->min_depth = p->min_raycast_depth;
  R->max_depth = p->max_raycast_depth;
  R->step = p->raycast_step;
  R->sample_sdf_threshold = p->sample_sdf_threshold;
  R->sdf_threshold = p->sdf_threshold;
  R->enable_gradients = p->enable_gradients;
  R->inv_vs = 1.0f / v->vp.voxel_size;
  std::memcpy(R->cTw.m, cTw, sizeof(float) * 16);
  invert4<float>(cTw, R->wTc.m);
  R->range = nullptr;
  R->tiles_x = (p->width + 7) / 8;
  R->tiles = R->tiles_x * ((p->height + 7) / 8);
  return MH_OK;
}

// The volume's tile-range buffer, grown to the raycast's size.
int attach_range(mh_volume* v, RayConst* R) {
  const uint32_t words = ray_range_words(R->cam.width, R->cam.height);
  if (words > v->ray_range_words) {
    if (v->ray_range) (void)hipFree(v->ray_range);
    v->ray_range = nullptr;
    v->ray_range_words = 0;
    HIP_TRY(dalloc(&v->ray_range, words));
    v->ray_range_words = words;
  }
  R->range = v->ray_range;
  return MH_OK;
}

}  // namespace

// ------------------------------------------------------------------ dumps
namespace {

struct BlockRec {
  i3 pos;
  int32_t ptr;
};

bool pos_less(const i3& a, const i3& b) {
  if (a.x != b.x) return a.x < b.x;
  if (a.y != b.y) return a.y < b.y;
  return a.z < b.z;
}

int sorted_blocks(mh_volume* v, std::vector<BlockRec>& out) {
  const Counters c = read_counters(v);
  const uint32_t hwm = c.block_hwm;
  std::vector<uint64_t> keys(hwm);
  if (hwm)
    HIP_TRY(hipMemcpy(keys.data(), v->V.block_keys, hwm * sizeof(uint64_t), hipMemcpyDeviceToHost));
  out.clear();
  for (uint32_t p = 0; p < hwm; ++p)
    if (keys[p] != kEmptyKey) out.push_back(BlockRec{unpack_key(keys[p]), (int32_t)p});
  std::sort(out.begin(), out.end(),
            [](const BlockRec& a, const BlockRec& b) { return pos_less(a.pos, b.pos); });
  return MH_OK;
}

template <typename T>
int gather_field(mh_volume* v, const T* dev, const std::vector<BlockRec>& blocks, uint32_t hwm,
                 size_t per_voxel, T* host) {
  if (!host || !dev) return MH_OK;
  std::vector<T> all((size_t)hwm * kBlockVoxels * per_voxel);
  if (!all.empty())
    HIP_TRY(hipMemcpy(all.data(), dev, all.size() * sizeof(T), hipMemcpyDeviceToHost));
  const size_t bs = (size_t)kBlockVoxels * per_voxel;
  for (size_t i = 0; i < blocks.size(); ++i)
    std::memcpy(host + i * bs, all.data() + (size_t)blocks[i].ptr * bs, bs * sizeof(T));
  return MH_OK;
}

}  // namespace

extern "C" {

const char* mh_last_error(void) { return g_err.c_str(); }

void mh_mc_params_default(mh_mc_params* out) {
  if (!out) return;
  out->sdf_threshold = 0.20f;
  out->rho_min = 0.1f;
  out->weight_min = -1.0f;
  out->snap_distance = 0.008;
  out->enable_sdf_gradient = 1;
}

int mh_volume_create(const mh_hash_params* hash, const mh_volume_params* vol,
                     const mh_mesh_params* mesh, const mh_sensor_params* sensor, int enable_psdf,
                     mh_volume** out) {
  if (!hash || !vol || !mesh || !sensor || !out) return fail(MH_ERR_ARG, "null argument");
  *out = nullptr;
  if (hash->bucket_count == 0 || hash->bucket_size == 0 ||
      (uint64_t)hash->bucket_count * hash->bucket_size != hash->entry_count)
    return fail(MH_ERR_ARG, "entry_count must equal bucket_count * bucket_size");
  if (hash->value_capacity == 0 || hash->value_capacity > (1u << 30))
    return fail(MH_ERR_ARG, "value_capacity out of range");
  if (mesh->max_vertex_count < 2 || mesh->max_triangle_count < 1)
    return fail(MH_ERR_ARG, "mesh pools too small");
  if (sensor->width == 0 || sensor->height == 0 || sensor->width > 16384 || sensor->height > 16384)
    return fail(MH_ERR_ARG, "sensor size out of range");
  if (!(vol->voxel_size > 0.0f)) return fail(MH_ERR_ARG, "voxel_size must be positive");

  mh_volume* v = new mh_volume();
  v->hp = *hash;
  v->vp = *vol;
  v->mp = *mesh;
  v->sp = *sensor;
  v->psdf = enable_psdf != 0;
  if (hipGetDevice(&v->device) != hipSuccess) {
    delete v;
    return fail(MH_ERR_HIP, "no HIP device");
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, v->device) == hipSuccess &&
      cus > 0)
    v->cus = cus;
  {
    // set-up of the allocation-free adjoint: this device's scratch address
    unsigned char* sa = nullptr;
    const int rc = scratch_address(v->device, &sa);
    if (rc) {
      delete v;
      return rc;
    }
  }

  v->clock = KernelClock{kernel_tick, v};
  View& V = v->V;
  const size_t cap = hash->value_capacity;
  const size_t nvox = cap * kBlockVoxels;
  const size_t npix = (size_t)sensor->width * sensor->height;
  bool ok = true;
  auto A = [&](auto** p, size_t n) {
    if (!ok) return;
    if (dalloc(p, n) != hipSuccess) {
      ok = false;
      return;
    }
    v->owned.push_back((void*)*p);
  };
  A(&V.entries, hash->entry_count);
  A(&V.block_keys, cap);
  A(&V.block_heap, cap);
  A(&V.occ, (size_t)1 << (kOccBits - 5));
  A(&V.occ2, (size_t)1 << (kOcc2Bits - 5));
  A(&V.ctr, 1);
  A(&V.ovf, kOverflowQueue);
  A(&V.sdf, nvox);
  A(&V.weight, nvox);
  A(&V.color, nvox);
  if (v->psdf) {
    A(&V.va, nvox);
    A(&V.vb, nvox);
  }
  A(&V.vptr, nvox * 3);
  A(&V.cube, nvox);
  A(&V.surfel, cap * 2);
  A(&V.block_mark, cap);
  A(&V.vertices, mesh->max_vertex_count);
  A(&V.vertex_heap, mesh->max_vertex_count);
  A(&V.vfree, mesh->max_vertex_count);
  A(&V.cand, cap);
  A(&v->mc_nbt, cap * 27);
  A(&V.cand_flag, cap);
  A(&V.life, cap);
  A(&V.gc_list, cap);
  A(&V.gc_free, cap);
  A(&V.depth, npix);
  A(&V.inlier, npix);
  A(&V.color_in, npix);
  A(&v->raw, npix);
  A(&v->raw_color, npix * 4);
  if (!ok) {
    for (void* p : v->owned) (void)hipFree(p);
    delete v;
    return fail(MH_ERR_HIP, "device allocation failed");
  }
  V.bucket_count = hash->bucket_count;
  V.bucket_size = hash->bucket_size;
  V.entry_count = hash->entry_count;
  V.linked_list_size = hash->linked_list_size;
  V.capacity = hash->value_capacity;
  V.max_vertices = mesh->max_vertex_count;
  V.max_triangles = mesh->max_triangle_count;
  V.geo = Geo{vol->voxel_size, vol->truncation_distance, vol->truncation_distance_scale,
              vol->sdf_upper_bound, (float)vol->weight_sample};
  V.cam = Cam{sensor->fx, sensor->fy, sensor->cx, sensor->cy, sensor->width, sensor->height,
              sensor->min_depth_range, sensor->max_depth_range};
  V.range_factor = sensor->range_factor;
  const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  set_pose(v, ident);
  *out = v;
  const int rc = mh_volume_reset(v);
  if (rc != MH_OK) {
    mh_volume_destroy(v);
    *out = nullptr;
    return rc;
  }
  return MH_OK;
}

int mh_volume_destroy(mh_volume* v) {
  if (!v) return MH_OK;
  (void)hipStreamSynchronize(v->stream);
  for (void* p : v->owned) (void)hipFree(p);
  if (v->rgrad_partials) (void)hipFree(v->rgrad_partials);
  if (v->rgrad_out) (void)hipFree(v->rgrad_out);
  if (v->ray_range) (void)hipFree(v->ray_range);
  if (v->ray_mark) (void)hipFree(v->ray_mark);
  for (hipEvent_t e : v->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : v->ev_used) (void)hipEventDestroy(e);
  mc_tune_free(v->mct);
  delete v;
  return MH_OK;
}

int mh_volume_reset(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  launch_reset(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(v->stream));
  v->have_frame = false;
  ++v->work_gen;
  return MH_OK;
}

int mh_volume_set_diagnostics(mh_volume* v, uint32_t flags) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (flags & ~(uint32_t)(MH_DIAG_KERNEL_TIMES | MH_DIAG_RAY_BLOCKS))
    return fail(MH_ERR_ARG, "unknown diagnostics flag");
  if ((flags & MH_DIAG_RAY_BLOCKS) && !v->ray_mark) {
    uint32_t* m = nullptr;
    HIP_TRY(dalloc(&m, v->hp.value_capacity));
    const hipError_t e = hipMemsetAsync(m, 0, (size_t)v->hp.value_capacity * 4, v->stream);
    if (e != hipSuccess) {
      (void)hipFree(m);
      return fail(MH_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    }
    v->ray_mark = m;
  }
  if (!(flags & MH_DIAG_KERNEL_TIMES) && v->open_name) kernel_tick(v, v->stream, nullptr);
  v->diag = flags;
  return MH_OK;
}

int mh_kernel_times(mh_volume* v, uint32_t capacity, mh_kernel_time* out, uint32_t* count,
                    int reset) {
  if (check_vol(v)) return MH_ERR_ARG;
  const int rc = harvest_times(v);
  if (rc) return rc;
  const uint32_t n = (uint32_t)v->kt_names.size();
  if (count) *count = n;
  if (out) {
    if (capacity < n) return fail(MH_ERR_CAPACITY, "kernel time buffer too small");
    for (uint32_t i = 0; i < n; ++i) {
      std::memset(out[i].name, 0, sizeof(out[i].name));
      std::strncpy(out[i].name, v->kt_names[i].c_str(), sizeof(out[i].name) - 1);
      out[i].launches = v->kt_launches[i];
      out[i].total_ms = v->kt_ms[i];
    }
  }
  if (reset) {
    v->kt_names.clear();
    v->kt_launches.clear();
    v->kt_ms.clear();
  }
  return MH_OK;
}

int mh_volume_set_stream(mh_volume* v, void* s) {
  if (check_vol(v)) return MH_ERR_ARG;
  v->stream = (hipStream_t)s;
  return MH_OK;
}

int mh_synchronize(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  return MH_OK;
}

int mh_get_status(mh_volume* v, mh_status* out) {
  if (check_vol(v) || !out) return fail(MH_ERR_ARG, "null argument");
  const Counters c = read_counters(v);
  out->flags = c.flags;
  out->dropped_inserts = c.dropped;
  out->dda_cap_hits = c.dda_cap;
  out->list_inserts = c.list_inserts;
  out->block_count = (uint32_t)((int64_t)v->hp.value_capacity - 1 - c.block_top);
  out->candidate_count = c.cand_count;
  out->vertex_count = (uint32_t)((int64_t)v->mp.max_vertex_count - 2 - c.vertex_top);
  const WorkStripe w = work_totals(c);
  out->triangle_count = (uint32_t)w.tri_live;
  if (w.tri_live > v->mp.max_triangle_count) out->flags |= MH_STATUS_TRIANGLE_HEAP_EXHAUSTED;
  out->voxel_updates = w.voxel_updates;
  out->blocks_integrated = w.blocks_integrated;
  out->mc_vertex_writes = w.mc_vertex_writes;
  out->mc_triangle_writes = w.mc_triangle_writes;
  out->mc_blocks = w.mc_blocks;
  out->ray_samples = w.ray_samples;
  out->ray_blocks = w.ray_blocks;
  return MH_OK;
}

int mh_frame_set_pose(mh_volume* v, const float wTc[16]) {
  if (check_vol(v) || !wTc) return fail(MH_ERR_ARG, "null argument");
  set_pose(v, wTc);
  return MH_OK;
}

int mh_frame_depth_u16_device(mh_volume* v, const uint16_t* depth_dev, const float wTc[16]) {
  if (check_vol(v) || !depth_dev || !wTc) return fail(MH_ERR_ARG, "null argument");
  set_pose(v, wTc);
  v->pending_raw = depth_dev;
  v->pending_color = nullptr;
  v->have_frame = true;
  v->V.has_color = 0;
  return MH_OK;
}

int mh_frame_color_bgra8_device(mh_volume* v, const uint8_t* bgra_dev) {
  if (check_vol(v) || !bgra_dev) return fail(MH_ERR_ARG, "null argument");
  if ((uintptr_t)bgra_dev % 4) return fail(MH_ERR_ARG, "colour buffer must be 4-byte aligned");
  if (!v->have_frame) return fail(MH_ERR_STATE, "upload the frame's depth first");
  if (v->pending_raw) {
    v->pending_color = bgra_dev;  // converted with the depth
  } else {
    launch_convert_color(v->V, bgra_dev, v->stream);
    HIP_TRY(hipGetLastError());
  }
  v->V.has_color = 1;
  return MH_OK;
}

int mh_frame_upload_color_bgra8(mh_volume* v, const uint8_t* bgra_host) {
  if (check_vol(v) || !bgra_host) return fail(MH_ERR_ARG, "null argument");
  const size_t n = (size_t)v->sp.width * v->sp.height * 4;
  HIP_TRY(hipMemcpyAsync(v->raw_color, bgra_host, n, hipMemcpyHostToDevice, v->stream));
  return mh_frame_color_bgra8_device(v, v->raw_color);
}

int mh_frame_upload_depth_u16(mh_volume* v, const uint16_t* depth_host, const float wTc[16]) {
  if (check_vol(v) || !depth_host || !wTc) return fail(MH_ERR_ARG, "null argument");
  const size_t n = (size_t)v->sp.width * v->sp.height;
  HIP_TRY(hipMemcpyAsync(v->raw, depth_host, n * sizeof(uint16_t), hipMemcpyHostToDevice,
                         v->stream));
  return mh_frame_depth_u16_device(v, v->raw, wTc);
}

int mh_frame_depth_device_ptr(mh_volume* v, const float** depth_dev) {
  if (check_vol(v) || !depth_dev) return fail(MH_ERR_ARG, "null argument");
  flush_frame(v);
  HIP_TRY(hipGetLastError());
  *depth_dev = v->V.depth;
  return MH_OK;
}

int mh_inlier_ratio_device_ptr(mh_volume* v, float** inlier_dev) {
  if (check_vol(v) || !inlier_dev) return fail(MH_ERR_ARG, "null argument");
  flush_frame(v);
  HIP_TRY(hipGetLastError());
  *inlier_dev = v->V.inlier;
  return MH_OK;
}

int mh_volume_field_device_ptr(mh_volume* v, int field, void** field_dev, uint64_t* count) {
  if (check_vol(v) || !field_dev) return fail(MH_ERR_ARG, "null argument");
  void* p = nullptr;
  switch (field) {
    case MH_FIELD_SDF: p = v->V.sdf; break;
    case MH_FIELD_WEIGHT: p = v->V.weight; break;
    case MH_FIELD_COLOR: p = v->V.color; break;
    case MH_FIELD_BETA_A: p = v->V.va; break;
    case MH_FIELD_BETA_B: p = v->V.vb; break;
    default: return fail(MH_ERR_ARG, "unknown field");
  }
  if (!p) return fail(MH_ERR_STATE, "volume has no PSDF fields");
  *field_dev = p;
  if (count) *count = (uint64_t)v->hp.value_capacity * kBlockVoxels;
  return MH_OK;
}

int mh_alloc_blocks(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  View Va = v->V;  // the frame's conversion runs inside the allocation kernel
  Va.raw = v->pending_raw;
  Va.raw_color = v->pending_color;
  v->pending_raw = nullptr;
  v->pending_color = nullptr;
  launch_alloc(Va, v->stream, v->kc());
  ++v->work_gen;  // allocation clears the worklist
  v->cand_cleared = true;
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

static int collect(mh_volume* v, int frustum, uint32_t* count_host) {
  launch_collect(v->V, frustum, v->stream, v->cus * 4, v->kc(), v->cand_cleared);
  v->cand_cleared = false;
  ++v->work_gen;
  HIP_TRY(hipGetLastError());
  if (count_host) *count_host = read_counters(v).cand_count;
  return MH_OK;
}

int mh_collect_visible(mh_volume* v, uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  return collect(v, 1, count_host);
}

int mh_collect_all(mh_volume* v, uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  return collect(v, 0, count_host);
}

int mh_collect_low_surfel_blocks(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  launch_gc_low_surfel(v->V, v->stream, v->cus * 4);
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_starve_blocks(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  launch_gc_starve(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_collect_garbage_blocks(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  launch_gc_flag_garbage(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_recycle_garbage_blocks(mh_volume* v, uint32_t* freed_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  launch_gc_recycle(v->V, v->stream, v->cus);
  ++v->work_gen;  // freed blocks may sit on the worklist
  HIP_TRY(hipGetLastError());
  if (freed_host) {
    Counters c = read_counters(v);
    *freed_host = c.gc_free_count;
  }
  return MH_OK;
}

int mh_integrate_tsdf(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  flush_frame(v);
  launch_integrate_tsdf(v->V, v->stream, v->cus * v->grid_per_cu, v->tiling, v->kc());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_predict_inlier(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->psdf) return fail(MH_ERR_STATE, "volume created without PSDF fields");
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  flush_frame(v);
  launch_predict_inlier(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_integrate_psdf(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->psdf) return fail(MH_ERR_STATE, "volume created without PSDF fields");
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  flush_frame(v);
  launch_integrate_psdf(v->V, v->stream, v->cus * v->psdf_grid_per_cu, kPsdfLanes[v->psdf_tiling]);
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

// Reads the last tuning trial's time; after the last trial sets *shape by
// the rule of mc_pick.
static int mc_tune_collect(McTune& t, hipStream_t stream, int* shape) {
  if (t.pending >= 0) {
    HIP_TRY(hipEventSynchronize(t.b));
    if (t.n_host) HIP_TRY(hipStreamSynchronize(stream));  // the worklist length copy
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
    t.ms.push_back(!t.n_host ? (double)ms : (*t.n_host > 0 ? ms / *t.n_host : (double)NAN));
    t.pending = -1;
  }
  if (t.active && t.trial >= t.n_trials) {
    const McPick pick = mc_pick(t.ms, kNumMcShapes);
    const int best = pick.shape;
    const double default_ms = pick.default_ms, best_ms = pick.best_ms;
    *shape = best;
    t.active = false;
    t.result.done = 1;
    t.result.shape = best;
    t.result.workgroups_per_cu = kMcShapes[best].grid_per_cu;
    t.result.static_quarters = kMcShapes[best].static_quarters;
    t.result.candidates = kNumMcShapes;
    t.result.best_ms = (float)best_ms;
    t.result.default_ms = (float)default_ms;
    t.result.blocks_per_ms = t.n_host && best_ms > 0.0 ? (float)(1.0 / best_ms) : 0.0f;
  }
  return MH_OK;
}

// Starts an online tuning of reps rounds (2 (kNumMcShapes - 1) trials each,
// and one more trial of shape 0).
static int mc_tune_start(McTune& t, hipStream_t stream, uint32_t reps, bool per_block, int* shape) {
  if (!t.a) HIP_TRY(hipEventCreate(&t.a));
  if (!t.b) HIP_TRY(hipEventCreate(&t.b));
  if (per_block && !t.n_host) HIP_TRY(hipHostMalloc((void**)&t.n_host, sizeof(uint32_t), hipHostMallocDefault));
  const int rc = mc_tune_collect(t, stream, shape);  // a trial of an earlier tuning still pending
  if (rc) return rc;
  t.active = true;
  t.trial = 0;
  t.n_trials = (reps ? reps : 1) * 2u * (kNumMcShapes - 1) + 1u;
  t.ms.clear();
  t.result = mh_tune_mc_result{};
  return MH_OK;
}

int mh_marching_cubes(mh_volume* v, const mh_mc_params* p) {
  if (check_vol(v)) return MH_ERR_ARG;
  McConst c;
  int rc = build_mc(v, p, &c);
  if (rc) return rc;
  auto& t = v->mct;
  if (t.active || t.pending >= 0) {
    rc = mc_tune_collect(t, v->stream, &v->mc_shape);
    if (rc) return rc;
  }
  if (t.active) {
    const int k = mc_trial_shape(t.trial++, kNumMcShapes);
    HIP_TRY(hipEventRecord(t.a, v->stream));
    launch_marching_cubes(v->V, c, v->mc_nbt, v->stream, v->cus, v->kc(), kMcShapes[k]);
    HIP_TRY(hipEventRecord(t.b, v->stream));
    HIP_TRY(hipMemcpyAsync(t.n_host, &v->V.ctr->cand_count, sizeof(uint32_t), hipMemcpyDeviceToHost, v->stream));
    t.pending = k;
  } else {
    launch_marching_cubes(v->V, c, v->mc_nbt, v->stream, v->cus, v->kc(), kMcShapes[v->mc_shape]);
  }
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_tune_marching_cubes(mh_volume* v, uint32_t reps) {
  if (check_vol(v)) return MH_ERR_ARG;
  return mc_tune_start(v->mct, v->stream, reps, true, &v->mc_shape);
}

int mh_marching_cubes_tuning(mh_volume* v, mh_tune_mc_result* out) {
  if (check_vol(v) || !out) return fail(MH_ERR_ARG, "null argument");
  const int rc = mc_tune_collect(v->mct, v->stream, &v->mc_shape);
  if (rc) return rc;
  *out = v->mct.result;
  if (!out->done) {
    out->candidates = kNumMcShapes;
    out->shape = v->mc_shape;
  }
  return MH_OK;
}

int mh_set_marching_cubes_shape(mh_volume* v, int shape) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (shape < 0 || shape >= kNumMcShapes) return fail(MH_ERR_ARG, "shape out of range");
  v->mc_shape = shape;
  v->mct.active = false;
  return MH_OK;
}

int mh_raycast(mh_volume* v, const mh_raycaster_params* p, const float cTw[16], float* depth_dev,
               float* vertex_dev, float* normal_dev) {
  if (check_vol(v) || !p || !cTw) return fail(MH_ERR_ARG, "null argument");
  RayConst R;
  int rc = build_ray(v, p, cTw, &R);
  if (rc) return rc;
  rc = attach_range(v, &R);
  if (rc) return rc;
  if (v->diag & MH_DIAG_RAY_BLOCKS) {
    View Vr = v->V;  // lookups record the blocks they find; then they are counted
    Vr.ray_mark = v->ray_mark;
    Vr.ray_stamp = ++v->ray_stamp;
    launch_raycast(Vr, R, depth_dev, vertex_dev, normal_dev, v->stream, v->kc(), v->ray_shape);
    launch_ray_mark_count(Vr, v->stream, v->cus * 2);
  } else {
    launch_raycast(v->V, R, depth_dev, vertex_dev, normal_dev, v->stream, v->kc(), v->ray_shape);
  }
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

static int ensure_pd(mh_volume* v) {
  View& V = v->V;
  if (V.pd_sdf0) return MH_OK;
  const size_t nvox = (size_t)v->hp.value_capacity * kBlockVoxels;
  const uint32_t np = (uint32_t)v->grid_wg256();
  Scratch sc;
  float *s0 = nullptr, *sb = nullptr, *w = nullptr, *p = nullptr;
  uint8_t* m = nullptr;
  int32_t* nbr = nullptr;
  double *part = nullptr, *en = nullptr;
  HIP_TRY(sc.get(&s0, nvox));
  HIP_TRY(sc.get(&sb, nvox));
  HIP_TRY(sc.get(&w, nvox));
  HIP_TRY(sc.get(&p, nvox * 3));
  HIP_TRY(sc.get(&m, nvox));
  HIP_TRY(sc.get(&nbr, (size_t)v->hp.value_capacity * 6));
  HIP_TRY(sc.get(&part, (size_t)np * 2));
  HIP_TRY(sc.get(&en, 4));
  HIP_TRY(hipMemsetAsync(m, 0, nvox, v->stream));
  HIP_TRY(hipMemsetAsync(s0, 0, nvox * 4, v->stream));
  HIP_TRY(hipMemsetAsync(sb, 0, nvox * 4, v->stream));
  HIP_TRY(hipMemsetAsync(w, 0, nvox * 4, v->stream));
  HIP_TRY(hipMemsetAsync(p, 0, nvox * 12, v->stream));
  // every buffer exists: publish them together
  sc.release_to(v->owned);
  V.pd_sdf0 = s0;
  V.pd_sdf_bar = sb;
  V.pd_w = w;
  V.pd_p = p;
  V.pd_mask = m;
  v->pd_nbr = nbr;
  v->pd_partials = part;
  v->pd_energy = en;
  v->pd_npartials = np;
  return MH_OK;
}

int mh_primal_dual_init(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  const int rc = ensure_pd(v);
  if (rc != MH_OK) return rc;
  launch_pd_neighbors(v->V, v->pd_nbr, v->stream, v->grid_wg256());
  launch_pd_init(v->V, v->pd_nbr, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  v->pd_work_gen = v->work_gen;
  return MH_OK;
}

int mh_primal_dual_iterate(mh_volume* v, float lambda, float sigma, float tau, double* energy_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->V.pd_sdf0) return fail(MH_ERR_STATE, "mh_primal_dual_init not called");
  // The face-neighbour table indexes the worklist as it was at init; any
  // call that has rebuilt the worklist since makes it stale.
  if (v->pd_work_gen != v->work_gen)
    return fail(MH_ERR_STATE, "worklist changed since mh_primal_dual_init");
  launch_pd_iterate(v->V, v->pd_nbr, lambda, sigma, tau, v->pd_partials, v->pd_npartials,
                    v->pd_energy, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  if (energy_host) {
    HIP_TRY(hipMemcpyAsync(energy_host, v->pd_energy, 3 * sizeof(double), hipMemcpyDeviceToHost,
                           v->stream));
    HIP_TRY(hipStreamSynchronize(v->stream));
  }
  return MH_OK;
}

// ------------------------------------------------------------------ dumps

int mh_dump_blocks(mh_volume* v, uint32_t capacity, int32_t* pos_host, float* sdf_host,
                   float* weight_host, uint32_t* color_host, float* a_host, float* b_host,
                   uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  std::vector<BlockRec> blocks;
  int rc = sorted_blocks(v, blocks);
  if (rc) return rc;
  if (count_host) *count_host = (uint32_t)blocks.size();
  const bool want = pos_host || sdf_host || weight_host || color_host || a_host || b_host;
  if (!want) return MH_OK;
  if (capacity < blocks.size()) return fail(MH_ERR_CAPACITY, "dump buffer too small");
  const uint32_t hwm = read_counters(v).block_hwm;
  if (pos_host)
    for (size_t i = 0; i < blocks.size(); ++i) {
      pos_host[3 * i] = blocks[i].pos.x;
      pos_host[3 * i + 1] = blocks[i].pos.y;
      pos_host[3 * i + 2] = blocks[i].pos.z;
    }
  if ((rc = gather_field(v, v->V.sdf, blocks, hwm, 1, sdf_host))) return rc;
  if ((rc = gather_field(v, v->V.weight, blocks, hwm, 1, weight_host))) return rc;
  if ((rc = gather_field(v, v->V.color, blocks, hwm, 1, color_host))) return rc;
  if ((rc = gather_field(v, v->V.va, blocks, hwm, 1, a_host))) return rc;
  if ((rc = gather_field(v, v->V.vb, blocks, hwm, 1, b_host))) return rc;
  return MH_OK;
}

int mh_dump_cube_indices(mh_volume* v, uint32_t capacity, uint16_t* cube_host,
                         uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  std::vector<BlockRec> blocks;
  int rc = sorted_blocks(v, blocks);
  if (rc) return rc;
  if (count_host) *count_host = (uint32_t)blocks.size();
  if (!cube_host) return MH_OK;
  if (capacity < blocks.size()) return fail(MH_ERR_CAPACITY, "dump buffer too small");
  return gather_field(v, v->V.cube, blocks, read_counters(v).block_hwm, 1, cube_host);
}

int mh_dump_primal_dual(mh_volume* v, uint32_t capacity, float* pdv_host, uint8_t* mask_host,
                        uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  std::vector<BlockRec> blocks;
  int rc = sorted_blocks(v, blocks);
  if (rc) return rc;
  if (count_host) *count_host = (uint32_t)blocks.size();
  if (!pdv_host && !mask_host) return MH_OK;
  if (!v->V.pd_sdf0) return fail(MH_ERR_STATE, "no primal-dual state");
  if (capacity < blocks.size()) return fail(MH_ERR_CAPACITY, "dump buffer too small");
  const uint32_t hwm = read_counters(v).block_hwm;
  if (pdv_host) {
    std::vector<float> f0(blocks.size() * kBlockVoxels), f1(f0.size()), f2(f0.size()),
        fp(f0.size() * 3);
    if ((rc = gather_field(v, v->V.pd_sdf0, blocks, hwm, 1, f0.data()))) return rc;
    if ((rc = gather_field(v, v->V.pd_sdf_bar, blocks, hwm, 1, f1.data()))) return rc;
    if ((rc = gather_field(v, v->V.pd_w, blocks, hwm, 1, f2.data()))) return rc;
    if ((rc = gather_field(v, v->V.pd_p, blocks, hwm, 3, fp.data()))) return rc;
    for (size_t b = 0; b < blocks.size(); ++b) {
      float* dst = pdv_host + b * 6 * kBlockVoxels;
      for (int i = 0; i < kBlockVoxels; ++i) {
        const size_t s = b * kBlockVoxels + i;
        dst[i] = f0[s];
        dst[kBlockVoxels + i] = f1[s];
        dst[2 * kBlockVoxels + i] = f2[s];
        dst[3 * kBlockVoxels + i] = fp[3 * s];
        dst[4 * kBlockVoxels + i] = fp[3 * s + 1];
        dst[5 * kBlockVoxels + i] = fp[3 * s + 2];
      }
    }
  }
  if (mask_host && (rc = gather_field(v, v->V.pd_mask, blocks, hwm, 1, mask_host))) return rc;
  return MH_OK;
}

int mh_dump_hash(mh_volume* v, uint32_t capacity, int32_t* rec_host, uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  std::vector<Entry> e(v->hp.entry_count);
  HIP_TRY(hipMemcpy(e.data(), v->V.entries, e.size() * sizeof(Entry), hipMemcpyDeviceToHost));
  uint32_t n = 0;
  for (size_t i = 0; i < e.size(); ++i) {
    if (e[i].key == kEmptyKey || e[i].ptr < 0) continue;
    if (rec_host) {
      if (n >= capacity) return fail(MH_ERR_CAPACITY, "dump buffer too small");
      const i3 p = unpack_key(e[i].key);
      const uint32_t home = hash_bucket(p, v->hp.bucket_count);
      int32_t* r = rec_host + 6 * (size_t)n;
      r[0] = (int32_t)i;
      r[1] = (int32_t)home;
      r[2] = (i / v->hp.bucket_size) == home ? 1 : 0;
      r[3] = p.x;
      r[4] = p.y;
      r[5] = p.z;
    }
    ++n;
  }
  if (count_host) *count_host = n;
  return MH_OK;
}

int mh_dump_candidates(mh_volume* v, uint32_t capacity, int32_t* pos_host, uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  const Counters c = read_counters(v);
  std::vector<int32_t> cand(c.cand_count);
  if (c.cand_count)
    HIP_TRY(hipMemcpy(cand.data(), v->V.cand, cand.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<uint64_t> keys(c.block_hwm);
  if (c.block_hwm)
    HIP_TRY(hipMemcpy(keys.data(), v->V.block_keys, keys.size() * sizeof(uint64_t),
                      hipMemcpyDeviceToHost));
  std::vector<i3> pos;
  for (int32_t p : cand) pos.push_back(unpack_key(keys[p]));
  std::sort(pos.begin(), pos.end(), pos_less);
  if (count_host) *count_host = (uint32_t)pos.size();
  if (!pos_host) return MH_OK;
  if (capacity < pos.size()) return fail(MH_ERR_CAPACITY, "dump buffer too small");
  for (size_t i = 0; i < pos.size(); ++i) {
    pos_host[3 * i] = pos[i].x;
    pos_host[3 * i + 1] = pos[i].y;
    pos_host[3 * i + 2] = pos[i].z;
  }
  return MH_OK;
}

int mh_dump_candidate_flags(mh_volume* v, uint32_t capacity, int32_t* rec_host, uint32_t* count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  const Counters c = read_counters(v);
  std::vector<int32_t> cand(c.cand_count);
  std::vector<uint8_t> flag(c.cand_count);
  if (c.cand_count) {
    HIP_TRY(hipMemcpy(cand.data(), v->V.cand, cand.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(flag.data(), v->V.cand_flag, flag.size(), hipMemcpyDeviceToHost));
  }
  std::vector<uint64_t> keys(c.block_hwm);
  if (c.block_hwm)
    HIP_TRY(hipMemcpy(keys.data(), v->V.block_keys, keys.size() * sizeof(uint64_t),
                      hipMemcpyDeviceToHost));
  std::vector<std::array<int32_t, 4>> rec;
  for (size_t k = 0; k < cand.size(); ++k) {
    const i3 p = unpack_key(keys[cand[k]]);
    rec.push_back({p.x, p.y, p.z, (int32_t)flag[k]});
  }
  std::sort(rec.begin(), rec.end());
  if (count_host) *count_host = (uint32_t)rec.size();
  if (!rec_host) return MH_OK;
  if (capacity < rec.size()) return fail(MH_ERR_CAPACITY, "dump buffer too small");
  std::memcpy(rec_host, rec.data(), rec.size() * sizeof(rec[0]));
  return MH_OK;
}

int mh_dump_mesh(mh_volume* v, uint32_t tri_capacity, int32_t* tri_host, int32_t* tri_cube_host,
                 uint32_t vtx_capacity, int32_t* vtx_key_host, float* vtx_val_host,
                 int32_t* vtx_ref_host, uint32_t* tri_count_host, uint32_t* vtx_count_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(v->stream));
  std::vector<BlockRec> blocks;
  int rc = sorted_blocks(v, blocks);
  if (rc) return rc;
  const uint32_t hwm = read_counters(v).block_hwm;
  const size_t nvox = (size_t)hwm * kBlockVoxels;
  std::vector<int32_t> vptr(nvox * 3);
  std::vector<uint16_t> cubes(nvox);
  if (nvox) {
    HIP_TRY(hipMemcpy(cubes.data(), v->V.cube, nvox * 2, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(vptr.data(), v->V.vptr, vptr.size() * 4, hipMemcpyDeviceToHost));
  }
  std::vector<Vertex> verts(v->mp.max_vertex_count);
  HIP_TRY(hipMemcpy(verts.data(), v->V.vertices, verts.size() * sizeof(Vertex),
                    hipMemcpyDeviceToHost));
  // vertex pointer -> canonical key (owner voxel, axis)
  struct VKey {
    int32_t k[4];
    int32_t vp;
  };
  std::vector<VKey> vkeys;
  for (const BlockRec& b : blocks) {
    for (int i = 0; i < kBlockVoxels; ++i) {
      const i3 o = devectorize((uint32_t)i);
      for (int ax = 0; ax < 3; ++ax) {
        const int32_t vp = vptr[((size_t)b.ptr * kBlockVoxels + i) * 3 + ax];
        if (vp < 0) continue;
        vkeys.push_back(VKey{{b.pos.x * 8 + o.x, b.pos.y * 8 + o.y, b.pos.z * 8 + o.z, ax}, vp});
      }
    }
  }
  std::sort(vkeys.begin(), vkeys.end(), [](const VKey& a, const VKey& b) {
    return std::lexicographical_compare(a.k, a.k + 4, b.k, b.k + 4);
  });
  if (vtx_count_host) *vtx_count_host = (uint32_t)vkeys.size();
  if (vtx_key_host || vtx_val_host || vtx_ref_host) {
    if (vtx_capacity < vkeys.size()) return fail(MH_ERR_CAPACITY, "vertex buffer too small");
    for (size_t i = 0; i < vkeys.size(); ++i) {
      const Vertex& vt = verts[vkeys[i].vp];
      if (vtx_key_host) std::memcpy(vtx_key_host + 4 * i, vkeys[i].k, sizeof(vkeys[i].k));
      if (vtx_val_host) {
        float* d = vtx_val_host + 10 * i;
        d[0] = vt.px; d[1] = vt.py; d[2] = vt.pz;
        d[3] = vt.nx; d[4] = vt.ny; d[5] = vt.nz;
        d[6] = vt.cr; d[7] = vt.cg; d[8] = vt.cb;
        d[9] = vt.radius;
      }
      if (vtx_ref_host) {
        // reference count from the four cube configurations around the edge
        const int32_t* k = vkeys[i].k;
        const int ax = k[3], b1 = ax == 0 ? 1 : 0, b2 = ax == 2 ? 1 : 2;
        int refs = 0;
        for (int q = 0; q < 4; ++q) {
          int d[3] = {0, 0, 0};
          d[b1] = q & 1;
          d[b2] = q >> 1;
          const i3 c{k[0] - d[0], k[1] - d[1], k[2] - d[2]};
          const i3 cb = voxel_to_block(c);
          auto it = std::lower_bound(blocks.begin(), blocks.end(), cb,
                                     [](const BlockRec& r, const i3& p) { return pos_less(r.pos, p); });
          if (it == blocks.end() || !(it->pos.x == cb.x && it->pos.y == cb.y && it->pos.z == cb.z))
            continue;
          const uint32_t li = (uint32_t)((c.x - cb.x * 8) + 8 * (c.y - cb.y * 8) + 64 * (c.z - cb.z * 8));
          refs += edge_refs(kHostMc, cubes[(size_t)it->ptr * kBlockVoxels + li] & 255, d[0], d[1], d[2], ax);
        }
        vtx_ref_host[i] = refs;
      }
    }
  }
  // triangles in (cube voxel, slot) order: blocks are sorted by position and
  // voxels within a block are visited in x-fastest order, so sort afterwards.
  struct TRec {
    int32_t cube[4];
    int32_t v[12];
  };
  // The live triangles of a cube are its configuration's table row over the
  // vertices of its edges; a vertex is named by its owner voxel and axis.
  std::vector<TRec> trecs;
  auto slot_of = [&](const i3& vox, int ax) -> int32_t {
    const i3 vb = voxel_to_block(vox);
    auto it = std::lower_bound(blocks.begin(), blocks.end(), vb,
                               [](const BlockRec& r, const i3& p) { return pos_less(r.pos, p); });
    if (it == blocks.end() || !(it->pos.x == vb.x && it->pos.y == vb.y && it->pos.z == vb.z))
      return kFreePtr;
    const uint32_t li = (uint32_t)((vox.x - vb.x * 8) + 8 * (vox.y - vb.y * 8) + 64 * (vox.z - vb.z * 8));
    return vptr[((size_t)it->ptr * kBlockVoxels + li) * 3 + ax];
  };
  for (const BlockRec& b : blocks) {
    for (int i = 0; i < kBlockVoxels; ++i) {
      const int cube = cubes[(size_t)b.ptr * kBlockVoxels + i] & 255;
      const int nt = kHostMc.ntri[cube];
      if (cube == 0 || cube == 255 || nt == 0) continue;
      const i3 o = devectorize((uint32_t)i);
      const i3 c{b.pos.x * 8 + o.x, b.pos.y * 8 + o.y, b.pos.z * 8 + o.z};
      for (int t = 0; t < nt; ++t) {
        TRec r{{c.x, c.y, c.z, t}, {}};
        for (int k = 0; k < 3; ++k) {
          const int e = kHostMc.tri[cube][3 * t + k];
          const int ax = edge_axis(e);
          const i3 ov{c.x + edge_owner_off(e, 0), c.y + edge_owner_off(e, 1), c.z + edge_owner_off(e, 2)};
          const bool live = slot_of(ov, ax) >= 0;
          r.v[4 * k + 0] = live ? ov.x : INT32_MIN;
          r.v[4 * k + 1] = live ? ov.y : INT32_MIN;
          r.v[4 * k + 2] = live ? ov.z : INT32_MIN;
          r.v[4 * k + 3] = ax;
        }
        trecs.push_back(r);
      }
    }
  }
  std::sort(trecs.begin(), trecs.end(), [](const TRec& a, const TRec& b) {
    return std::lexicographical_compare(a.cube, a.cube + 4, b.cube, b.cube + 4);
  });
  if (tri_count_host) *tri_count_host = (uint32_t)trecs.size();
  if (tri_host || tri_cube_host) {
    if (tri_capacity < trecs.size()) return fail(MH_ERR_CAPACITY, "triangle buffer too small");
    for (size_t i = 0; i < trecs.size(); ++i) {
      if (tri_host) std::memcpy(tri_host + 12 * i, trecs[i].v, sizeof(trecs[i].v));
      if (tri_cube_host) std::memcpy(tri_cube_host + 4 * i, trecs[i].cube, sizeof(trecs[i].cube));
    }
  }
  return MH_OK;
}

int mh_load_blocks(mh_volume* v, uint32_t n, const int32_t* pos_host, const float* sdf_host,
                   const float* weight_host, const uint32_t* color_host, const float* a_host,
                   const float* b_host) {
  if (check_vol(v) || !pos_host || !sdf_host || !weight_host) return fail(MH_ERR_ARG, "null argument");
  if ((a_host || b_host) && !v->psdf) return fail(MH_ERR_STATE, "volume has no PSDF fields");
  if (n == 0) return MH_OK;
  if (n > v->hp.value_capacity) return fail(MH_ERR_CAPACITY, "more blocks than value_capacity");
  std::vector<uint64_t> keys(n);
  for (uint32_t i = 0; i < n; ++i) {
    const i3 p{pos_host[3 * i], pos_host[3 * i + 1], pos_host[3 * i + 2]};
    if (!coord_ok(p)) return fail(MH_ERR_ARG, "block coordinate out of range");
    keys[i] = pack_key(p);
  }
  const size_t nv = (size_t)n * kBlockVoxels;
  Scratch sc;  // temporaries, freed on every return path
  uint64_t* dkeys = nullptr;
  float *dsdf = nullptr, *dw = nullptr, *da = nullptr, *db = nullptr;
  uint32_t* dc = nullptr;
  HIP_TRY(sc.get(&dkeys, n));
  HIP_TRY(sc.get(&dsdf, nv));
  HIP_TRY(sc.get(&dw, nv));
  HIP_TRY(hipMemcpyAsync(dkeys, keys.data(), n * 8, hipMemcpyHostToDevice, v->stream));
  HIP_TRY(hipMemcpyAsync(dsdf, sdf_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  HIP_TRY(hipMemcpyAsync(dw, weight_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  if (color_host) {
    HIP_TRY(sc.get(&dc, nv));
    HIP_TRY(hipMemcpyAsync(dc, color_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  }
  if (a_host) {
    HIP_TRY(sc.get(&da, nv));
    HIP_TRY(hipMemcpyAsync(da, a_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  }
  if (b_host) {
    HIP_TRY(sc.get(&db, nv));
    HIP_TRY(hipMemcpyAsync(db, b_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  }
  // the copies read pageable host memory and the kernels the device copies:
  // the stream is drained before the temporaries go
  launch_load_insert(v->V, dkeys, n, v->stream);
  launch_load_scatter(v->V, dkeys, n, dsdf, dw, dc, da, db, v->stream);
  HIP_TRY(hipGetLastError());
  const hipError_t e = hipStreamSynchronize(v->stream);
  if (e != hipSuccess) {
    sc.ptrs.clear();  // in-flight work may still use them: leak rather than free
    return fail(MH_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  }
  return MH_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuning
// Every integration tiling (kIntegrateTilings) x persistent-grid size is
// timed on the current worklist.  Integration is stateful, so the worklist
// blocks' fused state and the counters are saved once and restored after
// each trial: tuning changes no output, and every tiling gives bit-identical
// results (each voxel is updated by one lane with the same arithmetic).
namespace {

constexpr int kGridPerCu[] = {2, 4, 8, 16, 32};
static_assert(sizeof(kGridPerCu) / sizeof(kGridPerCu[0]) == MH_NUM_TUNE_GRIDS, "grid list");
static_assert(kNumIntegrateTilings == MH_NUM_TILINGS, "tiling list");
static_assert(kNumRayShapes == MH_NUM_RAY_SHAPES, "raycast shape list");
static_assert(kNumMcShapes == MH_NUM_MC_SHAPES, "marching-cubes shape list");
constexpr int kDefaultGridPerCu = 8;

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

template <class Run, class Restore>
int tune_loop(hipStream_t s, int cus, uint32_t reps, Run run, Restore restore, mh_tune_result* out,
              int* best_t, int* best_g) {
  Events ev;
  HIP_TRY(hipEventCreate(&ev.a));
  HIP_TRY(hipEventCreate(&ev.b));
  float best = INFINITY, def = INFINITY;
  int cands = 0;
  *best_t = kDefaultIntegrateTiling;
  *best_g = kDefaultGridPerCu;
  for (int t = 0; t < kNumIntegrateTilings; ++t) {
    for (int g : kGridPerCu) {
      float tmin = INFINITY;
      for (uint32_t r = 0; r < reps; ++r) {
        HIP_TRY(hipEventRecord(ev.a, s));
        run(t, cus * g);
        HIP_TRY(hipEventRecord(ev.b, s));
        HIP_TRY(hipGetLastError());
        const int rc = restore();
        if (rc) return rc;
        HIP_TRY(hipEventSynchronize(ev.b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
        tmin = fminf(tmin, ms);
      }
      ++cands;
      if (tmin < best) {
        best = tmin;
        *best_t = t;
        *best_g = g;
      }
      if (t == kDefaultIntegrateTiling && g == kDefaultGridPerCu) def = tmin;
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  if (out) {
    const Tiling tl = kIntegrateTilings[*best_t];
    out->tiling = *best_t;
    out->workgroup_size = tl.nt;
    out->blocks_per_workgroup = tl.bpw;
    out->voxels_per_lane = kBlockVoxels * tl.bpw / tl.nt;
    out->early_loads = tl.early ? 1 : 0;
    out->workgroups_per_cu = *best_g;
    out->candidates = cands;
    out->best_ms = best;
    out->default_ms = def;
    out->ray_shape = -1;
  }
  return MH_OK;
}

// Times run(shape) for every raycast tile shape (reps trials each, restore()
// after each) and returns the fastest; the kernels write only outputs and
// work counters, which restore() puts back.
template <class Run, class Restore>
int tune_ray_shapes(hipStream_t s, uint32_t reps, Run run, Restore restore, mh_tune_shape_result* out,
                    int* best_shape) {
  Events ev;
  HIP_TRY(hipEventCreate(&ev.a));
  HIP_TRY(hipEventCreate(&ev.b));
  float best = INFINITY, def = INFINITY;
  *best_shape = 0;
  for (int k = 0; k < kNumRayShapes; ++k) {
    float tmin = INFINITY;
    for (uint32_t r = 0; r < reps; ++r) {
      HIP_TRY(hipEventRecord(ev.a, s));
      run(k);
      HIP_TRY(hipEventRecord(ev.b, s));
      HIP_TRY(hipGetLastError());
      const int rc = restore();
      if (rc) return rc;
      HIP_TRY(hipEventSynchronize(ev.b));
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
      tmin = fminf(tmin, ms);
    }
    // shape 0 is measured first; another replaces the best only if it is
    // at least 3 % faster (a single frame's timing favours no shape by less)
    if (k == 0 || tmin < 0.97f * best) {
      best = tmin;
      *best_shape = k;
    }
    if (k == 0) def = tmin;
  }
  HIP_TRY(hipStreamSynchronize(s));
  if (out) {
    out->shape = *best_shape;
    out->tile_width = kRayTileWidths[*best_shape];
    out->tile_height = 64 / kRayTileWidths[*best_shape];
    out->candidates = kNumRayShapes;
    out->best_ms = best;
    out->default_ms = def;
  }
  return MH_OK;
}

int check_tiling(int tiling, int grid_per_cu) {
  if (tiling < 0 || tiling >= kNumIntegrateTilings) return fail(MH_ERR_ARG, "tiling out of range");
  if (grid_per_cu < 1 || grid_per_cu > 64) return fail(MH_ERR_ARG, "grid_per_cu out of range");
  return MH_OK;
}

}  // namespace

// ------------------------------------------------------------------ batches
// S independent scenes of one image size.  Each scene is an ordinary volume;
// a device table holds every scene's View, MarchingCubes and RayCaster
// constants for the current frame, so each stage is one launch over all
// scenes and the per-frame chain can be captured once as a HIP graph and
// replayed (every count it needs lives on the device).
struct mh_batch {
  std::vector<mh_volume*> vols;
  uint32_t n = 0, W = 0, H = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // the raycast branch of a frame (it reads only the fused TSDF and the
  // table, so it runs beside the meshing pass): forked from and joined back
  // into stream with the two events
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int cus = 256;
  SceneFrame* d_frames = nullptr;
  std::vector<SceneFrame> h_frames;
  static constexpr int kRing = 4;  // pinned staging slots for the frame table
  SceneFrame* pinned[kRing] = {};
  hipEvent_t ev[kRing] = {};
  int ring = 0;
  uint32_t* d_work = nullptr;
  uint32_t* d_work_count = nullptr;
  uint32_t* d_dyn = nullptr;       // marching cubes' on-demand item counter
  uint16_t* d_raw = nullptr;       // staging for host depth uploads
  uint32_t* d_range = nullptr;     // raycast tile ranges, ray_range_words(W, H) per scene
  int32_t* d_nbt = nullptr;        // marching cubes: 27 neighbours per concatenated-worklist item
  const uint16_t* raw = nullptr;   // depth of the current frame (n*W*H)
  bool have_frame = false;
  std::vector<mh_raycaster_params> rp;
  float *out_depth = nullptr, *out_vertex = nullptr, *out_normal = nullptr;
  bool have_ray = false;
  int tiling = kDefaultIntegrateTiling;
  int grid_per_cu = kDefaultGridPerCu;
  int ray_shape = 0;
  int mc_shape = 0;
  McTune mct;  // online marching-cubes tuning over graph replays
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_flags = 0;

  BatchDev dev() const { return BatchDev{d_frames, d_work, d_work_count, d_dyn, n, W, H, 0}; }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

void batch_free(mh_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_graph();
  mc_tune_free(b->mct);
  if (b->side) (void)hipStreamSynchronize(b->side);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  for (int k = 0; k < mh_batch::kRing; ++k) {
    if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
    if (b->pinned[k]) (void)hipHostFree(b->pinned[k]);
  }
  for (void* p : {(void*)b->d_frames, (void*)b->d_work, (void*)b->d_work_count, (void*)b->d_dyn, (void*)b->d_raw,
                  (void*)b->d_range, (void*)b->d_nbt})
    if (p) (void)hipFree(p);
  for (mh_volume* v : b->vols) mh_volume_destroy(v);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// The per-frame chain: convert -> alloc -> collect visible -> TSDF ->
// [marching cubes] -> [raycast], one launch per stage over all scenes.
// Whether a frame's chain runs the raycast beside the meshing pass (below).
bool batch_overlaps(const mh_batch* b, uint32_t flags) {
  const size_t ray_waves = (size_t)b->n * (((size_t)b->W * b->H + 63) / 64);
  const size_t resident = (size_t)kRayWavesPerSimd * 4 * (size_t)b->cus;
  return (flags & MH_RUN_MESH) && (flags & MH_RUN_RAYCAST) && ray_waves <= resident;
}

int batch_chain(mh_batch* b, uint32_t flags, bool integrate) {
  const BatchDev B = b->dev();
  launch_batch_alloc(B, b->stream);  // converts each scene's raw depth (View.raw) as it goes
  // the list pass cleared the counts; with meshing the collect pass also
  // stamps the blocks for the marching-cubes pass (no marking launch)
  launch_batch_collect(B, 1, b->stream, b->cus * 4, true, (flags & MH_RUN_MESH) != 0);
  if (!integrate) return MH_OK;
  launch_batch_integrate_tsdf(B, b->stream, b->cus * b->grid_per_cu, b->tiling);
  // The raycast needs only the fused TSDF and the table, the meshing pass
  // writes neither: when the raycast's waves (one per 64 pixels) fit on the
  // chip in one round, it runs on the side stream beside meshing, so each
  // fills the other's tail (the slowest tiles end ~3x after the median).
  // With more waves than that there is no idle tail to fill, and sharing
  // the chip only slowed both (64 scenes of 320 x 240: 14 %).  Captured, the
  // fork and join become graph edges.
  if (batch_overlaps(b, flags)) {
    HIP_TRY(hipEventRecord(b->fork, b->stream));
    HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0));
    launch_batch_raycast(B, b->out_depth, b->out_vertex, b->out_normal, b->side, b->ray_shape);
    launch_batch_marching_cubes(B, b->d_nbt, b->stream, b->cus, kMcShapes[b->mc_shape], true);
    HIP_TRY(hipEventRecord(b->join, b->side));
    HIP_TRY(hipStreamWaitEvent(b->stream, b->join, 0));
    return MH_OK;
  }
  if (flags & MH_RUN_MESH) launch_batch_marching_cubes(B, b->d_nbt, b->stream, b->cus, kMcShapes[b->mc_shape], true);
  if (flags & MH_RUN_RAYCAST) launch_batch_raycast(B, b->out_depth, b->out_vertex, b->out_normal, b->stream, b->ray_shape);
  return MH_OK;
}

int batch_tune(mh_batch* b, uint32_t reps, mh_tune_result* out) {
  HIP_TRY(hipStreamSynchronize(b->stream));
  std::vector<size_t> off(b->n + 1, 0);
  for (uint32_t s = 0; s < b->n; ++s) off[s + 1] = off[s] + read_counters(b->vols[s]).cand_count;
  const size_t total = off[b->n];
  if (total == 0) return fail(MH_ERR_EMPTY, "empty worklists");
  Scratch sc;
  float *ss = nullptr, *sw = nullptr;
  uint32_t* scol = nullptr;
  Counters* sctr = nullptr;
  HIP_TRY(sc.get(&ss, total * kBlockVoxels));
  HIP_TRY(sc.get(&sw, total * kBlockVoxels));
  HIP_TRY(sc.get(&scol, total * kBlockVoxels));
  HIP_TRY(sc.get(&sctr, b->n));
  for (uint32_t s = 0; s < b->n; ++s) {
    const size_t o = off[s] * kBlockVoxels;
    launch_state_save(b->vols[s]->V, ss + o, sw + o, scol + o, b->stream, b->cus);
    HIP_TRY(hipMemcpyAsync(sctr + s, b->vols[s]->V.ctr, sizeof(Counters), hipMemcpyDeviceToDevice,
                           b->stream));
  }
  const BatchDev B = b->dev();
  auto run = [&](int t, int g) { launch_batch_integrate_tsdf(B, b->stream, g, t); };
  auto restore = [&]() -> int {
    for (uint32_t s = 0; s < b->n; ++s) {
      const size_t o = off[s] * kBlockVoxels;
      launch_state_restore(b->vols[s]->V, ss + o, sw + o, scol + o, b->stream, b->cus);
      HIP_TRY(hipMemcpyAsync(b->vols[s]->V.ctr, sctr + s, sizeof(Counters),
                             hipMemcpyDeviceToDevice, b->stream));
    }
    return MH_OK;
  };
  int bt = 0, bg = 0;
  const int rc = tune_loop(b->stream, b->cus, reps, run, restore, out, &bt, &bg);
  if (rc) return rc;
  b->tiling = bt;
  b->grid_per_cu = bg;
  return MH_OK;
}

}  // namespace

extern "C" {

int mh_tune_raycast(mh_volume* v, const mh_raycaster_params* p, const float cTw[16], uint32_t reps,
                    mh_tune_shape_result* out) {
  if (check_vol(v) || !p || !cTw) return fail(MH_ERR_ARG, "null argument");
  if (reps == 0) reps = 3;
  RayConst R;
  int rc = build_ray(v, p, cTw, &R);
  if (rc) return rc;
  rc = attach_range(v, &R);
  if (rc) return rc;
  const size_t npix = (size_t)p->width * p->height;
  Scratch sc;
  float *d = nullptr, *n = nullptr;
  Counters* sctr = nullptr;
  HIP_TRY(sc.get(&d, npix));
  HIP_TRY(sc.get(&n, npix * 3));
  HIP_TRY(sc.get(&sctr, 1));
  if (R.range) launch_ray_range_one(v->V, R, v->stream);
  HIP_TRY(hipMemcpyAsync(sctr, v->V.ctr, sizeof(Counters), hipMemcpyDeviceToDevice, v->stream));
  auto run = [&](int k) { launch_raycast_march(v->V, R, d, nullptr, n, v->stream, k); };
  auto restore = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(v->V.ctr, sctr, sizeof(Counters), hipMemcpyDeviceToDevice, v->stream));
    return MH_OK;
  };
  int best = 0;
  rc = tune_ray_shapes(v->stream, reps, run, restore, out, &best);
  if (rc) return rc;
  v->ray_shape = best;
  return MH_OK;
}

int mh_set_raycast_shape(mh_volume* v, int shape) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (shape < 0 || shape >= kNumRayShapes) return fail(MH_ERR_ARG, "shape out of range");
  v->ray_shape = shape;
  return MH_OK;
}

int mh_set_integrate_tiling(mh_volume* v, int tiling, int grid_per_cu) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (v->psdf) {
    if (tiling < 0 || tiling >= kNumPsdfTilings) return fail(MH_ERR_ARG, "PSDF tiling out of range");
    if (grid_per_cu < 1 || grid_per_cu > 64) return fail(MH_ERR_ARG, "grid_per_cu out of range");
    v->psdf_tiling = tiling;
    v->psdf_grid_per_cu = grid_per_cu;
    return MH_OK;
  }
  const int rc = check_tiling(tiling, grid_per_cu);
  if (rc) return rc;
  v->tiling = tiling;
  v->grid_per_cu = grid_per_cu;
  return MH_OK;
}

// PSDF fusion tuning: workgroup sizes kPsdfLanes x workgroups per CU, the
// worklist's sdf / weight (1/sigma^2) / a / b and the counters restored
// after every trial.
static int tune_psdf(mh_volume* v, uint32_t n, uint32_t reps, mh_tune_result* out) {
  Scratch sc;
  float *ss = nullptr, *sw = nullptr, *sa = nullptr, *sb = nullptr;
  uint32_t *scol = nullptr, *scol2 = nullptr;
  Counters* sctr = nullptr;
  const size_t nv = (size_t)n * kBlockVoxels;
  HIP_TRY(sc.get(&ss, nv));
  HIP_TRY(sc.get(&sw, nv));
  HIP_TRY(sc.get(&sa, nv));
  HIP_TRY(sc.get(&sb, nv));
  HIP_TRY(sc.get(&scol, nv));
  HIP_TRY(sc.get(&scol2, nv));
  HIP_TRY(sc.get(&sctr, 1));
  View Vab = v->V;  // the Beta parameters through the same copy kernel
  Vab.sdf = v->V.va;
  Vab.weight = v->V.vb;
  launch_state_save(v->V, ss, sw, scol, v->stream, v->cus * 4);
  launch_state_save(Vab, sa, sb, scol2, v->stream, v->cus * 4);
  HIP_TRY(hipMemcpyAsync(sctr, v->V.ctr, sizeof(Counters), hipMemcpyDeviceToDevice, v->stream));
  auto restore = [&]() -> int {
    launch_state_restore(v->V, ss, sw, scol, v->stream, v->cus * 4);
    launch_state_restore(Vab, sa, sb, scol2, v->stream, v->cus * 4);
    HIP_TRY(hipMemcpyAsync(v->V.ctr, sctr, sizeof(Counters), hipMemcpyDeviceToDevice, v->stream));
    return MH_OK;
  };
  Events ev;
  HIP_TRY(hipEventCreate(&ev.a));
  HIP_TRY(hipEventCreate(&ev.b));
  float best = INFINITY, def = INFINITY;
  int bt = kDefaultPsdfTiling, bg = kDefaultGridPerCu, cands = 0;
  for (int t = 0; t < kNumPsdfTilings; ++t)
    for (int g : kGridPerCu) {
      float tmin = INFINITY;
      for (uint32_t r = 0; r < reps; ++r) {
        HIP_TRY(hipEventRecord(ev.a, v->stream));
        launch_integrate_psdf(v->V, v->stream, v->cus * g, kPsdfLanes[t]);
        HIP_TRY(hipEventRecord(ev.b, v->stream));
        HIP_TRY(hipGetLastError());
        const int rc = restore();
        if (rc) return rc;
        HIP_TRY(hipEventSynchronize(ev.b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
        tmin = fminf(tmin, ms);
      }
      ++cands;
      if (tmin < best) {
        best = tmin;
        bt = t;
        bg = g;
      }
      if (t == kDefaultPsdfTiling && g == kDefaultGridPerCu) def = tmin;
    }
  HIP_TRY(hipStreamSynchronize(v->stream));
  v->psdf_tiling = bt;
  v->psdf_grid_per_cu = bg;
  if (out) {
    out->tiling = bt;
    out->workgroup_size = kPsdfLanes[bt];
    out->blocks_per_workgroup = 1;
    out->voxels_per_lane = kBlockVoxels / kPsdfLanes[bt];
    out->workgroups_per_cu = bg;
    out->candidates = cands;
    out->best_ms = best;
    out->default_ms = def;
    out->early_loads = 0;
    out->ray_shape = -1;
  }
  return MH_OK;
}

int mh_tune_integrate(mh_volume* v, uint32_t reps, mh_tune_result* out) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  flush_frame(v);
  if (reps == 0) reps = 3;
  const uint32_t n = read_counters(v).cand_count;
  if (n == 0) return fail(MH_ERR_EMPTY, "empty worklist");
  if (v->psdf) return tune_psdf(v, n, reps, out);
  Scratch sc;
  float *ss = nullptr, *sw = nullptr;
  uint32_t* scol = nullptr;
  Counters* sctr = nullptr;
  HIP_TRY(sc.get(&ss, (size_t)n * kBlockVoxels));
  HIP_TRY(sc.get(&sw, (size_t)n * kBlockVoxels));
  HIP_TRY(sc.get(&scol, (size_t)n * kBlockVoxels));
  HIP_TRY(sc.get(&sctr, 1));
  launch_state_save(v->V, ss, sw, scol, v->stream, v->cus * 4);
  HIP_TRY(hipMemcpyAsync(sctr, v->V.ctr, sizeof(Counters), hipMemcpyDeviceToDevice, v->stream));
  auto run = [&](int t, int g) { launch_integrate_tsdf(v->V, v->stream, g, t); };
  auto restore = [&]() -> int {
    launch_state_restore(v->V, ss, sw, scol, v->stream, v->cus * 4);
    HIP_TRY(hipMemcpyAsync(v->V.ctr, sctr, sizeof(Counters), hipMemcpyDeviceToDevice, v->stream));
    return MH_OK;
  };
  int bt = 0, bg = 0;
  const int rc = tune_loop(v->stream, v->cus, reps, run, restore, out, &bt, &bg);
  if (rc) return rc;
  v->tiling = bt;
  v->grid_per_cu = bg;
  return MH_OK;
}

int mh_frame_get_pose(mh_volume* v, float wTc[16], float cTw[16]) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (wTc) std::memcpy(wTc, v->V.wTc.m, sizeof(float) * 16);
  if (cTw) std::memcpy(cTw, v->V.cTw.m, sizeof(float) * 16);
  return MH_OK;
}

int mh_batch_create(uint32_t n_scenes, const mh_hash_params* hash, const mh_volume_params* vol,
                    const mh_mesh_params* mesh, const mh_sensor_params* sensor, mh_batch** out) {
  if (!hash || !vol || !mesh || !sensor || !out) return fail(MH_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_scenes == 0 || n_scenes > kMaxScenes) return fail(MH_ERR_ARG, "n_scenes out of range");
  for (uint32_t s = 0; s < n_scenes; ++s) {
    if (sensor[s].width != sensor[0].width || sensor[s].height != sensor[0].height)
      return fail(MH_ERR_ARG, "all scenes of a batch need the same image size");
    if (hash[s].value_capacity > kMaxBatchCapacity)
      return fail(MH_ERR_ARG, "value_capacity above 2^24 in a batch");
  }
  mh_batch* b = new mh_batch();
  b->n = n_scenes;
  b->W = sensor[0].width;
  b->H = sensor[0].height;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
    b->cus = cus;
  auto bail = [&](int rc) {
    batch_free(b);
    return rc;
  };
  if (hipStreamCreate(&b->stream) != hipSuccess) {
    b->stream = nullptr;
    return bail(fail(MH_ERR_HIP, "hipStreamCreate failed"));
  }
  b->own_stream = true;
  if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess) {
    b->side = nullptr;
    return bail(fail(MH_ERR_HIP, "hipStreamCreate failed"));
  }
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
    return bail(fail(MH_ERR_HIP, "hipEventCreate failed"));
  size_t work_cap = 0;
  for (uint32_t s = 0; s < n_scenes; ++s) {
    mh_volume* v = nullptr;
    const int rc = mh_volume_create(&hash[s], &vol[s], &mesh[s], &sensor[s], 0, &v);
    if (rc) return bail(rc);
    v->stream = b->stream;
    b->vols.push_back(v);
    work_cap += hash[s].value_capacity;
  }
  const size_t npix = (size_t)b->W * b->H * n_scenes;
  if (dalloc(&b->d_frames, n_scenes) != hipSuccess || dalloc(&b->d_work, work_cap) != hipSuccess ||
      dalloc(&b->d_work_count, 1) != hipSuccess || dalloc(&b->d_dyn, 1) != hipSuccess ||
      dalloc(&b->d_raw, npix) != hipSuccess ||
      dalloc(&b->d_nbt, work_cap * 27) != hipSuccess)
    return bail(fail(MH_ERR_HIP, "device allocation failed"));
  for (int k = 0; k < mh_batch::kRing; ++k) {
    if (hipHostMalloc((void**)&b->pinned[k], sizeof(SceneFrame) * n_scenes, hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(b->ev[k], b->stream) != hipSuccess)
      return bail(fail(MH_ERR_HIP, "pinned staging allocation failed"));
  }
  b->h_frames.resize(n_scenes);
  b->rp.resize(n_scenes);
  *out = b;
  return MH_OK;
}

int mh_batch_destroy(mh_batch* b) {
  if (b) batch_free(b);
  return MH_OK;
}

int mh_batch_set_stream(mh_batch* b, void* hip_stream) {
  if (!b || !hip_stream) return fail(MH_ERR_ARG, "null argument (a batch needs a non-null stream)");
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->own_stream) HIP_TRY(hipStreamDestroy(b->stream));
  b->own_stream = false;
  b->stream = (hipStream_t)hip_stream;
  for (mh_volume* v : b->vols) v->stream = b->stream;
  b->drop_graph();
  return MH_OK;
}

int mh_batch_volume(mh_batch* b, uint32_t scene, mh_volume** out) {
  if (!b || !out || scene >= b->n) return fail(MH_ERR_ARG, "bad batch, scene or output");
  *out = b->vols[scene];
  return MH_OK;
}

int mh_batch_frame_u16(mh_batch* b, const uint16_t* depth, int on_device, const float* wTc) {
  if (!b || !depth || !wTc) return fail(MH_ERR_ARG, "null argument");
  const size_t npix = (size_t)b->W * b->H * b->n;
  if (on_device) {
    b->raw = depth;
  } else {
    HIP_TRY(hipMemcpyAsync(b->d_raw, depth, npix * sizeof(uint16_t), hipMemcpyHostToDevice, b->stream));
    b->raw = b->d_raw;
  }
  for (uint32_t s = 0; s < b->n; ++s) {
    set_pose(b->vols[s], wTc + 16 * s);
    b->vols[s]->V.has_color = 0;
  }
  b->have_frame = true;
  return MH_OK;
}

int mh_batch_set_raycast(mh_batch* b, const mh_raycaster_params* p, float* depth_dev,
                         float* vertex_dev, float* normal_dev) {
  if (!b || !p) return fail(MH_ERR_ARG, "null argument");
  for (uint32_t s = 0; s < b->n; ++s) {
    if (p[s].width != b->W || p[s].height != b->H)
      return fail(MH_ERR_ARG, "raycast image size must equal the batch's image size");
    RayConst R;
    const int rc = build_ray(b->vols[s], &p[s], b->vols[s]->V.cTw.m, &R);
    if (rc) return rc;
  }
  if (!b->d_range) HIP_TRY(dalloc(&b->d_range, (size_t)b->n * ray_range_words(b->W, b->H)));
  b->rp.assign(p, p + b->n);
  b->out_depth = depth_dev;
  b->out_vertex = vertex_dev;
  b->out_normal = normal_dev;
  b->have_ray = true;
  b->drop_graph();  // output pointers are baked into a captured graph
  return MH_OK;
}

int mh_batch_set_marching_cubes_shape(mh_batch* b, int shape) {
  if (!b) return fail(MH_ERR_ARG, "null batch");
  if (shape < 0 || shape >= kNumMcShapes) return fail(MH_ERR_ARG, "shape out of range");
  b->mc_shape = shape;
  b->mct.active = false;
  b->drop_graph();
  return MH_OK;
}

int mh_batch_tune_marching_cubes(mh_batch* b, uint32_t reps) {
  if (!b) return fail(MH_ERR_ARG, "null batch");
  int shape = b->mc_shape;
  const int rc = mc_tune_start(b->mct, b->stream, reps, false, &shape);
  if (shape != b->mc_shape) {
    b->mc_shape = shape;
    b->drop_graph();
  }
  return rc;
}

int mh_batch_marching_cubes_tuning(mh_batch* b, mh_tune_mc_result* out) {
  if (!b || !out) return fail(MH_ERR_ARG, "null argument");
  int shape = b->mc_shape;
  const int rc = mc_tune_collect(b->mct, b->stream, &shape);
  if (shape != b->mc_shape) {
    b->mc_shape = shape;
    b->drop_graph();
  }
  if (rc) return rc;
  *out = b->mct.result;
  if (!out->done) {
    out->candidates = kNumMcShapes;
    out->shape = b->mc_shape;
  }
  return MH_OK;
}

int mh_batch_set_integrate_tiling(mh_batch* b, int tiling, int grid_per_cu) {
  if (!b) return fail(MH_ERR_ARG, "null batch");
  const int rc = check_tiling(tiling, grid_per_cu);
  if (rc) return rc;
  b->tiling = tiling;
  b->grid_per_cu = grid_per_cu;
  b->drop_graph();
  return MH_OK;
}

int mh_batch_run(mh_batch* b, const mh_mc_params* mc, uint32_t flags, mh_tune_result* tune_out) {
  if (!b) return fail(MH_ERR_ARG, "null batch");
  if (!b->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  if ((flags & MH_RUN_RAYCAST) && !b->have_ray) return fail(MH_ERR_STATE, "mh_batch_set_raycast first");
  if ((flags & MH_RUN_GRAPH) && (flags & MH_RUN_TUNE))
    return fail(MH_ERR_ARG, "tuning runs eagerly; tune first, then replay the graph");
  // an online marching-cubes tuning times whole replays, one shape each (the
  // shape is set before the frame table, whose raycast constants depend on it)
  int trial = -1;
  if (flags & MH_RUN_GRAPH) {
    McTune& t = b->mct;
    if (t.active || t.pending >= 0) {
      int chosen = b->mc_shape;
      const int rc = mc_tune_collect(t, b->stream, &chosen);
      if (rc) return rc;
      if (chosen != b->mc_shape) {
        b->mc_shape = chosen;
        b->drop_graph();
      }
    }
    if (t.active && (flags & MH_RUN_MESH)) {
      trial = mc_trial_shape(t.trial++, kNumMcShapes);
      if (trial != b->mc_shape) {
        b->mc_shape = trial;
        b->drop_graph();
      }
    }
  }
  // this frame's table: Views (pose, raw depth), MC pass stamps, raycast poses
  for (uint32_t s = 0; s < b->n; ++s) {
    mh_volume* v = b->vols[s];
    SceneFrame& f = b->h_frames[s];
    std::memset(&f, 0, sizeof(f));
    if (flags & MH_RUN_MESH) {
      const int rc = build_mc(v, mc ? &mc[s] : nullptr, &f.mc);
      if (rc) return rc;
    }
    if (flags & MH_RUN_RAYCAST) {
      const int rc = build_ray(v, &b->rp[s], v->V.cTw.m, &f.ray);
      if (rc) return rc;
      f.ray.range = b->d_range + (size_t)s * ray_range_words(b->W, b->H);
      // Beside the meshing pass the march is the longer branch once the
      // surfel pass leaves room on the chip (fewer than all resident
      // workgroups): its waves then issue first on a shared SIMD.  Beside an
      // all-resident surfel pass the boost cost time instead.
      f.ray.boost = batch_overlaps(b, flags) && kMcShapes[b->mc_shape].grid_per_cu > 0;
    }
    f.V = v->V;
    f.V.raw = b->raw + (size_t)s * b->W * b->H;
    v->have_frame = true;
    ++v->work_gen;  // the chain rebuilds every scene's worklist
  }
  const int k = b->ring;
  b->ring = (k + 1) % mh_batch::kRing;
  HIP_TRY(hipEventSynchronize(b->ev[k]));  // the slot's previous upload has finished
  std::memcpy(b->pinned[k], b->h_frames.data(), sizeof(SceneFrame) * b->n);
  HIP_TRY(hipMemcpyAsync(b->d_frames, b->pinned[k], sizeof(SceneFrame) * b->n,
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipEventRecord(b->ev[k], b->stream));
  if (flags & MH_RUN_TUNE) {
    const int crc = batch_chain(b, flags, false);
    if (crc) return crc;
    HIP_TRY(hipGetLastError());
    const int rc = batch_tune(b, 3, tune_out);
    if (rc) return rc;
    b->drop_graph();
    const BatchDev B = b->dev();
    launch_batch_integrate_tsdf(B, b->stream, b->cus * b->grid_per_cu, b->tiling);
    if (flags & MH_RUN_MESH)
      launch_batch_marching_cubes(B, b->d_nbt, b->stream, b->cus, kMcShapes[b->mc_shape]);
    if (flags & MH_RUN_RAYCAST) {
      // tile shape: each trial writes the outputs and the work counters,
      // which are put back; the chosen shape's run below is the frame's
      Scratch sc;
      Counters* sctr = nullptr;
      HIP_TRY(sc.get(&sctr, b->n));
      for (uint32_t s = 0; s < b->n; ++s)
        HIP_TRY(hipMemcpyAsync(sctr + s, b->vols[s]->V.ctr, sizeof(Counters), hipMemcpyDeviceToDevice,
                               b->stream));
      auto run = [&](int k) { launch_batch_raycast(B, b->out_depth, b->out_vertex, b->out_normal, b->stream, k); };
      auto restore = [&]() -> int {
        for (uint32_t s = 0; s < b->n; ++s)
          HIP_TRY(hipMemcpyAsync(b->vols[s]->V.ctr, sctr + s, sizeof(Counters), hipMemcpyDeviceToDevice,
                                 b->stream));
        return MH_OK;
      };
      int best = 0;
      const int rc2 = tune_ray_shapes(b->stream, 3, run, restore, nullptr, &best);
      if (rc2) return rc2;
      b->ray_shape = best;
      if (tune_out) tune_out->ray_shape = best;
      launch_batch_raycast(B, b->out_depth, b->out_vertex, b->out_normal, b->stream, b->ray_shape);
    }
  } else if (flags & MH_RUN_GRAPH) {
    McTune& t = b->mct;
    const uint32_t shape = flags & (MH_RUN_MESH | MH_RUN_RAYCAST);
    if (!b->exec || b->graph_flags != shape) {
      b->drop_graph();
      HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int crc = batch_chain(b, flags, true);
      const hipError_t le = hipGetLastError();
      const hipError_t ee = hipStreamEndCapture(b->stream, &b->graph);
      if (crc) return crc;
      HIP_TRY(le);
      HIP_TRY(ee);
      HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_flags = shape;
    }
    if (trial >= 0) HIP_TRY(hipEventRecord(t.a, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    if (trial >= 0) {
      HIP_TRY(hipEventRecord(t.b, b->stream));
      t.pending = trial;
    }
  } else {
    const int crc = batch_chain(b, flags, true);
    if (crc) return crc;
  }
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_batch_synchronize(mh_batch* b) {
  if (!b) return fail(MH_ERR_ARG, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return MH_OK;
}

}  // extern "C"

// ------------------------------------------------------------ raycast adjoint
namespace {

// FP64 partial-sum rows (up to 30 slots each) for `rows` workgroups, and the
// 30-slot result.
hipError_t ensure_partials(mh_volume* v, uint32_t rows) {
  if (rows > v->rgrad_cap) {
    if (v->rgrad_partials) (void)hipFree(v->rgrad_partials);
    v->rgrad_partials = nullptr;
    v->rgrad_cap = 0;
    const hipError_t e = dalloc(&v->rgrad_partials, (size_t)rows * 30);
    if (e != hipSuccess) return e;
    v->rgrad_cap = rows;
  }
  if (!v->rgrad_out) return dalloc(&v->rgrad_out, 30);
  return hipSuccess;
}

// The adjoint from a built RayConst (its range buffer attached) into the
// given partial-sum rows and result row; synchronises the stream.
static_assert(sizeof(mh_raycast_grad) == 360, "mh_raycast_grad layout");

int raycast_grad_run(mh_volume* v, const RayConst& R, const float* obs, float* grad_sdf,
                     double* partials, double* res, mh_raycast_grad* out, float max_residual,
                     float max_slope) {
  launch_raycast_backward(v->V, R, obs, grad_sdf, partials, res, max_residual, max_slope, v->stream);
  HIP_TRY(hipGetLastError());
  double h[kGradSlots];
  HIP_TRY(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, v->stream));
  HIP_TRY(hipStreamSynchronize(v->stream));
  out->loss = h[0];
  for (int k = 0; k < 6; ++k) out->grad[k] = h[1 + k];
  int t = 7;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      out->hessian[6 * i + j] = h[t];
      out->hessian[6 * j + i] = h[t];
      ++t;
    }
  out->pixels = (uint32_t)h[28];
  out->outliers = (uint32_t)h[29];
  out->observed = (uint32_t)h[30];
  return MH_OK;
}

int raycast_grad(mh_volume* v, const mh_raycaster_params* p, const float cTw[16], const float* obs,
                 float* grad_sdf, mh_raycast_grad* out, float max_residual = INFINITY,
                 float max_slope = INFINITY) {
  RayConst R;
  int rc = build_ray(v, p, cTw, &R);
  if (rc) return rc;
  rc = attach_range(v, &R);
  if (rc) return rc;
  const uint32_t n = raycast_grad_partials(R);
  HIP_TRY(ensure_partials(v, n));
  return raycast_grad_run(v, R, obs, grad_sdf, v->rgrad_partials, v->rgrad_out, out, max_residual,
                          max_slope);
}

// ---------------------------------------------------- allocation-free adjoint
// mh_raycast_backward_noalloc keeps its scratch -- the tile-range words of
// the range pass, one FP64 partial-sum row per 8 x 8-pixel workgroup and the
// result row -- in one file-scope device array instead of buffers allocated
// on first use.  Each GPU holds its own copy of the array for the life of
// the process; its device address is looked up once per device
// (hipGetSymbolAddress) and only that address is handed to kernels and
// copies.  Images up to kNoAllocMaxWidth x kNoAllocMaxHeight fit.
constexpr uint32_t kNoAllocMaxWidth = 1280, kNoAllocMaxHeight = 1024;
constexpr uint32_t kNoAllocTiles = ((kNoAllocMaxWidth + 7) / 8) * ((kNoAllocMaxHeight + 7) / 8);
constexpr size_t kNoAllocRound = 256;
constexpr size_t round_up(size_t x) { return (x + kNoAllocRound - 1) / kNoAllocRound * kNoAllocRound; }
constexpr size_t kNoAllocPartialOff = 0;
constexpr size_t kNoAllocPartialBytes = round_up((size_t)kNoAllocTiles * kGradSlots * sizeof(double));
constexpr size_t kNoAllocResultOff = kNoAllocPartialOff + kNoAllocPartialBytes;
constexpr size_t kNoAllocResultBytes = round_up(kGradSlots * sizeof(double));
constexpr size_t kNoAllocRangeOff = kNoAllocResultOff + kNoAllocResultBytes;
constexpr size_t kNoAllocRangeBytes = round_up((2 * (size_t)kNoAllocTiles + 1) * sizeof(uint32_t));
constexpr size_t kNoAllocScratchBytes = kNoAllocRangeOff + kNoAllocRangeBytes;  // ~5.2 MB

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on device `dev`, which must be current.
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(MH_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(MH_ERR_HIP, "hipGetSymbolAddress returned no address");
    g_scratch_addr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_addr[dev];
  return MH_OK;
}

// exp of xi = (omega, upsilon) in se(3) as a row-major 4x4 (Rodrigues, and
// V = I + (1 - cos)/th^2 W + (th - sin)/th^3 W^2 for the translation).
void se3_exp(const double xi[6], double T[16]) {
  const double wx = xi[0], wy = xi[1], wz = xi[2];
  const double th2 = wx * wx + wy * wy + wz * wz, th = std::sqrt(th2);
  double A, B, C;
  if (th < 1e-6) {
    A = 1.0 - th2 / 6.0;
    B = 0.5 - th2 / 24.0;
    C = 1.0 / 6.0 - th2 / 120.0;
  } else {
    A = std::sin(th) / th;
    B = (1.0 - std::cos(th)) / th2;
    C = (th - std::sin(th)) / (th2 * th);
  }
  const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
  double W2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double a = 0;
      for (int k = 0; k < 3; ++k) a += W[3 * i + k] * W[3 * k + j];
      W2[3 * i + j] = a;
    }
  for (int i = 0; i < 3; ++i) {
    double t = 0;
    for (int j = 0; j < 3; ++j) {
      const double I = i == j ? 1.0 : 0.0;
      T[4 * i + j] = I + A * W[3 * i + j] + B * W2[3 * i + j];
      t += (I + B * W[3 * i + j] + C * W2[3 * i + j]) * xi[3 + j];
    }
    T[4 * i + 3] = t;
  }
  T[12] = T[13] = T[14] = 0.0;
  T[15] = 1.0;
}

void matmul4(const double* a, const double* b, double* c) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
      c[4 * i + j] = s;
    }
}

// Solves (H + lambda diag(H)) x = -g by Gaussian elimination with partial
// pivoting; false if singular.
bool solve6(const double* H, const double* g, double lambda, double* x) {
  double A[6][7];
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j < 6; ++j) A[i][j] = H[6 * i + j] + (i == j ? lambda * H[6 * i + j] : 0.0);
    A[i][6] = -g[i];
  }
  for (int c = 0; c < 6; ++c) {
    int piv = c;
    for (int r = c + 1; r < 6; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    if (!(std::fabs(A[piv][c]) > 0.0)) return false;
    for (int j = 0; j < 7; ++j) std::swap(A[c][j], A[piv][j]);
    for (int r = c + 1; r < 6; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int j = c; j < 7; ++j) A[r][j] -= f * A[c][j];
    }
  }
  for (int i = 5; i >= 0; --i) {
    double s = A[i][6];
    for (int j = i + 1; j < 6; ++j) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
  return true;
}

// Rays that nearly graze the surface (depth moving more than 3x a camera
// translation) carry unreliable Jacobians on a noisy SDF; refinement skips them.
constexpr float kMaxSlope = 3.0f;

int eval_pose(mh_volume* v, const mh_raycaster_params* p, const double T[16], const float* obs,
              float gate, mh_raycast_grad* out) {
  double inv[16];
  invert4<double>(T, inv);
  float cTw[16];
  for (int i = 0; i < 16; ++i) cTw[i] = (float)inv[i];
  return raycast_grad(v, p, cTw, obs, nullptr, out, gate, kMaxSlope);
}

// Residual gate of the pose refinement: pixels with |r| above it (occlusion
// boundaries, surfaces the volume does not hold) leave the least-squares
// problem.  The gate narrows as the pose settles: 0.20 m for the first 40 %
// of the steps (the start may be centimetres and degrees off), 0.10 m for
// the next 30 %, then 0.05 m.
float gate_at(uint32_t step, uint32_t steps) {
  if (10 * step < 4 * steps) return 0.20f;
  if (10 * step < 7 * steps) return 0.10f;
  return 0.05f;
}

// A truncated quadratic over every observed pixel: r^2 / 2 for a pixel in
// the sum, gate^2 / 2 for one that is gated out or that the pose loses (no
// hit, or a grazing ray).  The count of observed pixels does not depend on
// the pose, so a step cannot lower the cost by dropping pixels.
double truncated_cost(const mh_raycast_grad& g, float gate) {
  if (g.observed == 0) return INFINITY;
  const double left = (double)g.observed - (double)g.pixels;
  return (g.loss + 0.5 * (double)gate * (double)gate * left) / (double)g.observed;
}


}  // namespace

extern "C" {

int mh_raycast_backward(mh_volume* v, const mh_raycaster_params* p, const float cTw[16],
                        const float* obs_depth_dev, float* grad_sdf_dev, mh_raycast_grad* out) {
  if (check_vol(v) || !p || !cTw || !obs_depth_dev || !out) return fail(MH_ERR_ARG, "null argument");
  return raycast_grad(v, p, cTw, obs_depth_dev, grad_sdf_dev, out);
}

int mh_raycast_backward_noalloc(mh_volume* v, const mh_raycaster_params* p, const float cTw[16],
                                const float* obs_depth_dev, float* grad_sdf_dev,
                                mh_raycast_grad* out) {
  if (check_vol(v) || !p || !cTw || !obs_depth_dev || !out) return fail(MH_ERR_ARG, "null argument");
  if (p->width > kNoAllocMaxWidth || p->height > kNoAllocMaxHeight)
    return fail(MH_ERR_CAPACITY, "image larger than the allocation-free scratch (1280 x 1024)");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != v->device) return fail(MH_ERR_STATE, "the volume's device must be current");
  unsigned char* base = nullptr;
  int rc = scratch_address(dev, &base);
  if (rc) return rc;
  RayConst R;
  rc = build_ray(v, p, cTw, &R);
  if (rc) return rc;
  R.range = reinterpret_cast<uint32_t*>(base + kNoAllocRangeOff);  // cleared by the range pass
  return raycast_grad_run(v, R, obs_depth_dev, grad_sdf_dev,
                          reinterpret_cast<double*>(base + kNoAllocPartialOff),
                          reinterpret_cast<double*>(base + kNoAllocResultOff), out, INFINITY,
                          INFINITY);
}

int mh_refine_pose(mh_volume* v, const mh_raycaster_params* p, const float wTc_init[16],
                   const float* obs_depth_dev, uint32_t iterations, float wTc_out[16],
                   double* cost_trace) {
  if (check_vol(v) || !p || !wTc_init || !obs_depth_dev || !wTc_out)
    return fail(MH_ERR_ARG, "null argument");
  double T[16];
  for (int i = 0; i < 16; ++i) T[i] = wTc_init[i];
  float gate = gate_at(0, iterations);
  mh_raycast_grad cur;
  int rc = eval_pose(v, p, T, obs_depth_dev, gate, &cur);
  if (rc) return rc;
  if (cost_trace) cost_trace[0] = truncated_cost(cur, gate);
  double lambda = 1e-2;  // damped first steps: the start may be degrees off
  for (uint32_t it = 0; it < iterations; ++it) {
    double xi[6];
    bool improved = false;
    if (cur.pixels >= 6 && solve6(cur.hessian, cur.grad, lambda, xi)) {
      double E[16], Tn[16];
      se3_exp(xi, E);
      matmul4(E, T, Tn);
      mh_raycast_grad nxt;
      rc = eval_pose(v, p, Tn, obs_depth_dev, gate, &nxt);
      if (rc) return rc;
      if (truncated_cost(nxt, gate) < truncated_cost(cur, gate)) {
        std::memcpy(T, Tn, sizeof(T));
        cur = nxt;
        lambda = std::max(lambda * 0.1, 1e-9);
        improved = true;
      }
    }
    if (!improved) lambda *= 10.0;
    if (cost_trace) cost_trace[it + 1] = truncated_cost(cur, gate);
    const float g2 = gate_at(it + 1, iterations);
    if (g2 != gate && it + 1 < iterations) {  // re-linearise at the new gate
      gate = g2;
      rc = eval_pose(v, p, T, obs_depth_dev, gate, &cur);
      if (rc) return rc;
    }
  }
  // Each gate's steps only lower that gate's cost; the start is kept if the
  // result costs more than it at the final gate.
  double S[16];
  for (int i = 0; i < 16; ++i) S[i] = wTc_init[i];
  mh_raycast_grad at_start;
  rc = eval_pose(v, p, S, obs_depth_dev, gate, &at_start);
  if (rc) return rc;
  const bool keep_start = truncated_cost(at_start, gate) < truncated_cost(cur, gate);
  for (int i = 0; i < 16; ++i) wTc_out[i] = keep_start ? wTc_init[i] : (float)T[i];
  return MH_OK;
}

int mh_dump_voxel_field(mh_volume* v, const float* field_dev, uint32_t capacity, float* out_host,
                        uint32_t* count_host) {
  if (check_vol(v) || !field_dev) return fail(MH_ERR_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(v->stream));
  std::vector<BlockRec> blocks;
  const int rc = sorted_blocks(v, blocks);
  if (rc) return rc;
  if (count_host) *count_host = (uint32_t)blocks.size();
  if (!out_host) return MH_OK;
  if (capacity < blocks.size()) return fail(MH_ERR_CAPACITY, "buffer too small");
  return gather_field(v, field_dev, blocks, read_counters(v).block_hwm, 1, out_host);
}

}  // extern "C"

// ------------------------------------------------------------ localisation
namespace {

// PointToSurface (localizing/point_to_psdf.cu:84-147) on the current frame.
int point_to_psdf(mh_volume* v, const float wTc[16], int jacobian, mh_point_to_psdf_result* out) {
  flush_frame(v);
  HIP_TRY(ensure_partials(v, point_to_psdf_partials(v->V)));
  M4 T;
  std::memcpy(T.m, wTc, sizeof(T.m));
  launch_point_to_psdf(v->V, T, jacobian == MH_P2S_CORRECTED ? 1 : 0, v->rgrad_partials,
                       v->rgrad_out, v->stream);
  HIP_TRY(hipGetLastError());
  double h[kP2SSlots];
  HIP_TRY(hipMemcpyAsync(h, v->rgrad_out, sizeof(h), hipMemcpyDeviceToHost, v->stream));
  HIP_TRY(hipStreamSynchronize(v->stream));
  int t = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      out->A[6 * i + j] = h[t];
      out->A[6 * j + i] = h[t];
      ++t;
    }
  for (int i = 0; i < 6; ++i) out->b[i] = h[21 + i];
  out->err = h[27];
  out->count = (uint32_t)h[28];
  return MH_OK;
}

}  // namespace

extern "C" {

int mh_point_to_psdf(mh_volume* v, const float wTc[16], int jacobian, mh_point_to_psdf_result* out) {
  if (check_vol(v) || !wTc || !out) return fail(MH_ERR_ARG, "null argument");
  if (jacobian != MH_P2S_REFERENCE && jacobian != MH_P2S_CORRECTED)
    return fail(MH_ERR_ARG, "jacobian must be MH_P2S_REFERENCE or MH_P2S_CORRECTED");
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  return point_to_psdf(v, wTc, jacobian, out);
}

int mh_localize(mh_volume* v, uint32_t iterations, float lambda, int jacobian, float wTc[16],
                double* err_trace, uint32_t* count_trace, uint32_t* iterations_done) {
  if (check_vol(v) || !wTc) return fail(MH_ERR_ARG, "null argument");
  if (jacobian != MH_P2S_REFERENCE && jacobian != MH_P2S_CORRECTED)
    return fail(MH_ERR_ARG, "jacobian must be MH_P2S_REFERENCE or MH_P2S_CORRECTED");
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  uint32_t it = 0;
  for (; it < iterations; ++it) {
    mh_point_to_psdf_result r;
    const int rc = point_to_psdf(v, wTc, jacobian, &r);
    if (rc) return rc;
    if (err_trace) err_trace[it] = r.err;
    if (count_trace) count_trace[it] = r.count;
    if (r.count == 0) break;
    // SolveAndConvertDeltaXi (main_engine.cc:28-54): (A + lambda I) dxi = -b,
    // dT = exp(dxi) with dxi = (translation, rotation), wTc <- dT wTc.
    double H[36], x[6];
    for (int k = 0; k < 36; ++k) H[k] = r.A[k] + (k % 7 == 0 ? (double)lambda : 0.0);
    if (!solve6(H, r.b, 0.0, x)) break;
    const double xi[6] = {x[3], x[4], x[5], x[0], x[1], x[2]};
    double E[16], T[16], Tn[16];
    se3_exp(xi, E);
    for (int k = 0; k < 16; ++k) T[k] = wTc[k];
    matmul4(E, T, Tn);
    for (int k = 0; k < 16; ++k) wTc[k] = (float)Tn[k];
  }
  if (iterations_done) *iterations_done = it;
  set_pose(v, wTc);  // Sensor::set_transform
  return MH_OK;
}

}  // extern "C"

// ------------------------------------------------------------ mesh compaction
namespace {

int ensure_compact(mh_volume* v) {
  if (v->cm_used) return MH_OK;
  const size_t nv = v->mp.max_vertex_count, cap = v->hp.value_capacity;
  const size_t nt = v->mp.max_triangle_count;
  const size_t chunks = std::max((nv + 1023) / 1024, (cap + 1023) / 1024) + 1;
  Scratch sc;
  uint32_t *used, *remap, *blk, *blk_off, *sums, *total;
  float *pos, *nrm, *col;
  int32_t* tri;
  HIP_TRY(sc.get(&used, nv));
  HIP_TRY(sc.get(&remap, nv));
  HIP_TRY(sc.get(&blk, cap));
  HIP_TRY(sc.get(&blk_off, cap));
  HIP_TRY(sc.get(&sums, chunks));
  HIP_TRY(sc.get(&total, 2));
  HIP_TRY(sc.get(&pos, 3 * nv));
  HIP_TRY(sc.get(&nrm, 3 * nv));
  HIP_TRY(sc.get(&col, 3 * nv));
  HIP_TRY(sc.get(&tri, 3 * nt));
  // every buffer exists: publish them together
  sc.release_to(v->owned);
  v->cm_used = used;
  v->cm_remap = remap;
  v->cm_blk = blk;
  v->cm_blk_off = blk_off;
  v->cm_sums = sums;
  v->cm_total = total;
  v->cm_pos = pos;
  v->cm_nrm = nrm;
  v->cm_col = col;
  v->cm_tri = tri;
  return MH_OK;
}

}  // namespace

extern "C" {

int mh_compress_mesh(mh_volume* v, uint32_t* vertex_count, uint32_t* triangle_count) {
  if (check_vol(v)) return MH_ERR_ARG;
  int rc = ensure_compact(v);
  if (rc) return rc;
  const uint32_t n = read_counters(v).cand_count;
  const uint32_t nv = v->mp.max_vertex_count;
  HIP_TRY(hipMemsetAsync(v->cm_used, 0, (size_t)nv * sizeof(uint32_t), v->stream));
  const int grid = v->grid_wg512();
  launch_compact_count(v->V, n, v->cm_used, v->cm_blk, v->stream, grid);
  launch_scan(v->cm_blk, v->cm_blk_off, n, v->cm_sums, v->cm_total + 0, v->stream);
  launch_scan(v->cm_used, v->cm_remap, nv, v->cm_sums, v->cm_total + 1, v->stream);
  HIP_TRY(hipGetLastError());
  uint32_t tot[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(tot, v->cm_total, sizeof(tot), hipMemcpyDeviceToHost, v->stream));
  HIP_TRY(hipStreamSynchronize(v->stream));
  if (tot[0] > v->mp.max_triangle_count)
    return fail(MH_ERR_CAPACITY, "more live triangles than max_triangle_count");
  launch_compact_emit(v->V, n, v->cm_blk_off, v->cm_remap, v->cm_tri, v->cm_used, v->cm_pos,
                      v->cm_nrm, v->cm_col, v->stream, grid);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(v->stream));
  v->cm_triangles = tot[0];
  v->cm_vertices = tot[1];
  if (vertex_count) *vertex_count = tot[1];
  if (triangle_count) *triangle_count = tot[0];
  return MH_OK;
}

int mh_compact_mesh_get(mh_volume* v, uint32_t vertex_capacity, float* pos_host, float* normal_host,
                        float* color_host, uint32_t triangle_capacity, int32_t* tri_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->cm_used) return fail(MH_ERR_STATE, "mh_compress_mesh first");
  if ((pos_host || normal_host || color_host) && vertex_capacity < v->cm_vertices)
    return fail(MH_ERR_CAPACITY, "vertex buffers too small");
  if (tri_host && triangle_capacity < v->cm_triangles)
    return fail(MH_ERR_CAPACITY, "triangle buffer too small");
  const size_t vb = (size_t)v->cm_vertices * 3 * sizeof(float);
  if (pos_host && vb) HIP_TRY(hipMemcpy(pos_host, v->cm_pos, vb, hipMemcpyDeviceToHost));
  if (normal_host && vb) HIP_TRY(hipMemcpy(normal_host, v->cm_nrm, vb, hipMemcpyDeviceToHost));
  if (color_host && vb) HIP_TRY(hipMemcpy(color_host, v->cm_col, vb, hipMemcpyDeviceToHost));
  const size_t tb = (size_t)v->cm_triangles * 3 * sizeof(int32_t);
  if (tri_host && tb) HIP_TRY(hipMemcpy(tri_host, v->cm_tri, tb, hipMemcpyDeviceToHost));
  return MH_OK;
}

// SavePly (io/mesh_writer.cc:75-150): ASCII PLY, vertex x y z nx ny nz and
// colour int(255 * c) per channel, faces "3 i j k"; numbers as the default
// C++ stream formatting.  int() of a NaN colour (plain-TSDF vertices carry
// ValToRGB of 0/0) is undefined in C++; here it is 0, other values saturate.
static int ply_channel(float c) {
  const float x = 255.0f * c;
  if (!(x == x)) return 0;
  if (x >= 2147483647.0f) return 2147483647;
  if (x <= -2147483648.0f) return -2147483647 - 1;
  return (int)x;
}
int mh_save_ply(mh_volume* v, const char* path) {
  if (check_vol(v) || !path) return fail(MH_ERR_ARG, "null argument");
  if (!v->cm_used) return fail(MH_ERR_STATE, "mh_compress_mesh first");
  const uint32_t nv = v->cm_vertices, nt = v->cm_triangles;
  std::vector<float> pos((size_t)3 * nv), nrm((size_t)3 * nv), col((size_t)3 * nv);
  std::vector<int32_t> tri((size_t)3 * nt);
  const int rc = mh_compact_mesh_get(v, nv, pos.data(), nrm.data(), col.data(), nt, tri.data());
  if (rc) return rc;
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(MH_ERR_ARG, std::string("cannot open ") + path);
  std::ostringstream ss;
  ss << "ply\nformat ascii 1.0\n";
  ss << "element vertex " << nv << "\n";
  ss << "property float x\nproperty float y\nproperty float z\n"
        "property float nx\nproperty float ny\nproperty float nz\n"
        "property uchar red\nproperty uchar green\nproperty uchar blue\n";
  ss << "element face " << nt << "\n";
  ss << "property list uchar int vertex_index\nend_header\n";
  for (uint32_t i = 0; i < nv; ++i) {
    const float* p = &pos[3 * (size_t)i];
    const float* n = &nrm[3 * (size_t)i];
    const float* c = &col[3 * (size_t)i];
    ss << p[0] << " " << p[1] << " " << p[2] << " " << n[0] << " " << n[1] << " " << n[2] << " "
       << ply_channel(c[0]) << " " << ply_channel(c[1]) << " " << ply_channel(c[2]) << "\n";
  }
  for (uint32_t i = 0; i < nt; ++i)
    ss << "3 " << tri[3 * (size_t)i] << " " << tri[3 * (size_t)i + 1] << " " << tri[3 * (size_t)i + 2]
       << "\n";
  const std::string out = ss.str();
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  std::fclose(f);
  if (!ok) return fail(MH_ERR_ARG, std::string("write failed: ") + path);
  return MH_OK;
}

}  // extern "C"
