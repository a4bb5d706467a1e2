// capi.hip -- the C ABI (include/meshhashing.h): volume lifecycle, frame
// upload, stage launchers, canonical dumps and restore.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mh_kernels.h"
#include "mh_mc_tables.h"

using namespace mh;

static constexpr McTables kHostMc = build_mc_tables();

namespace {

thread_local std::string g_err;

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

}  // namespace

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
  bool have_frame = false;
  int32_t* pd_nbr = nullptr;
  uint32_t pd_nbr_cap = 0;
  double* pd_partials = nullptr;
  double* pd_energy = nullptr;
  uint32_t pd_npartials = 0;
  uint32_t mc_pass = 0;
  std::vector<void*> owned;

  int grid_wg256() const { return cus * 8; }
  int grid_wg512() const { return cus * 4; }
};

namespace {

int check_vol(mh_volume* v) {
  if (!v) return fail(MH_ERR_ARG, "null volume");
  return MH_OK;
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
  A(&V.cand, cap);
  A(&V.depth, npix);
  A(&V.inlier, npix);
  A(&v->raw, npix);
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
  delete v;
  return MH_OK;
}

int mh_volume_reset(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  launch_reset(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(v->stream));
  v->have_frame = false;
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
  out->triangle_count = (uint32_t)c.tri_live;
  if (c.tri_live > v->mp.max_triangle_count) out->flags |= MH_STATUS_TRIANGLE_HEAP_EXHAUSTED;
  out->voxel_updates = c.voxel_updates;
  out->blocks_integrated = c.blocks_integrated;
  out->mc_vertex_writes = c.mc_vertex_writes;
  out->mc_triangle_writes = c.mc_triangle_writes;
  out->mc_blocks = c.mc_blocks;
  out->ray_samples = c.ray_samples;
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
  launch_convert_depth(v->V, depth_dev, v->stream);
  HIP_TRY(hipGetLastError());
  v->have_frame = true;
  return MH_OK;
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
  *depth_dev = v->V.depth;
  return MH_OK;
}

int mh_inlier_ratio_device_ptr(mh_volume* v, float** inlier_dev) {
  if (check_vol(v) || !inlier_dev) return fail(MH_ERR_ARG, "null argument");
  *inlier_dev = v->V.inlier;
  return MH_OK;
}

int mh_alloc_blocks(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  launch_alloc(v->V, v->stream);
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

static int collect(mh_volume* v, int frustum, uint32_t* count_host) {
  launch_collect(v->V, frustum, v->stream, v->cus * 4);
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

int mh_integrate_tsdf(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  launch_integrate_tsdf(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_predict_inlier(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->psdf) return fail(MH_ERR_STATE, "volume created without PSDF fields");
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  launch_predict_inlier(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_integrate_psdf(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->psdf) return fail(MH_ERR_STATE, "volume created without PSDF fields");
  if (!v->have_frame) return fail(MH_ERR_STATE, "no frame uploaded");
  launch_integrate_psdf(v->V, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_marching_cubes(mh_volume* v, const mh_mc_params* p) {
  if (check_vol(v)) return MH_ERR_ARG;
  mh_mc_params d;
  mh_mc_params_default(&d);
  if (!p) p = &d;
  if (!p->enable_sdf_gradient)
    return fail(MH_ERR_ARG, "flat per-triangle normals are not supported");
  McConst c;
  c.sdf_threshold = p->sdf_threshold;
  c.rho_min = p->rho_min;
  const float inv = 1.0f / v->vp.voxel_size;
  c.weight_min = p->weight_min < 0.0f ? inv * inv : p->weight_min;
  c.snap = p->snap_distance;
  c.flat_normals = 0;
  c.pass = ++v->mc_pass;
  launch_marching_cubes(v->V, c, v->stream, v->grid_wg512());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_raycast(mh_volume* v, const mh_raycaster_params* p, const float cTw[16], float* depth_dev,
               float* vertex_dev, float* normal_dev) {
  if (check_vol(v) || !p || !cTw) return fail(MH_ERR_ARG, "null argument");
  if (p->width == 0 || p->height == 0 || p->width > 16384 || p->height > 16384 ||
      !(p->raycast_step > 0.0f) || !(p->min_raycast_depth > 0.0f) ||
      !(p->max_raycast_depth >= p->min_raycast_depth) ||
      (p->max_raycast_depth - p->min_raycast_depth) / p->raycast_step > 65536.0f)
    return fail(MH_ERR_ARG, "bad raycaster params (sizes, depth range, or > 65536 steps per ray)");
  RayConst R;
  R.cam = Cam{p->fx, p->fy, p->cx, p->cy, p->width, p->height, p->min_raycast_depth,
              p->max_raycast_depth};
  R.min_depth = p->min_raycast_depth;
  R.max_depth = p->max_raycast_depth;
  R.step = p->raycast_step;
  R.sample_sdf_threshold = p->sample_sdf_threshold;
  R.sdf_threshold = p->sdf_threshold;
  R.enable_gradients = p->enable_gradients;
  R.inv_vs = 1.0f / v->vp.voxel_size;
  std::memcpy(R.cTw.m, cTw, sizeof(float) * 16);
  invert4<float>(cTw, R.wTc.m);
  launch_raycast(v->V, R, depth_dev, vertex_dev, normal_dev, v->stream);
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

static int ensure_pd(mh_volume* v) {
  View& V = v->V;
  if (!V.pd_sdf0) {
    const size_t nvox = (size_t)v->hp.value_capacity * kBlockVoxels;
    HIP_TRY(dalloc(&V.pd_sdf0, nvox));
    v->owned.push_back(V.pd_sdf0);
    HIP_TRY(dalloc(&V.pd_sdf_bar, nvox));
    v->owned.push_back(V.pd_sdf_bar);
    HIP_TRY(dalloc(&V.pd_w, nvox));
    v->owned.push_back(V.pd_w);
    HIP_TRY(dalloc(&V.pd_p, nvox * 3));
    v->owned.push_back(V.pd_p);
    HIP_TRY(dalloc(&V.pd_mask, nvox));
    v->owned.push_back(V.pd_mask);
    HIP_TRY(hipMemsetAsync(V.pd_mask, 0, nvox, v->stream));
    HIP_TRY(hipMemsetAsync(V.pd_sdf0, 0, nvox * 4, v->stream));
    HIP_TRY(hipMemsetAsync(V.pd_sdf_bar, 0, nvox * 4, v->stream));
    HIP_TRY(hipMemsetAsync(V.pd_w, 0, nvox * 4, v->stream));
    HIP_TRY(hipMemsetAsync(V.pd_p, 0, nvox * 12, v->stream));
    HIP_TRY(dalloc(&v->pd_nbr, (size_t)v->hp.value_capacity * 6));
    v->owned.push_back(v->pd_nbr);
    v->pd_npartials = (uint32_t)v->grid_wg256();
    HIP_TRY(dalloc(&v->pd_partials, (size_t)v->pd_npartials * 2));
    v->owned.push_back(v->pd_partials);
    HIP_TRY(dalloc(&v->pd_energy, 4));
    v->owned.push_back(v->pd_energy);
  }
  return MH_OK;
}

int mh_primal_dual_init(mh_volume* v) {
  if (check_vol(v)) return MH_ERR_ARG;
  const int rc = ensure_pd(v);
  if (rc != MH_OK) return rc;
  launch_pd_neighbors(v->V, v->pd_nbr, v->stream, v->grid_wg256());
  launch_pd_init(v->V, v->pd_nbr, v->stream, v->grid_wg256());
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_primal_dual_iterate(mh_volume* v, float lambda, float sigma, float tau, double* energy_host) {
  if (check_vol(v)) return MH_ERR_ARG;
  if (!v->V.pd_sdf0) return fail(MH_ERR_STATE, "mh_primal_dual_init not called");
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
  uint64_t* dkeys = nullptr;
  float *dsdf = nullptr, *dw = nullptr, *da = nullptr, *db = nullptr;
  uint32_t* dc = nullptr;
  HIP_TRY(dalloc(&dkeys, n));
  HIP_TRY(dalloc(&dsdf, nv));
  HIP_TRY(dalloc(&dw, nv));
  HIP_TRY(hipMemcpyAsync(dkeys, keys.data(), n * 8, hipMemcpyHostToDevice, v->stream));
  HIP_TRY(hipMemcpyAsync(dsdf, sdf_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  HIP_TRY(hipMemcpyAsync(dw, weight_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  if (color_host) {
    HIP_TRY(dalloc(&dc, nv));
    HIP_TRY(hipMemcpyAsync(dc, color_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  }
  if (a_host) {
    HIP_TRY(dalloc(&da, nv));
    HIP_TRY(hipMemcpyAsync(da, a_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  }
  if (b_host) {
    HIP_TRY(dalloc(&db, nv));
    HIP_TRY(hipMemcpyAsync(db, b_host, nv * 4, hipMemcpyHostToDevice, v->stream));
  }
  launch_load_insert(v->V, dkeys, n, v->stream);
  launch_load_scatter(v->V, dkeys, n, dsdf, dw, dc, da, db, v->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(v->stream));
  for (void* p : {(void*)dkeys, (void*)dsdf, (void*)dw, (void*)dc, (void*)da, (void*)db})
    if (p) (void)hipFree(p);
  return MH_OK;
}

}  // extern "C"
