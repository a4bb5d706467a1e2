// mh_kernels.h -- kernel sources (one volume or a batch of scenes) and the
// host launchers of the HIP kernels (one group per .hip file).
#pragma once
#include "mh_core.h"

namespace mh {

// meshing.hip constants
struct McConst {
  float sdf_threshold;
  float rho_min;
  float weight_min;
  double snap;
  int flat_normals;
  uint32_t pass;  // 1, 2, ...: per-call stamp for single-writer vertex payloads
};

// raycast.hip constants
struct RayConst {
  Cam cam;
  float min_depth, max_depth, step;
  float sample_sdf_threshold, sdf_threshold;
  int enable_gradients;
  float inv_vs;  // 1 / voxel_size (fast empty-space test only)
  M4 cTw, wTc;
  // Camera-z range of the allocated blocks over each 8 x 8 pixel tile
  // (2 words per tile, float bits: zmin, zmax; then one word that, when
  // non-zero, disables the range test), or null.  Filled by the range pass.
  uint32_t* range;
  uint32_t tiles_x, tiles;
  // the march's waves raise their issue priority (a batch's raycast beside a
  // meshing pass that leaves room on the chip)
  int boost;
};
// Words of a range buffer for a W x H raycast.
inline uint32_t ray_range_words(uint32_t w, uint32_t h) { return 2u * ((w + 7) / 8) * ((h + 7) / 8) + 1u; }

// Everything a batched kernel needs about one scene for one frame.
struct SceneFrame {
  View V;
  McConst mc;
  RayConst ray;
};

// A batch work item: scene << kSceneShift | block pool index.
constexpr uint32_t kSceneShift = 24;
constexpr uint32_t kMaxScenes = 256;
constexpr uint32_t kMaxBatchCapacity = 1u << kSceneShift;

// Kernel sources.  A kernel templated on its source runs either over one
// volume (One: the View is a kernel argument, the worklist is the volume's
// own) or over a batch of independent scenes (Many: per-scene state lives in
// a device table of SceneFrames; pixel kernels take the scene from
// blockIdx.z, worklist kernels from the concatenated (scene, block) list).
// Items of one scene are processed by exactly the per-scene code.
// item(ci) needs a wave-uniform ci; lane_item(ci) takes any ci.
struct One {
  View V;
  __device__ __forceinline__ const View& view(uint32_t) const { return V; }
  __device__ __forceinline__ uint32_t count() const { return V.ctr->cand_count; }
  __device__ __forceinline__ uint32_t* next_ctr() const { return &V.ctr->mc_next; }
  __device__ __forceinline__ int32_t item(uint32_t ci, uint32_t* scene) const {
    *scene = 0;
    return V.cand[ci];
  }
  __device__ __forceinline__ int32_t lane_item(uint32_t ci, uint32_t* scene) const {
    return item(ci, scene);
  }
};

struct Many {
  const SceneFrame* __restrict__ f;
  const uint32_t* __restrict__ work;
  const uint32_t* work_count;
  uint32_t* dyn;  // items handed out on demand (marching cubes), after the static ones
  __device__ __forceinline__ const View& view(uint32_t s) const { return f[s].V; }
  __device__ __forceinline__ uint32_t count() const { return *work_count; }
  __device__ __forceinline__ uint32_t* next_ctr() const { return dyn; }
  __device__ __forceinline__ int32_t item(uint32_t ci, uint32_t* scene) const {
    const uint32_t w = __builtin_amdgcn_readfirstlane(work[ci]);  // ci is wave-uniform
    *scene = w >> kSceneShift;
    return (int32_t)(w & (kMaxBatchCapacity - 1u));
  }
  __device__ __forceinline__ int32_t lane_item(uint32_t ci, uint32_t* scene) const {
    const uint32_t w = work[ci];
    *scene = w >> kSceneShift;
    return (int32_t)(w & (kMaxBatchCapacity - 1u));
  }
};

// A list of one volume's blocks other than its worklist (garbage blocks).
struct Listed {
  View V;
  const int32_t* items;
  const uint32_t* n;
  __device__ __forceinline__ const View& view(uint32_t) const { return V; }
  __device__ __forceinline__ uint32_t count() const { return *n; }
  __device__ __forceinline__ int32_t item(uint32_t ci, uint32_t* scene) const {
    *scene = 0;
    return items[ci];
  }
  __device__ __forceinline__ int32_t lane_item(uint32_t ci, uint32_t* scene) const {
    return item(ci, scene);
  }
};

// Batched kernels also take the scene table as a __restrict__ kernel
// argument: stores through the per-scene pointers then provably leave the
// table unchanged, so per-scene constants load into scalar registers.
__device__ __forceinline__ One with_table(const One& o, const SceneFrame*) { return o; }
__device__ __forceinline__ Listed with_table(const Listed& o, const SceneFrame*) { return o; }
__device__ __forceinline__ Many with_table(const Many& m, const SceneFrame* t) {
  return Many{t, m.work, m.work_count, m.dyn};
}
inline const SceneFrame* table_of(const One&) { return nullptr; }
inline const SceneFrame* table_of(const Listed&) { return nullptr; }
inline const SceneFrame* table_of(const Many& m) { return m.f; }

// Device-side description of a batch, for the batched launchers.
struct BatchDev {
  SceneFrame* frames;   // device table, nscenes entries
  uint32_t* work;       // concatenated worklist
  uint32_t* work_count;
  uint32_t* dyn;        // marching cubes' on-demand item counter
  uint32_t nscenes;
  uint32_t width, height;  // common image size
  int psdf;             // scenes carry PSDF fields (all or none)
  Many src() const { return Many{frames, work, work_count, dyn}; }
};

// Integration tilings: workgroup size x blocks per workgroup (each block
// gets NT / BPW lanes, a multiple of 64, and 512 * BPW / NT voxels per lane).
struct Tiling {
  int nt, bpw;
  bool early;  // state loaded with the depth gathers, keys fetched per wave
};
constexpr Tiling kIntegrateTilings[] = {
    {64, 1, false},  {128, 1, false}, {256, 1, false}, {512, 1, false}, {128, 2, false},
    {256, 2, false}, {256, 4, false}, {512, 2, false}, {512, 4, false}, {512, 8, false},
    {64, 1, true},   {128, 1, true},  {256, 1, true},  {512, 1, true},  {128, 2, true},
    {256, 2, true},  {256, 4, true},  {512, 4, true}};
constexpr int kNumIntegrateTilings = sizeof(kIntegrateTilings) / sizeof(kIntegrateTilings[0]);
constexpr int kDefaultIntegrateTiling = 2;  // 256 lanes, one block, 2 voxels per lane

// Optional per-kernel timing (the reference's per-stage timers, util/timer.h).
// A launcher that runs several kernels calls tick(s, name) before each of
// them and tick(s, nullptr) after the last; the host side records HIP events
// at those points.  Null: no timing.
struct KernelClock {
  void (*fn)(void* ctx, hipStream_t s, const char* next);
  void* ctx;
  void tick(hipStream_t s, const char* next) const { fn(ctx, s, next); }
};
inline void tick(const KernelClock* c, hipStream_t s, const char* next) {
  if (c) c->tick(s, next);
}

// mapping.hip
void launch_reset(const View& V, hipStream_t s, int grid);
void launch_convert_depth(const View& V, const uint16_t* raw, hipStream_t s);
void launch_convert_color(const View& V, const uint8_t* bgra, hipStream_t s);
void launch_alloc(const View& V, hipStream_t s, const KernelClock* kc = nullptr);
void launch_collect(const View& V, int frustum, hipStream_t s, int grid, const KernelClock* kc = nullptr,
                    bool count_cleared = false);
void launch_integrate_tsdf(const View& V, hipStream_t s, int grid, int tiling,
                           const KernelClock* kc = nullptr);
void launch_predict_inlier(const View& V, hipStream_t s, int grid);
// PSDF fusion: workgroup sizes the tuner tries (each lane fuses 512 / size
// voxels of a block; one block per workgroup at a time).
constexpr int kPsdfLanes[] = {128, 256, 512};
constexpr int kNumPsdfTilings = sizeof(kPsdfLanes) / sizeof(kPsdfLanes[0]);
constexpr int kDefaultPsdfTiling = 1;
void launch_integrate_psdf(const View& V, hipStream_t s, int grid, int lanes = 256);
void launch_load_insert(const View& V, const uint64_t* keys, uint32_t n, hipStream_t s);
void launch_load_scatter(const View& V, const uint64_t* keys, uint32_t n, const float* sdf,
                         const float* weight, const uint32_t* color, const float* a,
                         const float* b, hipStream_t s);
// worklist state save / restore (autotuning runs the stateful integration
// repeatedly from one starting state)
void launch_state_save(const View& V, float* sdf, float* weight, uint32_t* color, hipStream_t s,
                       int grid);
void launch_state_restore(const View& V, const float* sdf, const float* weight,
                          const uint32_t* color, hipStream_t s, int grid);

void launch_batch_alloc(const BatchDev& B, hipStream_t s);
// stamp_mc: also stamp the collected blocks with the frame table's MC pass
// (then launch_batch_marching_cubes(..., marks_stamped = true))
void launch_batch_collect(const BatchDev& B, int frustum, hipStream_t s, int grid, bool counts_cleared = false,
                          bool stamp_mc = false);
void launch_batch_integrate_tsdf(const BatchDev& B, hipStream_t s, int grid, int tiling);

// block recycling (mapping/recycle.cu, core/hash_table.h:183-255)
constexpr int32_t kBlockLife = 3;  // BLOCK_LIFE, core/block.h:13
void launch_gc_low_surfel(const View& V, hipStream_t s, int grid);
void launch_gc_starve(const View& V, hipStream_t s, int grid);
void launch_gc_flag_garbage(const View& V, hipStream_t s, int grid);
void launch_gc_recycle(const View& V, hipStream_t s, int cus);

// meshing.hip
void launch_gc_release_triangles(const View& V, hipStream_t s, int grid);
void launch_gc_recycle_vertices(const View& V, hipStream_t s, int cus);
// nbt: the 27-neighbour table, 27 int32 per worklist item (for a batch, per
// item of the concatenated worklist)
// Marching-cubes launch shape: surfel workgroups per CU (0: all resident)
// and the quarters of the worklist handed out statically; output-invariant.
struct McShape {
  int grid_per_cu;
  int static_quarters;
};
constexpr McShape kMcShapes[] = {{0, 3}, {0, 2}, {0, 4}, {2, 3}, {2, 2}, {2, 4}, {1, 3}, {1, 2}, {1, 4}};
constexpr int kNumMcShapes = sizeof(kMcShapes) / sizeof(kMcShapes[0]);
void launch_marching_cubes(const View& V, const McConst& p, int32_t* nbt, hipStream_t s, int cus,
                           const KernelClock* kc = nullptr, McShape shape = kMcShapes[0]);
// marks_stamped: the collect pass stamped the marks (no marking launch)
void launch_batch_marching_cubes(const BatchDev& B, int32_t* nbt, hipStream_t s, int cus,
                                 McShape shape = kMcShapes[0], bool marks_stamped = false);
// mesh compaction (CompressMesh) and an exclusive scan of n uint32 (chunk_sums:
// ceil(n / 1024) entries)
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* chunk_sums,
                 uint32_t* total, hipStream_t s);
void launch_compact_count(const View& V, uint32_t n, uint32_t* used, uint32_t* blk_tris,
                          hipStream_t s, int grid);
void launch_compact_emit(const View& V, uint32_t n, const uint32_t* blk_off, const uint32_t* remap,
                         int32_t* tri_out, const uint32_t* used, float* pos, float* nrm, float* col,
                         hipStream_t s, int grid);

// raycast.hip
// Raycast tile shapes: tile width in pixels (tile = width x 64 / width).
constexpr int kRayTileWidths[] = {8, 16, 4, 32};
constexpr int kNumRayShapes = sizeof(kRayTileWidths) / sizeof(kRayTileWidths[0]);
// raycast waves resident per SIMD (the kernel is held to 96 registers)
constexpr int kRayWavesPerSimd = 5;
void launch_raycast(const View& V, const RayConst& p, float* depth, float* vertex, float* normal,
                    hipStream_t s, const KernelClock* kc = nullptr, int shape = 0);
// the march alone (the tile ranges already filled)
void launch_raycast_march(const View& V, const RayConst& p, float* depth, float* vertex, float* normal,
                          hipStream_t s, int shape);
void launch_ray_range_one(const View& V, const RayConst& p, hipStream_t s);
// diagnostics: add the distinct blocks the last raycast sampled (V.ray_mark
// == V.ray_stamp) to the work counters
void launch_ray_mark_count(const View& V, hipStream_t s, int grid);
void launch_batch_raycast(const BatchDev& B, float* depth, float* vertex, float* normal,
                          hipStream_t s, int shape = 0);
// adjoint: 31 FP64 sums (loss, dL/dxi[6], J^T J upper triangle[21], pixels,
// gated-out pixels, observed pixels)
constexpr int kGradSlots = 31;
uint32_t raycast_grad_partials(const RayConst& p);
void launch_raycast_backward(const View& V, const RayConst& p, const float* obs, float* grad_sdf,
                             double* partials, double* out, float max_residual, float max_slope,
                             hipStream_t s);
// point-to-PSDF normal equations: 29 FP64 sums (A upper triangle[21], b[6],
// err, count) over the current frame's depth
constexpr int kP2SSlots = 29;
uint32_t point_to_psdf_partials(const View& V);
void launch_point_to_psdf(const View& V, const M4& wTc, int corrected, double* partials, double* out,
                          hipStream_t s);

// primal_dual.hip
void launch_pd_neighbors(const View& V, int32_t* nbr, hipStream_t s, int grid);
void launch_pd_init(const View& V, const int32_t* nbr, hipStream_t s, int grid);
void launch_pd_iterate(const View& V, const int32_t* nbr, float lambda, float sigma, float tau,
                       double* partials, uint32_t npartials, double* energy, hipStream_t s,
                       int grid);

}  // namespace mh
