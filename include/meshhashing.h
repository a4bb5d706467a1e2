/*
 * meshhashing.h -- C ABI of the MI355X (gfx950) reconstruction hot path.
 *
 * Every entry point replaces one free function or lifecycle method of the
 * reference MeshHashing engine; the reference file:line it stands in for is
 * given beside each declaration.  Plain C: POD parameter structs, an opaque
 * volume handle, plain pointers and sizes, int status codes (MH_OK == 0).
 *
 * Threading: a volume is used by one host thread at a time.  All kernels of a
 * volume run on the stream given by mh_volume_set_stream (default: the null
 * stream).  Entry points that return data to host memory synchronise that
 * stream; every other entry point only enqueues work.
 *
 * Pointer conventions: a parameter named *_dev is device memory (HBM), one
 * named *_host is host memory.  A NULL output pointer means "not wanted".
 */
#ifndef MESHHASHING_H
#define MESHHASHING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
  MH_OK = 0,
  MH_ERR_ARG = -1,      /* bad argument (NULL handle, size, range)           */
  MH_ERR_HIP = -2,      /* a HIP runtime call failed                         */
  MH_ERR_CAPACITY = -3, /* a buffer given by the caller is too small         */
  MH_ERR_STATE = -4,    /* call out of order (e.g. no frame uploaded yet)    */
  MH_ERR_EMPTY = -5     /* empty worklist; mirrors the reference's -1 return
                           (mapping/update_bayesian.cu:275-276,
                            meshing/marching_cubes.cu:336-337)               */
};

/* ---- device status word bits (mh_status.flags) -------------------------- */
enum {
  MH_STATUS_BLOCK_HEAP_EXHAUSTED = 1u << 0,    /* hash_table.h:279-286        */
  MH_STATUS_INSERT_DROPPED = 1u << 1,          /* list window full, :149-177  */
  MH_STATUS_DDA_CAP = 1u << 2,                 /* allocate.cu:72 (1024 steps) */
  MH_STATUS_VERTEX_HEAP_EXHAUSTED = 1u << 3,   /* mesh.h:53-59                */
  MH_STATUS_TRIANGLE_HEAP_EXHAUSTED = 1u << 4, /* mesh.h:67-73                */
  MH_STATUS_COORD_RANGE = 1u << 5,  /* block coordinate outside +-2^20        */
  MH_STATUS_OVERFLOW_QUEUE_FULL = 1u << 6, /* too many list inserts / frame   */
  MH_STATUS_STENCIL_RANGE = 1u << 7 /* MC stencil left its staged shell       */
};

/* ---- parameter PODs: same fields as core/params.h:16-74 ----------------- */
typedef struct mh_hash_params {   /* HashParams, core/params.h:16-24        */
  uint32_t bucket_count;
  uint32_t bucket_size;
  uint32_t entry_count;           /* must equal bucket_count * bucket_size  */
  uint32_t linked_list_size;
  uint32_t value_capacity;        /* block pool size                        */
} mh_hash_params;

typedef struct mh_mesh_params {   /* MeshParams, core/params.h:26-29        */
  uint32_t max_vertex_count;
  uint32_t max_triangle_count;
} mh_mesh_params;

typedef struct mh_volume_params { /* VolumeParams, core/params.h:31-40      */
  float voxel_size;
  float truncation_distance_scale;
  float truncation_distance;
  float sdf_upper_bound;
  uint32_t weight_sample;
  uint32_t weight_upper_bound;    /* unused by the reference as well        */
} mh_volume_params;

typedef struct mh_sensor_params { /* SensorParams, core/params.h:62-74      */
  float fx, fy, cx, cy;
  uint32_t width, height;
  float min_depth_range;
  float max_depth_range;
  float range_factor;             /* metres per raw depth unit              */
} mh_sensor_params;

typedef struct mh_raycaster_params { /* RayCasterParams, core/params.h:42-57 */
  float fx, fy, cx, cy;
  uint32_t width, height;
  float min_raycast_depth;
  float max_raycast_depth;
  float raycast_step;
  float sample_sdf_threshold;
  float sdf_threshold;
  int32_t enable_gradients;
} mh_raycaster_params;

/* Marching-cubes thresholds.  The reference hard-codes them
 * (meshing/marching_cubes.cu:20-21,96-99,124-128); they are parameters here.
 * mh_mc_params_default() returns the reference's values. */
typedef struct mh_mc_params {
  float sdf_threshold;   /* |sdf| of every corner must be <= this (0.20)     */
  float rho_min;         /* a/(a+b) of every corner must be >= this (0.1)    */
  float weight_min;      /* inv_sigma2 of every corner >= this ((1/vs)^2);
                            a negative value selects (1/voxel_size)^2         */
  double snap_distance;  /* vertex snaps to an endpoint with |sdf| < this
                            (0.008; |sdf| is widened to double and compared
                            with it, as the reference's double literal)     */
  int32_t enable_sdf_gradient; /* 0: flat per-triangle normals (mesh.h:98)  */
} mh_mc_params;

typedef struct mh_status {
  uint32_t flags;            /* MH_STATUS_* bits, sticky until reset         */
  uint32_t dropped_inserts;  /* blocks lost to a full linked-list window     */
  uint32_t dda_cap_hits;     /* pixels whose DDA hit the 1024-step cap       */
  uint32_t list_inserts;     /* blocks placed in an overflow linked list     */
  uint32_t block_count;      /* allocated blocks                             */
  uint32_t candidate_count;  /* blocks in the current worklist               */
  uint32_t vertex_count;     /* live vertices                                */
  uint32_t triangle_count;   /* live triangles                               */
  /* work counters accumulated since the last reset */
  uint64_t voxel_updates;      /* voxels whose fused state changed          */
  uint64_t blocks_integrated;  /* worklist blocks visited by fusion         */
  uint64_t mc_vertex_writes;   /* vertex records written by MC              */
  uint64_t mc_triangle_writes; /* triangles (re)assigned by MC              */
  uint64_t mc_blocks;          /* worklist blocks meshed                    */
  uint64_t ray_samples;        /* valid trilinear samples of the raycaster  */
  uint64_t ray_blocks;         /* distinct blocks sampled, summed over
                                  raycasts run with MH_DIAG_RAY_BLOCKS      */
} mh_status;

typedef struct mh_volume mh_volume;

/* ---- lifecycle ----------------------------------------------------------
 * HashTable::{Resize,Reset,Free}   core/hash_table.h:25-33, hash_table.cu:57-143
 * BlockArray::{Resize,Reset,Free}  core/block_array.h:14-28
 * EntryArray::Resize               core/entry_array.h:13-23
 * Mesh::{Resize,Reset,Free}        core/mesh.h:21-24, core/mesh.cu:74-137
 * Sensor::Sensor                   sensor/rgbd_sensor.cu:15-49
 * enable_psdf allocates the Beta-inlier fields (a, b) of PSDF fusion.     */
int mh_volume_create(const mh_hash_params* hash, const mh_volume_params* vol,
                     const mh_mesh_params* mesh, const mh_sensor_params* sensor,
                     int enable_psdf, mh_volume** out);
int mh_volume_destroy(mh_volume* v);
int mh_volume_reset(mh_volume* v);
int mh_volume_set_stream(mh_volume* v, void* hip_stream);
int mh_synchronize(mh_volume* v);
int mh_get_status(mh_volume* v, mh_status* out);
/* Diagnostics, off by default; neither changes a result.  The reference
 * times its stages with host timers (util/timer.h, written out by
 * engine/logging_engine.cc:61-93); here:
 * MH_DIAG_KERNEL_TIMES: each kernel the volume's stage entry points launch
 *   is bracketed by HIP events on the volume's stream; mh_kernel_times sums
 *   them per kernel name (launches, total ms) and synchronises.
 * MH_DIAG_RAY_BLOCKS: mh_raycast records which blocks its lookups find and
 *   adds their number to mh_status.ray_blocks (one more kernel per call). */
#define MH_DIAG_KERNEL_TIMES 1u
#define MH_DIAG_RAY_BLOCKS 2u
int mh_volume_set_diagnostics(mh_volume* v, uint32_t flags);
typedef struct mh_kernel_time {
  char name[48];
  uint64_t launches;
  double total_ms;
} mh_kernel_time;
int mh_kernel_times(mh_volume* v, uint32_t capacity, mh_kernel_time* out,
                    uint32_t* count, int reset);
void mh_mc_params_default(mh_mc_params* out);
const char* mh_last_error(void);

/* ---- frame input: Sensor::Process + set_transform ------------------------
 * sensor/rgbd_sensor.cu:118-137, sensor/preprocess.cu:25-42 (depth
 * conversion: d = range_factor * (int16)raw, valid iff min <= d <= max, else
 * -inf), sensor/rgbd_sensor.h:51-54 (cTw = inverse(wTc), FP32).
 * wTc is camera->world, row-major m11..m44.  The PSDF inlier ratio is reset to
 * 0.1 (sensor/preprocess.cu:10-23) for every frame.  The conversion is fused
 * into the next mh_alloc_blocks (one pass over the image); a stage that
 * reads the frame before any allocation converts it first.  A device depth
 * buffer must therefore stay valid until that stage has run.              */
int mh_frame_upload_depth_u16(mh_volume* v, const uint16_t* depth_host,
                              const float wTc[16]);
int mh_frame_depth_u16_device(mh_volume* v, const uint16_t* depth_dev,
                              const float wTc[16]);
int mh_frame_set_pose(mh_volume* v, const float wTc[16]);
/* Colour of the current frame (Sensor::Process's ConvertColorFormat,
 * sensor/preprocess.cu:44-58): 8-bit BGRA, W*H*4 bytes.  A pixel is valid
 * iff its B, G and R are all non-zero; fusion then uses
 * uchar(255 * (c / 255.0f)) per channel (update_simple.cu:68-70), an invalid
 * pixel contributes (0, 0, 0).  Applies to the frame uploaded last; a depth
 * upload clears it, and without colour TSDF fusion uses (0, 255, 0)
 * (update_simple.cu:71-73).                                                */
int mh_frame_upload_color_bgra8(mh_volume* v, const uint8_t* bgra_host);
int mh_frame_color_bgra8_device(mh_volume* v, const uint8_t* bgra_dev);
/* Converted metric depth of the current frame (device, W*H floats). */
int mh_frame_depth_device_ptr(mh_volume* v, const float** depth_dev);

/* ---- mapping ------------------------------------------------------------ */
/* AllocBlockArray, mapping/allocate.h:16-20 / allocate.cu:8-125.  Lossless:
 * every block the per-pixel DDA reaches inside the frustum is inserted.   */
int mh_alloc_blocks(mh_volume* v);
/* CollectBlocksInFrustum / CollectAllBlocks, core/collect_block_array.h:16-31.
 * count_host may be NULL (no host synchronisation).                        */
int mh_collect_visible(mh_volume* v, uint32_t* count_host);
int mh_collect_all(mh_volume* v, uint32_t* count_host);
/* UpdateBlocksSimple, mapping/update_simple.h:21-27.                       */
int mh_integrate_tsdf(mh_volume* v);
/* PredictOutlierRatio, UpdateBlocksBayesian, mapping/update_bayesian.h:33-47 */
int mh_predict_inlier(mh_volume* v);
int mh_integrate_psdf(mh_volume* v);
/* Inlier ratio image of the current frame (device, W*H floats). */
int mh_inlier_ratio_device_ptr(mh_volume* v, float** inlier_dev);
/* Device storage of a voxel field: value_capacity * 512 values, block-pool
 * index major, voxel x + 8 y + 64 z within a block (the BlockArray's device
 * memory, core/block_array.h:14-28).  For interop, e.g. viewing the SDF as a
 * framework tensor; writes through it change the volume.  count (may be
 * NULL) receives the number of values.  PSDF fields need enable_psdf.     */
#define MH_FIELD_SDF 0     /* float */
#define MH_FIELD_WEIGHT 1  /* float: weight (TSDF) or inv_sigma2 (PSDF) */
#define MH_FIELD_COLOR 2   /* uint32: R | G << 8 | B << 16 */
#define MH_FIELD_BETA_A 3  /* float */
#define MH_FIELD_BETA_B 4  /* float */
int mh_volume_field_device_ptr(mh_volume* v, int field, void** field_dev,
                               uint64_t* count);

/* ---- block recycling: mapping/recycle.h:11-35 ---------------------------
 * Each works on the current worklist, as the reference (MainEngine::Meshing
 * and ::Recycle, engine/main_engine.cc:156-192).
 * CollectLowSurfelBlocks (recycle.cu:81-115): a block with < 10 inner and
 *   no boundary surfels (counted by the last mh_marching_cubes) ages by one,
 *   others are renewed (life 3, core/block.h:13); a block whose life is over
 *   is flagged when no allocated face neighbour is still alive.  Ageing and
 *   flagging are separate passes (deterministic).
 * StarveOccupiedBlockArray (recycle.cu:15-24): weight -= 1, floored at 0.
 * CollectGarbageBlockArray (recycle.cu:27-77): flag blocks whose observed
 *   voxels all have |sdf| >= the truncation at 5 m, or with none observed.
 * RecycleGarbageBlockArray (recycle.cu:117-179, 231-260): flagged blocks
 *   lose their triangles and the vertices they own left unreferenced; a
 *   flagged block with no vertex left is removed from the hash table
 *   (HashTable::FreeEntry, hash_table.h:183-255) and cleared.  freed_host
 *   (may be NULL, synchronises): blocks freed.                            */
int mh_collect_low_surfel_blocks(mh_volume* v);
int mh_starve_blocks(mh_volume* v);
int mh_collect_garbage_blocks(mh_volume* v);
int mh_recycle_garbage_blocks(mh_volume* v, uint32_t* freed_host);

/* ---- meshing: MarchingCubes, meshing/marching_cubes.h:18-24 ------------- */
int mh_marching_cubes(mh_volume* v, const mh_mc_params* p);

/* ---- raycasting: RayCaster::Cast, visualization/ray_caster.h:37-41 -------
 * cTw is world->camera (the reference's argument); wTc = inverse(cTw) in
 * FP32 as ray_caster.cu:195.  Outputs (device, caller-owned, W*H pixels):
 * depth_dev metric depth (-inf: no hit), vertex_dev camera-space point
 * (3 floats per pixel), normal_dev camera-space normal (3 floats per pixel).
 * Any output may be NULL.                                                  */
int mh_raycast(mh_volume* v, const mh_raycaster_params* p, const float cTw[16],
               float* depth_dev, float* vertex_dev, float* normal_dev);

/* ---- denoising: PrimalDualInit / PrimalDualIterate, optimize/primal_dual.h:14-29
 * Runs over the current worklist (collect_all first, as the reference's
 * driver does).  energy_host (3 doubles, may be NULL): data term, TV term and
 * err_data*lambda/2 + err_tv of the iteration's pass 1 (primal_dual.cu:207-214).
 * init builds a face-neighbour table over the worklist; iterate returns
 * MH_ERR_STATE if any call has rebuilt or cleared the worklist since (alloc,
 * collect, recycle, reset, a batch run): call init again after those.     */
int mh_primal_dual_init(mh_volume* v);
int mh_primal_dual_iterate(mh_volume* v, float lambda, float sigma, float tau,
                           double* energy_host);

/* ---- canonical dumps and restore (host buffers) --------------------------
 * Blocks are returned sorted by (x, y, z).  Pass NULL buffers to query the
 * count only.  capacity = number of blocks the buffers hold.
 * pos: 3 int32 per block; fields: 512 values per block (voxel index
 * x + 8y + 64z, geometry/geometry_helper.h:113-120).  color is RGBA8.     */
int mh_dump_blocks(mh_volume* v, uint32_t capacity, int32_t* pos_host,
                   float* sdf_host, float* weight_host, uint32_t* color_host,
                   float* a_host, float* b_host, uint32_t* count_host);
/* Upload n blocks (same layout as mh_dump_blocks) into an empty volume.
 * a_host / b_host may be NULL.  The blocks are inserted into the hash
 * table with the same insertion rule as mh_alloc_blocks.                   */
int mh_load_blocks(mh_volume* v, uint32_t n, const int32_t* pos_host,
                   const float* sdf_host, const float* weight_host,
                   const uint32_t* color_host, const float* a_host,
                   const float* b_host);
/* Hash layout: one record per occupied entry, sorted by entry index:
 * entry index, home bucket, in-bucket flag (1) or list (0), pos xyz.
 * rec_host holds 6 int32 per record.                                       */
int mh_dump_hash(mh_volume* v, uint32_t capacity, int32_t* rec_host,
                 uint32_t* count_host);
/* Current worklist as block positions (3 int32 each), sorted.              */
int mh_dump_candidates(mh_volume* v, uint32_t capacity, int32_t* pos_host,
                       uint32_t* count_host);
/* Flags of the last low-surfel or garbage pass (EntryArray::flag,
 * core/entry_array.h:29-31) as (x, y, z, flag) int32 rows in position order.     */
int mh_dump_candidate_flags(mh_volume* v, uint32_t capacity, int32_t* rec_host,
                            uint32_t* count_host);
/* Mesh in canonical form.  Triangles: 12 int32 each -- the owner voxel
 * (x, y, z) and edge axis of each of its three vertices, in table order;
 * a vertex whose slot is not allocated has x = y = z = INT32_MIN --
 * sorted by (cube voxel, slot).  tri_cube_host (optional): 4 int32 per
 * triangle: cube voxel x, y, z and triangle slot.  Vertices: one record per
 * live vertex, key (4 int32: owner voxel xyz, axis) + 10 floats (pos xyz,
 * normal xyz, color rgb, radius) + ref count, sorted by key.              */
int mh_dump_mesh(mh_volume* v, uint32_t tri_capacity, int32_t* tri_host,
                 int32_t* tri_cube_host, uint32_t vtx_capacity,
                 int32_t* vtx_key_host, float* vtx_val_host,
                 int32_t* vtx_ref_host, uint32_t* tri_count_host,
                 uint32_t* vtx_count_host);
/* Cube indices of every voxel of every allocated block, block order as
 * mh_dump_blocks: 512 (curr | prev << 8) per block.                        */
int mh_dump_cube_indices(mh_volume* v, uint32_t capacity, uint16_t* cube_host,
                         uint32_t* count_host);
/* Primal-dual variables in mh_dump_blocks order: sdf0, sdf_bar, w_pd, p xyz
 * (6 floats per voxel, 512 voxels per block, field-major per block) and the
 * mask (1 byte per voxel).                                                  */
int mh_dump_primal_dual(mh_volume* v, uint32_t capacity, float* pdv_host,
                        uint8_t* mask_host, uint32_t* count_host);

/* FP32 pose of the current frame: wTc as given, cTw its inverse as used by
 * every decision (sensor/rgbd_sensor.h:51-54).  Either output may be NULL. */
int mh_frame_get_pose(mh_volume* v, float wTc[16], float cTw[16]);

/* ---- integration autotuning (no counterpart in the reference) ------------
 * UpdateBlocksSimple's tiling -- lanes per workgroup, worklist blocks fused
 * per workgroup (hence voxels per lane) and persistent workgroups per CU --
 * is chosen by timing every candidate on the current worklist.  Results do
 * not depend on the tiling: each voxel is updated by one lane with the same
 * arithmetic.  mh_tune_integrate runs after mh_collect_* and before
 * mh_integrate_tsdf; it saves and restores the worklist's fused state and
 * the counters around every trial, so it changes no output.
 * On a PSDF volume it tunes mh_integrate_psdf instead: tiling t picks 128,
 * 256 or 512 lanes per workgroup (one block at a time), times 5 grids per CU
 * (15 candidates); mh_set_integrate_tiling takes the same indices there.   */
typedef struct mh_tune_result {
  int32_t tiling;               /* index of the chosen tiling               */
  int32_t workgroup_size;       /* lanes per workgroup                      */
  int32_t blocks_per_workgroup; /* worklist blocks per workgroup at a time  */
  int32_t voxels_per_lane;
  int32_t workgroups_per_cu;    /* persistent grid = CUs x this             */
  int32_t candidates;           /* configurations timed                     */
  float best_ms;                /* fastest trial of the chosen tiling       */
  float default_ms;             /* fastest trial of the default tiling      */
  int32_t early_loads;          /* 1: voxel state loaded with the depth
                                   gathers, block keys fetched per wave    */
  int32_t ray_shape;            /* mh_batch_run with MH_RUN_TUNE and
                                   MH_RUN_RAYCAST: the raycast tile shape
                                   chosen (mh_tune_raycast); otherwise -1   */
} mh_tune_result;
#define MH_NUM_TILINGS 18
#define MH_NUM_TUNE_GRIDS 5 /* persistent-grid sizes timed per tiling */
int mh_tune_integrate(mh_volume* v, uint32_t reps, mh_tune_result* out);
int mh_set_integrate_tiling(mh_volume* v, int tiling, int workgroups_per_cu);

/* ---- raycast tile shape (no counterpart in the reference, whose CastKernel
 * runs 16 x 16-thread blocks, ray_caster.cu:190-224) ----------------------
 * Each 64-lane workgroup of mh_raycast marches a tile of W x 64/W pixels,
 * W = 8, 16, 4 or 32 (shapes 0..3); every shape gives identical results.
 * mh_tune_raycast times the march of each shape on one raycast at cTw
 * (outputs to scratch, work counters restored, reps trials each, 0 = 3)
 * and keeps the fastest for the volume's later raycasts (shape 0 unless
 * another is at least 3 % faster).  Synchronises.                         */
typedef struct mh_tune_shape_result {
  int32_t shape;        /* index of the chosen shape                        */
  int32_t tile_width;   /* pixels                                           */
  int32_t tile_height;
  int32_t candidates;   /* shapes timed                                     */
  float best_ms;        /* fastest trial of the chosen shape                */
  float default_ms;     /* fastest trial of shape 0 (8 x 8)                 */
} mh_tune_shape_result;
#define MH_NUM_RAY_SHAPES 4
int mh_tune_raycast(mh_volume* v, const mh_raycaster_params* p, const float cTw[16],
                    uint32_t reps, mh_tune_shape_result* out);
int mh_set_raycast_shape(mh_volume* v, int shape);

/* ---- marching-cubes launch shape (no counterpart in the reference, whose
 * SurfelExtractionKernel runs one 512-thread block per worklist block,
 * marching_cubes.cu:339-381) ---------------------------------------------
 * The surfel pass runs 512-lane workgroups over the worklist, persistent:
 * shape k sets the workgroups per CU (0: as many as are resident) and the
 * quarters of the worklist handed out statically (the rest on demand);
 * MH_NUM_MC_SHAPES shapes, shape 0 the default, all with the same mesh.
 * Meshing changes the volume, so tuning runs online: mh_tune_marching_cubes
 * makes the next reps x 2 (MH_NUM_MC_SHAPES - 1) + 1 calls of
 * mh_marching_cubes alternate shape 0 with each other shape in turn (timed
 * with HIP events per worklist block, since frames differ in size; every
 * call still meshes as asked).  Each candidate trial is compared with the
 * faster of the shape-0 trials just before and after it, since frames also
 * differ in content; the call after the last trial keeps the candidate
 * whose worst such ratio is least, if it is at most 0.97, else shape 0.
 * mh_marching_cubes_tuning reports the progress (done = 0) or the result;
 * mh_set_marching_cubes_shape sets a shape (and ends any tuning).         */
typedef struct mh_tune_mc_result {
  int32_t done;               /* 1 once the shape is chosen                 */
  int32_t shape;              /* the chosen (or current) shape              */
  int32_t workgroups_per_cu;  /* 0: all resident                            */
  int32_t static_quarters;    /* of the worklist handed out statically      */
  int32_t candidates;
  float best_ms;              /* default_ms x the chosen shape's worst ratio
                                 (default_ms for shape 0)                    */
  float default_ms;           /* median time per worklist block (ms) of the
                                 shape-0 trials                              */
  float blocks_per_ms;        /* 1 / best_ms                                 */
} mh_tune_mc_result;
#define MH_NUM_MC_SHAPES 9
int mh_tune_marching_cubes(mh_volume* v, uint32_t reps);
int mh_marching_cubes_tuning(mh_volume* v, mh_tune_mc_result* out);
int mh_set_marching_cubes_shape(mh_volume* v, int shape);

/* ---- localisation: PointToSurface, localizing/point_to_psdf.h -----------
 * Normal equations of the current frame's depth against the volume at the
 * camera->world pose wTc (PointToSurfaceKernel, point_to_psdf.cu:19-82):
 * every pixel with odd x and odd y and a depth that is finite, non-zero and
 * below sdf_upper_bound, whose world point has an SDF value and gradient
 * (GetSpatialValue / GetSpatialSDFGradient, spatial_query.h:23-97), adds
 * J J^T to A, d J to b, d^2 to err and 1 to count, with d the interpolated
 * SDF and J = (g, p x g) for the tangent (translation, rotation).
 * MH_P2S_REFERENCE keeps the reference's J, whose third entry is g.x
 * (point_to_psdf.cu:69-74); MH_P2S_CORRECTED uses g.z there.  Sums are
 * FP64 in a fixed order (the reference: float atomics).  Synchronises.   */
#define MH_P2S_REFERENCE 0
#define MH_P2S_CORRECTED 1
typedef struct {
  double A[36];    /* row-major 6x6 */
  double b[6];
  double err;      /* sum of d^2 */
  uint32_t count;  /* pixels used */
} mh_point_to_psdf_result;
int mh_point_to_psdf(mh_volume* v, const float wTc[16], int jacobian,
                     mh_point_to_psdf_result* out);
/* MainEngine::Localizing (engine/main_engine.cc:28-98): `iterations` steps
 * of (A + lambda I) dxi = -b, wTc <- exp(dxi) wTc, stopping early when no
 * pixel is used; the reference calls it with lambda = 1e5 and 20 steps
 * (app/slam.cc:94).  wTc is updated in place and becomes the frame's pose
 * (Sensor::set_transform).  err_trace / count_trace (may be NULL) get each
 * step's err and count.  The solve and exponential are FP64.            */
int mh_localize(mh_volume* v, uint32_t iterations, float lambda, int jacobian, float wTc[16],
                double* err_trace, uint32_t* count_trace, uint32_t* iterations_done);

/* ---- mesh compaction and PLY output -------------------------------------
 * CompressMesh (visualization/compress_mesh.cu:137-172) over the current
 * worklist (mh_collect_all + mh_marching_cubes first for the whole mesh, as
 * MainEngine::FinalLog, engine/main_engine.cc:299-311): every live triangle
 * whose three vertex slots are allocated, with the vertices they use
 * renumbered densely.  Triangles come in worklist x voxel x table-slot order,
 * vertices in vertex-pool order.  Synchronises the volume's stream.       */
int mh_compress_mesh(mh_volume* v, uint32_t* vertex_count, uint32_t* triangle_count);
/* The compacted mesh: 3 floats per vertex (position, normal, colour) and 3
 * int32 vertex indices per triangle; any buffer may be NULL.              */
int mh_compact_mesh_get(mh_volume* v, uint32_t vertex_capacity, float* pos_host,
                        float* normal_host, float* color_host,
                        uint32_t triangle_capacity, int32_t* tri_host);
/* SavePly (io/mesh_writer.cc:75-150): ASCII PLY of the compacted mesh.
 * Colour channels are int(255 c), with a NaN channel written as 0.        */
int mh_save_ply(mh_volume* v, const char* path);

/* ---- differentiable raycast (no counterpart in the reference) ------------
 * Loss L = 1/2 sum over pixels of (depth - obs)^2, over the pixels where
 * mh_raycast (same params and cTw) hits and obs_depth_dev (W*H metres,
 * device) holds a finite positive value.  The backward pass recomputes the
 * march and differentiates the discrete algorithm: the bracketing samples
 * and the three regula-falsi steps, through the trilinear interpolant.
 * Pose: xi = (omega, upsilon) perturbs the camera pose on the left,
 * wTc <- exp(xi^) wTc, with wTc = inverse(cTw) (FP32).  grad_sdf_dev (may be
 * NULL): capacity*512 floats laid out as the volume's SDF (pool index major),
 * zeroed by the caller; dL/dsdf is added to it (atomically, so sums of many
 * pixels agree with FP64 to float rounding).  Per-pixel values are FP32;
 * the sums below are FP64 in a fixed order.                                */
typedef struct mh_raycast_grad {
  double loss;          /* L                                                 */
  double grad[6];       /* dL/dxi                                            */
  double hessian[36];   /* sum J^T J (Gauss-Newton), J = d depth / d xi      */
  uint32_t pixels;      /* pixels in the sum                                 */
  uint32_t outliers;    /* hit + observation, left out by a residual gate   */
  uint32_t observed;    /* pixels with a finite positive observation, hit or
                           not (the pose refinement charges the lost ones)   */
} mh_raycast_grad;
int mh_raycast_backward(mh_volume* v, const mh_raycaster_params* p,
                        const float cTw[16], const float* obs_depth_dev,
                        float* grad_sdf_dev, mh_raycast_grad* out);
/* Allocation-free form of mh_raycast_backward (no counterpart in the
 * reference).  For callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc:
 * it calls no hipMalloc, hipFree or allocating library routine.  Its
 * scratch (tile ranges, per-workgroup partial sums, the result row) is one
 * static device array of ~5.2 MB per GPU, held until the process ends, so
 * images are limited to 1280 x 1024 (larger: MH_ERR_CAPACITY).  The
 * volume's device must be current.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race;
 * serialise them.  Same arguments, results and synchronisation as
 * mh_raycast_backward.                                                     */
int mh_raycast_backward_noalloc(mh_volume* v, const mh_raycaster_params* p,
                                const float cTw[16], const float* obs_depth_dev,
                                float* grad_sdf_dev, mh_raycast_grad* out);
/* Pose refinement: Levenberg-damped Gauss-Newton over xi, iterations steps
 * from wTc_init, on a truncated quadratic over every observed pixel: r^2 / 2
 * for a pixel within the residual gate, gate^2 / 2 for one gated out or
 * lost (no hit; a ray close to grazing the surface, its depth moving more
 * than 3x a camera translation along some axis).  The gate is 0.20 m for the
 * first 40 % of the steps, then 0.10 m, then 0.05 m for the last 30 %; a
 * step is kept only if it lowers the cost.  The result is wTc_init if that
 * costs less at the final gate.  cost_trace (may be NULL) receives the cost
 * per observed pixel before the first and after every step (iterations + 1
 * values, each at that step's gate).                                       */
int mh_refine_pose(mh_volume* v, const mh_raycaster_params* p,
                   const float wTc_init[16], const float* obs_depth_dev,
                   uint32_t iterations, float wTc_out[16], double* cost_trace);
/* A caller's per-voxel device field (capacity*512 floats, pool-index major,
 * e.g. grad_sdf_dev) in mh_dump_blocks order.                             */
int mh_dump_voxel_field(mh_volume* v, const float* field_dev, uint32_t capacity,
                        float* out_host, uint32_t* count_host);

/* ---- batched scenes (no counterpart in the reference) --------------------
 * Many independent scenes (uneven block counts and voxel sizes, one common
 * image size) fused per stage: every stage of a frame is one launch over
 * all scenes (pixel stages index scenes by grid z, block stages run over
 * the concatenated (scene, block) worklist).  Each scene's results are
 * those of running the same frames through the single-volume entry points
 * (mh_frame_upload_depth_u16, mh_alloc_blocks, mh_collect_visible,
 * mh_integrate_tsdf, mh_marching_cubes, mh_raycast with the frame's cTw);
 * mh_batch_volume exposes each scene's volume for status and dumps.  TSDF
 * scenes only.  The batch owns a HIP stream (replaceable, never NULL).     */
typedef struct mh_batch mh_batch;
enum {
  MH_RUN_MESH = 1u << 0,    /* marching cubes after fusion                 */
  MH_RUN_RAYCAST = 1u << 1, /* raycast after fusion (mh_batch_set_raycast) */
  MH_RUN_GRAPH = 1u << 2,   /* replay the captured per-frame HIP graph
                               (captured on first use or shape change)      */
  MH_RUN_TUNE = 1u << 3     /* autotune integration on this frame first    */
};
/* Arrays of n_scenes parameter structs; all sensors share width/height.    */
int mh_batch_create(uint32_t n_scenes, const mh_hash_params* hash,
                    const mh_volume_params* vol, const mh_mesh_params* mesh,
                    const mh_sensor_params* sensor, mh_batch** out);
int mh_batch_destroy(mh_batch* b);
int mh_batch_set_stream(mh_batch* b, void* hip_stream);
int mh_batch_volume(mh_batch* b, uint32_t scene, mh_volume** out);
/* Depth of every scene (n*W*H uint16, scene-major; host memory, or device
 * memory that stays valid until the run completes when on_device != 0) and
 * the poses wTc (n*16 floats).                                             */
int mh_batch_frame_u16(mh_batch* b, const uint16_t* depth, int on_device,
                       const float* wTc);
/* Raycaster params per scene and device outputs (n*W*H depth, n*W*H*3
 * vertex and normal floats, scene-major; any may be NULL).                 */
int mh_batch_set_raycast(mh_batch* b, const mh_raycaster_params* p,
                         float* depth_dev, float* vertex_dev, float* normal_dev);
int mh_batch_set_integrate_tiling(mh_batch* b, int tiling, int workgroups_per_cu);
/* The marching-cubes shape of every scene of a batch (as chosen by
 * mh_tune_marching_cubes on a representative volume); ends a tuning.       */
int mh_batch_set_marching_cubes_shape(mh_batch* b, int shape);
/* Online tuning of the batch's marching-cubes shape over its graph replays
 * (no counterpart in the reference; a replayed chain can run faster with
 * another shape than eager launches do): the next
 * reps x 2 (MH_NUM_MC_SHAPES - 1) + 1 replays with MH_RUN_MESH alternate
 * shape 0 with each other shape, each timed whole (HIP events around the
 * replay), and the replay after them keeps the shape chosen by the rule of
 * mh_tune_marching_cubes.  best_ms and default_ms are then ms per replay
 * and blocks_per_ms is 0.  Every replay still runs the full chain.         */
int mh_batch_tune_marching_cubes(mh_batch* b, uint32_t reps);
int mh_batch_marching_cubes_tuning(mh_batch* b, mh_tune_mc_result* out);
/* One frame of every scene: convert -> alloc -> collect visible -> TSDF
 * [-> marching cubes (mc: n structs, or NULL for defaults)] [-> raycast].
 * tune_out receives the tuning result with MH_RUN_TUNE (may be NULL).     */
int mh_batch_run(mh_batch* b, const mh_mc_params* mc, uint32_t flags,
                 mh_tune_result* tune_out);
int mh_batch_synchronize(mh_batch* b);

/* ---- synthetic depth sequences (seeded, deterministic, CPU) --------------
 * Analytic scenes: an axis-aligned room seen from inside plus spheres and
 * boxes.  kind: 0 = circular trajectory around the room centre with 3
 * spheres; 1 = spline trajectory with 8 spheres/boxes.  The scene and the
 * trajectory depend on seed only.                                          */
typedef struct mh_synth_spec {
  uint64_t seed;
  int32_t kind;
  uint32_t width, height;
  float fx, fy, cx, cy;
  uint32_t frames;
  float hole_fraction;     /* pixels set to 0                              */
  float outlier_fraction;  /* valid pixels replaced by U[0.3, 5] m         */
  int32_t noise;           /* 1: sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2    */
} mh_synth_spec;
/* Pose of frame f (wTc, row-major).                                         */
int mh_synth_pose(const mh_synth_spec* s, uint32_t frame, float wTc[16]);
/* Depth of frame f in mm (uint16, W*H, 0 = no measurement).               */
int mh_synth_depth(const mh_synth_spec* s, uint32_t frame, uint16_t* depth_host);
/* Exact (noise-free) z-depth in metres, for tests (+inf: no hit).          */
int mh_synth_depth_exact(const mh_synth_spec* s, uint32_t frame,
                         const float wTc[16], double* depth_host);
/* Colour of frame f: 8-bit BGRA per pixel (W*H*4 bytes), a procedural
 * texture of the surface point, 0 where there is no surface.               */
int mh_synth_color(const mh_synth_spec* s, uint32_t frame, uint8_t* bgra_host);

#ifdef __cplusplus
}
#endif
#endif /* MESHHASHING_H */
