// mh_kernels.h -- host launchers of the HIP kernels (one per .hip file).
#pragma once
#include "mh_core.h"

namespace mh {

// mapping.hip
void launch_reset(const View& V, hipStream_t s, int grid);
void launch_convert_depth(const View& V, const uint16_t* raw, hipStream_t s);
void launch_alloc(const View& V, hipStream_t s);
void launch_collect(const View& V, int frustum, hipStream_t s, int grid);
void launch_integrate_tsdf(const View& V, hipStream_t s, int grid);
void launch_predict_inlier(const View& V, hipStream_t s, int grid);
void launch_integrate_psdf(const View& V, hipStream_t s, int grid);
void launch_load_insert(const View& V, const uint64_t* keys, uint32_t n, hipStream_t s);
void launch_load_scatter(const View& V, const uint64_t* keys, uint32_t n, const float* sdf,
                         const float* weight, const uint32_t* color, const float* a,
                         const float* b, hipStream_t s);

// meshing.hip
struct McConst {
  float sdf_threshold;
  float rho_min;
  float weight_min;
  double snap;
  int flat_normals;
  uint32_t pass;  // 1, 2, ...: per-call stamp for single-writer vertex payloads
};
void launch_marching_cubes(const View& V, const McConst& p, hipStream_t s, int grid);

// raycast.hip
struct RayConst {
  Cam cam;
  float min_depth, max_depth, step;
  float sample_sdf_threshold, sdf_threshold;
  int enable_gradients;
  float inv_vs;  // 1 / voxel_size (fast empty-space test only)
  M4 cTw, wTc;
};
void launch_raycast(const View& V, const RayConst& p, float* depth, float* vertex, float* normal,
                    hipStream_t s);

// primal_dual.hip
void launch_pd_neighbors(const View& V, int32_t* nbr, hipStream_t s, int grid);
void launch_pd_init(const View& V, const int32_t* nbr, hipStream_t s, int grid);
void launch_pd_iterate(const View& V, const int32_t* nbr, float lambda, float sigma, float tau,
                       double* partials, uint32_t npartials, double* energy, hipStream_t s,
                       int grid);

}  // namespace mh
