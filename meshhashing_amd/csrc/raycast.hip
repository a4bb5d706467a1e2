// raycast.hip -- fixed-step ray march over the hashed TSDF with regula-falsi
// refinement.
//
// Reference: visualization/ray_caster.cu:16-150 (CastKernel),
// geometry/isosurface_intersection.h:13-55 (LinearIntersection,
// BisectionIntersection), geometry/spatial_query.h:23-97 (trilinear value and
// gradient), geometry/voxel_query.h:102-129 (world-point voxel lookup).
//
// One lane per pixel, 8x8 pixels per wave.  Samples, the crossing test and
// the three regula-falsi steps follow the reference's FP32 expressions; each
// lane keeps the last two block lookups (including misses) so that runs of
// samples in one block, or in one unallocated block, cost no hash probe.
// The depth output is metric z (t * dir_c.z); the reference colour-maps it
// (ray_caster.cu:96-98) and keeps metric depth only in the vertex image.
#include "mh_kernels.h"

namespace mh {

struct BlockCache {
  uint64_t k0, k1;
  int32_t p0, p1;
};

__device__ __forceinline__ int32_t cached_find(const View& V, BlockCache& c, i3 b) {
  const uint64_t k = pack_key(b);
  if (k == c.k0) return c.p0;
  if (k == c.k1) {
    const uint64_t tk = c.k0;
    const int32_t tp = c.p0;
    c.k0 = c.k1;
    c.p0 = c.p1;
    c.k1 = tk;
    c.p1 = tp;
    return c.p0;
  }
  const int32_t p = coord_ok(b) ? hash_find(V, b) : kFreePtr;
  c.k1 = c.k0;
  c.p1 = c.p0;
  c.k0 = k;
  c.p0 = p;
  return p;
}

__device__ __forceinline__ int vox1(float w, float vs) {
  const float p = w / vs;
  return f2i(p + (float)sgn(p) * 0.5f);
}

// Lowest-corner voxel of a sample along one axis from a multiply by 1/vs
// instead of the IEEE division; accepted only when the quotient is more than
// 1e-3 voxel away from every rounding tie (and |q| < 2048, where the
// multiply is within 2 ulp of the quotient), so it equals the exact result.
__device__ __forceinline__ bool fast_corner(float pc, float inv_vs, int* c) {
  const float q = pc * inv_vs;
  const float aq = fabsf(q);
  if (!(aq < 2048.0f) || aq < 1e-3f) return false;
  const float f = aq - floorf(aq);
  if (fabsf(f - 0.5f) < 1e-3f) return false;
  *c = f2i(q + (float)sgn(q) * 0.5f);
  return true;
}

// GetSpatialValue restricted to the SDF (the only field the march uses).
__device__ __forceinline__ bool ray_sdf(const View& V, BlockCache& cache, f3 pos, float inv_vs,
                                        float* out) {
  const float vs = V.geo.voxel_size;
  const float hv = 0.5f * vs;
  const float pcx = pos.x - hv, pcy = pos.y - hv, pcz = pos.z - hv;
  {
    // empty-space fast path: the lowest corner's block is missing
    int c0x, c0y, c0z;
    if (fast_corner(pcx, inv_vs, &c0x) && fast_corner(pcy, inv_vs, &c0y) &&
        fast_corner(pcz, inv_vs, &c0z) &&
        cached_find(V, cache, i3{voxel_to_block1(c0x), voxel_to_block1(c0y), voxel_to_block1(c0z)}) < 0)
      return false;
  }
  const float qx = pos.x / vs, qy = pos.y / vs, qz = pos.z / vs;
  const float rx = qx - floorf(qx), ry = qy - floorf(qy), rz = qz - floorf(qz);
  const int cx[2] = {vox1(pcx + 0.0f, vs), vox1(pcx + vs, vs)};
  const int cy[2] = {vox1(pcy + 0.0f, vs), vox1(pcy + vs, vs)};
  const int cz[2] = {vox1(pcz + 0.0f, vs), vox1(pcz + vs, vs)};
  const int bx[2] = {voxel_to_block1(cx[0]), voxel_to_block1(cx[1])};
  const int by[2] = {voxel_to_block1(cy[0]), voxel_to_block1(cy[1])};
  const int bz[2] = {voxel_to_block1(cz[0]), voxel_to_block1(cz[1])};
  // The sample is invalid if any corner's block is missing; in empty space
  // the lowest corner's block already says so (one cached probe).
  const int32_t p000 = cached_find(V, cache, i3{bx[0], by[0], bz[0]});
  if (p000 < 0) return false;
  int32_t pb[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const bool same = bx[mx] == bx[0] && by[my] == by[0] && bz[mz] == bz[0];
    pb[i] = same ? p000 : cached_find(V, cache, i3{bx[mx], by[my], bz[mz]});
    if (pb[i] < 0) return false;
  }
  float sdf = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const uint32_t li = (uint32_t)((cx[mx] - bx[mx] * 8) + 8 * (cy[my] - by[my] * 8) +
                                   64 * (cz[mz] - bz[mz] * 8));
    const float w = (mx ? rx : 1.0f - rx) * (my ? ry : 1.0f - ry) * (mz ? rz : 1.0f - rz);
    sdf += w * V.sdf[(size_t)pb[i] * kBlockVoxels + li];
  }
  *out = sdf;
  return true;
}

// March one pixel; returns the number of valid samples taken.
__device__ __forceinline__ uint32_t cast_pixel(const View& V, const RayConst& R, uint32_t x, uint32_t y,
                               float* res_d, f3* res_v, f3* res_n) {
  uint32_t nsamp = 0;
  float out_d = -INFINITY;
  f3 out_v = mk3(-INFINITY, -INFINITY, -INFINITY);
  f3 out_n = mk3(-INFINITY, -INFINITY, -INFINITY);

  const f3 dir_c = normalize(reproject(R.cam, x, y, 1.0f));
  const float t_min = R.min_depth / dir_c.z;
  const float t_max = R.max_depth / dir_c.z;
  const f3 o = xform(R.wTc, mk3(0.0f, 0.0f, 0.0f));
  const f3 dir = normalize(xform_dir(R.wTc, dir_c));

  BlockCache cache{kEmptyKey, kEmptyKey, kFreePtr, kFreePtr};
  float prev_sdf = 0.0f, prev_t = 0.0f;
  bool prev_ok = false;
  for (float t = t_min; t < t_max; t += R.step) {
    const f3 p = mk3(o.x + t * dir.x, o.y + t * dir.y, o.z + t * dir.z);
    float s;
    if (!ray_sdf(V, cache, p, R.inv_vs, &s)) continue;
    ++nsamp;
    if (prev_ok && prev_sdf > 0.0f && s < 0.0f) {
      // BisectionIntersection: three regula-falsi steps.
      float l = prev_t, r = t, m = (l + r) * 0.5f;
      float l_sdf = prev_sdf, r_sdf = s;
      bool found = true;
      for (int k = 0; k < 3; ++k) {
        m = l + (l_sdf / (l_sdf - r_sdf)) * (r - l);
        float m_sdf;
        if (!ray_sdf(V, cache, mk3(o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z), R.inv_vs,
                     &m_sdf)) {
          found = false;
          break;
        }
        if (l_sdf * m_sdf > 0.0f) {
          l = m;
          l_sdf = m_sdf;
        } else {
          r = m;
          r_sdf = m_sdf;
        }
      }
      if (found && fabsf(prev_sdf - s) < R.sample_sdf_threshold && fabsf(s) < R.sdf_threshold) {
        const float d = m * dir_c.z;
        out_d = d;
        out_v = reproject(R.cam, x, y, d);
        if (R.enable_gradients) {
          const f3 iso = mk3(o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z);
          const float h = 0.5f * V.geo.voxel_size;
          bool valid = true;
          float sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
#pragma unroll
          for (int k = 0; k < 3 && valid; ++k) {
            const f3 dp = mk3(k == 0 ? h : 0.0f, k == 1 ? h : 0.0f, k == 2 ? h : 0.0f);
            valid = ray_sdf(V, cache, iso - dp, R.inv_vs, &sn[k]);
            if (!valid) break;
            valid = ray_sdf(V, cache, iso + dp, R.inv_vs, &sp[k]);
          }
          const float vs = V.geo.voxel_size;
          const f3 grad = mk3((sp[0] - sn[0]) / vs, (sp[1] - sn[1]) / vs, (sp[2] - sn[2]) / vs);
          const float l2 = length(grad);
          f3 nw = (l2 > 0.0f && valid) ? grad / l2 : mk3(0.0f, 0.0f, 0.0f);
          nw = -nw;
          out_n = xform_dir(R.cTw, nw);
        }
        break;
      }
    }
    prev_sdf = s;
    prev_t = t;
    prev_ok = true;
  }
  *res_d = out_d;
  *res_v = out_v;
  *res_n = out_n;
  return nsamp;
}

__device__ __forceinline__ const RayConst& ray_of(const One&, uint32_t, const RayConst& r) { return r; }
__device__ __forceinline__ const RayConst& ray_of(const Many& m, uint32_t s, const RayConst&) {
  return m.f[s].ray;
}

// Scene = blockIdx.z; a batch's outputs are scene-major (W*H per scene).
template <class Src>
__global__ __launch_bounds__(64) void raycast_kernel(Src src_, const SceneFrame* __restrict__ tab, RayConst R1, float* __restrict__ depth_out,
                                                     float* __restrict__ vertex_out,
                                                     float* __restrict__ normal_out) {
  const Src src = with_table(src_, tab);
  const uint32_t scene = blockIdx.z;
  const View& V = src.view(scene);
  const RayConst& R = ray_of(src, scene, R1);
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  uint32_t ns = 0;
  if (x < R.cam.width && y < R.cam.height) {
    const size_t pix = (size_t)scene * R.cam.width * R.cam.height + y * R.cam.width + x;
    float d;
    f3 vv, nn;
    ns = cast_pixel(V, R, x, y, &d, &vv, &nn);
    if (depth_out) depth_out[pix] = d;
    if (vertex_out) {
      vertex_out[3 * pix + 0] = vv.x;
      vertex_out[3 * pix + 1] = vv.y;
      vertex_out[3 * pix + 2] = vv.z;
    }
    if (normal_out) {
      normal_out[3 * pix + 0] = nn.x;
      normal_out[3 * pix + 1] = nn.y;
      normal_out[3 * pix + 2] = nn.z;
    }
  }
  wave_add_u64(&V.ctr->ray_samples, ns);
}

void launch_raycast(const View& V, const RayConst& p, float* depth, float* vertex, float* normal,
                    hipStream_t s) {
  dim3 grid((p.cam.width + 7) / 8, (p.cam.height + 7) / 8);
  raycast_kernel<<<grid, 64, 0, s>>>(One{V}, nullptr, p, depth, vertex, normal);
}

void launch_batch_raycast(const BatchDev& B, float* depth, float* vertex, float* normal,
                          hipStream_t s) {
  dim3 grid((B.width + 7) / 8, (B.height + 7) / 8, B.nscenes);
  raycast_kernel<<<grid, 64, 0, s>>>(B.src(), B.frames, RayConst{}, depth, vertex, normal);
}

}  // namespace mh
