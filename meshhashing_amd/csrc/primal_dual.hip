// primal_dual.hip -- TV/Huber primal-dual denoising of the fused SDF.
//
// Reference: optimize/primal_dual.cu:10-223 (PrimalDualInitKernel,
// PrimalDualIteratePass1Kernel, PrimalDualIteratePass2Kernel, Huber),
// geometry/primal_dual_query.h:15-234 (forward-difference gradients,
// backward-difference divergence), core/voxel.h:62-82 (PrimalDualVariables).
//
// MI355X layout: the six face neighbours of every worklist block are looked
// up once per solve into a table (the block set does not change while
// iterating), so a stencil access outside the block is one table read plus a
// coalesced voxel read instead of a hash probe.  Both passes are Jacobi
// steps; the energies are reduced per workgroup into FP64 partials and summed
// in a fixed order, so the whole iteration is deterministic.
#include "mh_kernels.h"

namespace mh {

// EPSILON is the double literal 1e-6 (core/common.h:60): float weights are
// widened before the comparison.
__device__ __forceinline__ bool below_eps(float w) { return (double)w < 1e-6; }
__device__ __forceinline__ bool above_eps(float w) { return (double)w > 1e-6; }

// Face-neighbour table: nbr[6*c + k], k = +x, +y, +z, -x, -y, -z.
__global__ __launch_bounds__(256) void pd_neighbors_kernel(View V, int32_t* nbr) {
  const uint32_t n = V.ctr->cand_count;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t j = gid; j < 6 * n; j += gridDim.x * blockDim.x) {
    const uint32_t c = j / 6, k = j % 6;
    const i3 b = unpack_key(V.block_keys[V.cand[c]]);
    const int s = k < 3 ? 1 : -1;
    const int ax = (int)(k % 3);
    const i3 q = i3{b.x + (ax == 0 ? s : 0), b.y + (ax == 1 ? s : 0), b.z + (ax == 2 ? s : 0)};
    nbr[j] = hash_find(V, q);
  }
}

// Address of the voxel at local offset o + d (d one of +-e_k), or -1 when
// its block is not allocated.
__device__ __forceinline__ int64_t nb_index(const int32_t* nbr6, int32_t self, i3 o, int ax,
                                            int s) {
  int c[3] = {o.x, o.y, o.z};
  c[ax] += s;
  int32_t p = self;
  if (c[ax] < 0 || c[ax] > 7) {
    p = nbr6[s > 0 ? ax : 3 + ax];
    c[ax] -= 8 * s;
    if (p < 0) return -1;
  }
  return (int64_t)p * kBlockVoxels + c[0] + 8 * c[1] + 64 * c[2];
}

// PrimalDualInitKernel (primal_dual.cu:15-43) with GetInitSDFGradient
// (primal_dual_query.h:44-86).
__global__ __launch_bounds__(256) void pd_init_kernel(View V, const int32_t* nbr) {
  const uint32_t n = V.ctr->cand_count;
  const float vs = V.geo.voxel_size;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const int32_t* nb = nbr + 6 * ci;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      if (below_eps(V.weight[idx])) continue;
      const i3 o = devectorize(i);
      const float c = V.sdf[idx];
      float g[3] = {0.0f, 0.0f, 0.0f};
      bool ok = true;
      for (int k = 0; k < 3 && ok; ++k) {
        const int64_t j = nb_index(nb, ptr, o, k, +1);
        if (j < 0 || below_eps(V.weight[j])) {
          ok = false;
          break;
        }
        g[k] = V.sdf[j];
      }
      const f3 grad = ok ? mk3(g[0] - c, g[1] - c, g[2] - c) : mk3(0.0f, 0.0f, 0.0f);
      V.pd_p[3 * idx + 0] = 0.0f;
      V.pd_p[3 * idx + 1] = 0.0f;
      V.pd_p[3 * idx + 2] = 0.0f;
      V.pd_sdf_bar[idx] = 0.0f;
      V.pd_w[idx] = expf(length(grad) / vs);
      V.pd_sdf0[idx] = c;
      V.pd_mask[idx] = 1;
    }
  }
}

__device__ __forceinline__ float huber(float x, float alpha) {
  return (fabsf(x) < alpha) ? 0.5f * (x * x) / alpha : (fabsf(x) - 0.5f * alpha);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

// PrimalDualIteratePass1Kernel (primal_dual.cu:45-109): energies, dual
// ascent p += sigma grad(sdf_bar), Huber shrink, projection onto |p| <= 1.
__global__ __launch_bounds__(256) void pd_pass1_kernel(View V, const int32_t* nbr, float sigma,
                                                       double* partials) {
  __shared__ double red[8];
  const float alpha = 0.02f;
  const uint32_t n = V.ctr->cand_count;
  double e_data = 0.0, e_tv = 0.0;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const int32_t* nb = nbr + 6 * ci;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      const float w = V.weight[idx];
      if (below_eps(w)) continue;
      const i3 o = devectorize(i);
      const float sdf = V.sdf[idx];
      float dd = fabsf(sdf - V.pd_sdf0[idx]);
      dd *= dd;
      if (above_eps(w)) e_data += (double)dd;
      // GetSDFGradient (primal_dual_query.h:88-133) and GetPrimalGradient
      // (:189-234): forward differences; any missing, unobserved or
      // unmasked neighbour zeroes the whole gradient.
      float gs[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
      bool ok = true;
      for (int k = 0; k < 3 && ok; ++k) {
        const int64_t j = nb_index(nb, ptr, o, k, +1);
        if (j < 0 || below_eps(V.weight[j]) || !V.pd_mask[j]) {
          ok = false;
          break;
        }
        gs[k] = V.sdf[j];
        gb[k] = V.pd_sdf_bar[j];
      }
      const float sb = V.pd_sdf_bar[idx];
      const f3 grad_s = ok ? mk3(gs[0] - sdf, gs[1] - sdf, gs[2] - sdf) : mk3(0.0f, 0.0f, 0.0f);
      const f3 grad_b = ok ? mk3(gb[0] - sb, gb[1] - sb, gb[2] - sb) : mk3(0.0f, 0.0f, 0.0f);
      e_tv += (double)huber(length(grad_s), alpha);
      f3 p = mk3(V.pd_p[3 * idx], V.pd_p[3 * idx + 1], V.pd_p[3 * idx + 2]);
      p = p + sigma * grad_b;
      p = p / (1.0f + sigma * alpha);
      p = p / fmaxf(1.0f, length(p));
      V.pd_p[3 * idx + 0] = p.x;
      V.pd_p[3 * idx + 1] = p.y;
      V.pd_p[3 * idx + 2] = p.z;
    }
  }
  const double sd = block_sum(e_data, red);
  const double st = block_sum(e_tv, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x + 0] = sd;
    partials[2 * blockIdx.x + 1] = st;
  }
}

// PrimalDualIteratePass2Kernel (primal_dual.cu:111-151) with
// GetDualDivergence (primal_dual_query.h:135-187).
__global__ __launch_bounds__(256) void pd_pass2_kernel(View V, const int32_t* nbr, float lambda,
                                                       float tau) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const int32_t* nb = nbr + 6 * ci;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      if (below_eps(V.weight[idx])) continue;
      const i3 o = devectorize(i);
      const float prev = V.sdf[idx];
      float div = 0.0f;
      if (V.pd_mask[idx]) {
        const f3 pc = mk3(V.pd_p[3 * idx], V.pd_p[3 * idx + 1], V.pd_p[3 * idx + 2]);
        float pn[3] = {0, 0, 0};
        bool ok = true;
        for (int k = 0; k < 3 && ok; ++k) {
          const int64_t j = nb_index(nb, ptr, o, k, -1);
          if (j < 0 || below_eps(V.weight[j]) || !V.pd_mask[j]) {
            ok = false;
            break;
          }
          pn[k] = V.pd_p[3 * j + k];
        }
        if (ok) div = (pn[0] - pc.x) + (pn[1] - pc.y) + (pn[2] - pc.z);
      }
      const float w_pd = V.pd_w[idx];
      const float sdf0 = V.pd_sdf0[idx];
      const float lam = lambda * w_pd;
      float s = prev - tau * div;
      s = (s + lam * tau * sdf0) / (1.0f + lam * tau);
      if (w_pd > 2.0f) s = sdf0;
      V.sdf[idx] = s;
      V.pd_sdf_bar[idx] = 2.0f * s - prev;
    }
  }
}

// Fixed-order sum of the per-workgroup partials.
__global__ void pd_energy_kernel(const double* partials, uint32_t n, float lambda, double* energy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double d = 0.0, t = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    d += partials[2 * i];
    t += partials[2 * i + 1];
  }
  energy[0] = d;
  energy[1] = t;
  energy[2] = d * (double)lambda / 2.0 + t;
}

void launch_pd_neighbors(const View& V, int32_t* nbr, hipStream_t s, int grid) {
  pd_neighbors_kernel<<<grid, 256, 0, s>>>(V, nbr);
}
void launch_pd_init(const View& V, const int32_t* nbr, hipStream_t s, int grid) {
  pd_init_kernel<<<grid, 256, 0, s>>>(V, nbr);
}
void launch_pd_iterate(const View& V, const int32_t* nbr, float lambda, float sigma, float tau,
                       double* partials, uint32_t npartials, double* energy, hipStream_t s,
                       int grid) {
  const int g = grid < (int)npartials ? grid : (int)npartials;
  pd_pass1_kernel<<<g, 256, 0, s>>>(V, nbr, sigma, partials);
  pd_energy_kernel<<<1, 64, 0, s>>>(partials, (uint32_t)g, lambda, energy);
  pd_pass2_kernel<<<grid, 256, 0, s>>>(V, nbr, lambda, tau);
}

}  // namespace mh
