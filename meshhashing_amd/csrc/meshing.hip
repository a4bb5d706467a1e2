// meshing.hip -- incremental marching cubes over the worklist blocks.
//
// Reference: meshing/marching_cubes.cu:16-382 (SurfelExtractionKernel,
// TriangleExtractionKernel, RecycleTrianglesKernel, RecycleVerticesKernel),
// geometry/spatial_query.h:23-97 (trilinear value and gradient),
// core/mesh.h:53-107 (vertex/triangle pools and ref counts),
// visualization/color_util.h:13-37 (ValToRGB).
//
// MI355X layout: one 512-lane workgroup per block, one cube per lane.  The
// block and its 26 neighbours are staged once as an 11^3 voxel shell in LDS
// (local coordinates -1..9), which covers every corner, every trilinear
// sample and every gradient stencil of the block's cubes, so the 8 + 56
// hashed lookups per vertex of the reference become 26 lookups per block.
//
// Vertices are canonical: vertex (owner voxel o, axis a) is always computed
// from the endpoints o and o + e_a in that order, with the owner cube's
// corner positions.  In the reference up to four cubes write the same vertex
// with different endpoint orders and corner positions (last writer wins);
// here every writer stores the same bits.
#include "mh_kernels.h"
#include "mh_mc_tables.h"

namespace mh {

__constant__ McTables kMc = build_mc_tables();

constexpr int kShell = 11;
constexpr int kShellN = kShell * kShell * kShell;

template <bool PSDF>
struct Shell {
  float sdf[kShellN];
  float w[kShellN];
  float a[PSDF ? kShellN : 1];
  float b[PSDF ? kShellN : 1];
  uint8_t ok[kShellN];
};

__device__ __forceinline__ int sidx(int x, int y, int z) {
  return (x + 1) + kShell * (y + 1) + kShell * kShell * (z + 1);
}
__device__ __forceinline__ int nb27(int dx, int dy, int dz) {
  return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}

template <bool PSDF>
__device__ void stage_shell(const View& V, i3 b, int32_t self, Shell<PSDF>& S, int32_t* nb) {
  if (threadIdx.x < 27) {
    const int t = (int)threadIdx.x;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    nb[t] = (dx | dy | dz) == 0 ? self : hash_find(V, i3{b.x + dx, b.y + dy, b.z + dz});
  }
  __syncthreads();
  for (int s = (int)threadIdx.x; s < kShellN; s += (int)blockDim.x) {
    const int lx = s % kShell - 1, ly = (s / kShell) % kShell - 1, lz = s / (kShell * kShell) - 1;
    const int dx = lx < 0 ? -1 : (lx >= 8 ? 1 : 0);
    const int dy = ly < 0 ? -1 : (ly >= 8 ? 1 : 0);
    const int dz = lz < 0 ? -1 : (lz >= 8 ? 1 : 0);
    const int32_t p = nb[nb27(dx, dy, dz)];
    if (p < 0) {
      S.ok[s] = 0;
      continue;
    }
    const size_t idx = (size_t)p * kBlockVoxels + (size_t)((lx - 8 * dx) + 8 * (ly - 8 * dy) +
                                                           64 * (lz - 8 * dz));
    S.sdf[s] = V.sdf[idx];
    S.w[s] = V.weight[idx];
    if (PSDF) {
      S.a[s] = V.va[idx];
      S.b[s] = V.vb[idx];
    }
    S.ok[s] = 1;
  }
  __syncthreads();
}

// WorldToVoxeli along one axis.
__device__ __forceinline__ int vox1(float w, float vs) {
  const float p = w / vs;
  return f2i(p + (float)sgn(p) * 0.5f);
}

struct SVal {
  float sdf, a, b, radius;
};

// GetSpatialValue (spatial_query.h:23-66) against the staged shell.  base is
// the block's voxel origin.  Corner voxels come from rounding the corner's
// world position, exactly as the reference, so ties and negative
// coordinates pick the same voxels.
template <bool PSDF>
__device__ bool shell_value(const Shell<PSDF>& S, const Geo& g, i3 base, f3 pos, SVal* out,
                            bool* oob) {
  const float vs = g.voxel_size;
  const float hv = 0.5f * vs;
  const float pcx = pos.x - hv, pcy = pos.y - hv, pcz = pos.z - hv;
  const float qx = pos.x / vs, qy = pos.y / vs, qz = pos.z / vs;
  const float rx = qx - floorf(qx), ry = qy - floorf(qy), rz = qz - floorf(qz);
  const int cx[2] = {vox1(pcx + 0.0f, vs) - base.x, vox1(pcx + vs, vs) - base.x};
  const int cy[2] = {vox1(pcy + 0.0f, vs) - base.y, vox1(pcy + vs, vs) - base.y};
  const int cz[2] = {vox1(pcz + 0.0f, vs) - base.z, vox1(pcz + vs, vs) - base.z};
  float sdf = 0.0f, a = 0.0f, b = 0.0f, rad = 0.0f;
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const int lx = cx[mx], ly = cy[my], lz = cz[mz];
    if (lx < -1 || lx > 9 || ly < -1 || ly > 9 || lz < -1 || lz > 9) {
      *oob = true;
      return false;
    }
    const int s = sidx(lx, ly, lz);
    if (!S.ok[s]) return false;
    const float w = (mx ? rx : 1.0f - rx) * (my ? ry : 1.0f - ry) * (mz ? rz : 1.0f - rz);
    sdf += w * S.sdf[s];
    if (PSDF) {
      a += w * S.a[s];
      b += w * S.b[s];
    }
    rad += w * sqrtf(1.0f / S.w[s]);
  }
  out->sdf = sdf;
  out->a = a;
  out->b = b;
  out->radius = rad;
  return true;
}

// ValToRGB (color_util.h:13-37); the 0.25/0.5/0.75 literals are doubles
// there, so the band tests and two of the channel formulas run in double.
__device__ __forceinline__ f3 val_to_rgb(float v, float vmin, float vmax) {
  f3 c = mk3(1.0f, 1.0f, 1.0f);
  if (v < vmin) v = vmin;
  if (v > vmax) v = vmax;
  const float dv = vmax - vmin;
  const double vd = v, mn = vmin, dd = dv;
  if (vd < mn + 0.25 * dd) {
    c.x = 0.0f;
    c.y = 4.0f * (v - vmin) / dv;
  } else if (vd < mn + 0.5 * dd) {
    c.x = 0.0f;
    c.z = (float)(1.0 + 4.0 * (mn + 0.25 * dd - vd) / dd);
  } else if (vd < mn + 0.75 * dd) {
    c.x = (float)(4.0 * (vd - mn - 0.5 * dd) / dd);
    c.z = 0.0f;
  } else {
    c.y = (float)(1.0 + 4.0 * (mn + 0.75 * dd - vd) / dd);
    c.z = 0.0f;
  }
  return c;
}

// AllocateVertexWithMutex's payload (marching_cubes.cu:45-72).
template <bool PSDF>
__device__ void vertex_payload(const Shell<PSDF>& S, const Geo& g, i3 base, f3 p, Vertex* vt,
                               bool* oob) {
  SVal sv{};
  const bool ok = shell_value(S, g, base, p, &sv, oob);
  vt->px = p.x;
  vt->py = p.y;
  vt->pz = p.z;
  if (ok) {
    const float inv_sigma2 = 1.0f / (sv.radius * sv.radius);
    vt->radius = sqrtf(1.0f / inv_sigma2);
    const f3 c = val_to_rgb(sv.a / (sv.a + sv.b), 0.4f, 1.0f);
    vt->cr = c.x;
    vt->cg = c.y;
    vt->cb = c.z;
  } else {  // the reference reads an uninitialised query here
    vt->radius = 0.0f;
    vt->cr = vt->cg = vt->cb = 0.0f;
  }
  // GetSpatialSDFGradient (spatial_query.h:68-97): central differences at
  // +-vs/2, with the reference's short-circuit on the first invalid sample.
  const float h = 0.5f * g.voxel_size;
  bool valid = true;
  float sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
  for (int k = 0; k < 3 && valid; ++k) {
    const f3 dp = mk3(k == 0 ? h : 0.0f, k == 1 ? h : 0.0f, k == 2 ? h : 0.0f);
    SVal q{};
    valid = shell_value(S, g, base, p - dp, &q, oob);
    sn[k] = q.sdf;
    if (!valid) break;
    valid = shell_value(S, g, base, p + dp, &q, oob);
    sp[k] = q.sdf;
  }
  const float vs = g.voxel_size;
  const f3 grad = mk3((sp[0] - sn[0]) / vs, (sp[1] - sn[1]) / vs, (sp[2] - sn[2]) / vs);
  const float l = length(grad);
  const f3 n = (l > 0.0f && valid) ? grad / l : mk3(0.0f, 0.0f, 0.0f);
  vt->nx = n.x;
  vt->ny = n.y;
  vt->nz = n.z;
}

__device__ __forceinline__ int32_t pop_heap(int32_t* top, const int32_t* heap, const View& V,
                                            uint32_t bit) {
  const int32_t a = atomicSub(top, 1);
  if (a < 0) {
    atomicAdd(top, 1);
    flag(V, bit);
    return kFreePtr;
  }
  return heap[a];
}
__device__ __forceinline__ void push_heap(int32_t* top, int32_t* heap, int32_t p) {
  const int32_t a = atomicAdd(top, 1);
  heap[a + 1] = p;
}

// ---------------------------------------------------------------- pass 1
// SurfelExtractionKernel (marching_cubes.cu:75-170): cube index of every
// cube of the block, and the vertex of every active edge.
template <bool PSDF>
__global__ __launch_bounds__(512) void mc_surfel_kernel(View V, McConst mc) {
  __shared__ Shell<PSDF> S;
  __shared__ int32_t nb[27];
  const uint32_t n = V.ctr->cand_count;
  const Geo g = V.geo;
  uint32_t nwrites = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&V.ctr->mc_blocks, (unsigned long long)n);
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const i3 b = unpack_key(V.block_keys[ptr]);
    stage_shell<PSDF>(V, b, ptr, S, nb);
    const i3 base = i3{b.x * kBlockSide, b.y * kBlockSide, b.z * kBlockSide};
    const uint32_t i = threadIdx.x;
    const i3 o = devectorize(i);
    const size_t idx = (size_t)ptr * kBlockVoxels + i;
    const uint32_t prev = V.cube[idx] & 255u;
    int cube = 0;
    bool valid = true;
    for (int k = 0; k < 8; ++k) {
      const int s = sidx(o.x + corner_off(k, 0), o.y + corner_off(k, 1), o.z + corner_off(k, 2));
      if (!S.ok[s]) { valid = false; break; }
      const float d = S.sdf[s];
      if (fabsf(d) > mc.sdf_threshold) { valid = false; break; }
      const float a = PSDF ? S.a[s] : 0.0f, bb = PSDF ? S.b[s] : 0.0f;
      const float rho = a / (a + bb);
      if (rho < mc.rho_min || S.w[s] < mc.weight_min) { valid = false; break; }
      if (d < 0.0f) cube |= 1 << k;
    }
    if (!valid) cube = 0;
    V.cube[idx] = (uint16_t)((uint32_t)cube | (prev << 8));
    bool oob = false;
    if (cube != 0 && cube != 255) {
      const unsigned em = kMc.edges[cube];
      for (int e = 0; e < 12; ++e) {
        if (!(em & (1u << e))) continue;
        const int ax = edge_axis(e);
        const int ox = o.x + edge_owner_off(e, 0), oy = o.y + edge_owner_off(e, 1),
                  oz = o.z + edge_owner_off(e, 2);
        const float vA = S.sdf[sidx(ox, oy, oz)];
        const float vB = S.sdf[sidx(ox + (ax == 0), oy + (ax == 1), oz + (ax == 2))];
        const f3 pA = voxel_to_world(g, i3{base.x + ox, base.y + oy, base.z + oz});
        const f3 pB = mk3(pA.x + (ax == 0 ? g.voxel_size : 0.0f), pA.y + (ax == 1 ? g.voxel_size : 0.0f),
                          pA.z + (ax == 2 ? g.voxel_size : 0.0f));
        // VertexIntersection (marching_cubes.cu:16-28), canonical order.
        f3 p;
        if ((double)fabsf(vA) < mc.snap) {
          p = pA;
        } else if ((double)fabsf(vB) < mc.snap) {
          p = pB;
        } else {
          const float mu = (0.0f - vA) / (vB - vA);
          p = mk3(pA.x + mu * (pB.x - pA.x), pA.y + mu * (pB.y - pA.y), pA.z + mu * (pB.z - pA.z));
        }
        const int dx = ox >> 3, dy = oy >> 3, dz = oz >> 3;
        const int32_t optr = nb[nb27(dx, dy, dz)];
        const uint32_t oi = (uint32_t)((ox - 8 * dx) + 8 * (oy - 8 * dy) + 64 * (oz - 8 * dz));
        int32_t* slot = &V.vptr[((size_t)optr * kBlockVoxels + oi) * 3 + ax];
        int32_t cur = *slot;
        int32_t vp = cur;
        bool write = cur >= 0;
        if (cur == kFreePtr) {
          const int32_t old = atomicCAS(slot, kFreePtr, kLockPtr);
          if (old == kFreePtr) {
            vp = pop_heap(&V.ctr->vertex_top, V.vertex_heap, V, MH_STATUS_VERTEX_HEAP_EXHAUSTED);
            atomicExch(slot, vp);
            write = vp >= 0;
          } else {
            vp = old;
            write = old >= 0;
          }
        }
        if (write) {
          Vertex vt;
          vertex_payload<PSDF>(S, g, base, p, &vt, &oob);
          Vertex* dst = &V.vertices[vp];
          *reinterpret_cast<float4*>(&dst->px) = make_float4(vt.px, vt.py, vt.pz, vt.radius);
          dst->nx = vt.nx;
          dst->ny = vt.ny;
          dst->nz = vt.nz;
          *reinterpret_cast<float4*>(&dst->cr) = make_float4(vt.cr, vt.cg, vt.cb, 0.0f);
          ++nwrites;
        }
      }
    }
    if (oob) flag(V, MH_STATUS_STENCIL_RANGE);
    __syncthreads();
  }
  wave_add_u64(&V.ctr->mc_vertex_writes, nwrites);
}

// ---------------------------------------------------------------- pass 2
// TriangleExtractionKernel (marching_cubes.cu:180-275).
__global__ __launch_bounds__(512) void mc_triangle_kernel(View V) {
  __shared__ int32_t nb[8];
  __shared__ int counts[2];
  const uint32_t n = V.ctr->cand_count;
  uint32_t nassign = 0;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const i3 b = unpack_key(V.block_keys[ptr]);
    const int t8 = (int)threadIdx.x;
    if (t8 < 8)
      nb[t8] = t8 == 0 ? ptr : hash_find(V, i3{b.x + (t8 & 1), b.y + ((t8 >> 1) & 1), b.z + ((t8 >> 2) & 1)});
    if (t8 == 0) { counts[0] = 0; counts[1] = 0; }
    __syncthreads();
    const uint32_t i = threadIdx.x;
    const i3 o = devectorize(i);
    const size_t idx = (size_t)ptr * kBlockVoxels + i;
    const int nv = (V.vptr[3 * idx] >= 0) + (V.vptr[3 * idx + 1] >= 0) + (V.vptr[3 * idx + 2] >= 0);
    if (nv) {
      const bool inner = o.x >= 1 && o.y >= 1 && o.z >= 1 && o.x < 7 && o.y < 7 && o.z < 7;
      atomicAdd(&counts[inner ? 0 : 1], nv);
    }
    const int cube = V.cube[idx] & 255;
    if (cube != 0 && cube != 255) {
      int32_t vps[12];
      const unsigned em = kMc.edges[cube];
      for (int e = 0; e < 12; ++e) {
        vps[e] = kFreePtr;
        if (!(em & (1u << e))) continue;
        const int ax = edge_axis(e);
        const int ox = o.x + edge_owner_off(e, 0), oy = o.y + edge_owner_off(e, 1),
                  oz = o.z + edge_owner_off(e, 2);
        const int dx = ox >> 3, dy = oy >> 3, dz = oz >> 3;
        const int32_t optr = nb[dx + 2 * dy + 4 * dz];
        const uint32_t oi = (uint32_t)((ox - 8 * dx) + 8 * (oy - 8 * dy) + 64 * (oz - 8 * dz));
        vps[e] = optr < 0 ? kFreePtr : V.vptr[((size_t)optr * kBlockVoxels + oi) * 3 + ax];
      }
      const int nt = kMc.ntri[cube];
      for (int t = 0; t < nt; ++t) {
        const int32_t v0 = vps[kMc.tri[cube][3 * t]], v1 = vps[kMc.tri[cube][3 * t + 1]],
                      v2 = vps[kMc.tri[cube][3 * t + 2]];
        if (v0 < 0 || v1 < 0 || v2 < 0) continue;  // vertex pool exhausted in pass 1
        int32_t tp = V.tptr[5 * idx + t];
        if (tp == kFreePtr) {
          tp = pop_heap(&V.ctr->triangle_top, V.triangle_heap, V, MH_STATUS_TRIANGLE_HEAP_EXHAUSTED);
          if (tp < 0) continue;
        } else {
          atomicSub(&V.vertices[V.triangles[3 * tp + 0]].ref_count, 1);
          atomicSub(&V.vertices[V.triangles[3 * tp + 1]].ref_count, 1);
          atomicSub(&V.vertices[V.triangles[3 * tp + 2]].ref_count, 1);
        }
        V.tptr[5 * idx + t] = tp;
        V.triangles[3 * tp + 0] = v0;
        V.triangles[3 * tp + 1] = v1;
        V.triangles[3 * tp + 2] = v2;
        atomicAdd(&V.vertices[v0].ref_count, 1);
        atomicAdd(&V.vertices[v1].ref_count, 1);
        atomicAdd(&V.vertices[v2].ref_count, 1);
        ++nassign;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      V.surfel[2 * ptr + 0] = counts[0];
      V.surfel[2 * ptr + 1] = counts[1];
    }
    __syncthreads();
  }
  wave_add_u64(&V.ctr->mc_triangle_writes, nassign);
}

// ---------------------------------------------------------------- recycling
// RecycleTrianglesKernel (marching_cubes.cu:278-301).
__global__ __launch_bounds__(256) void mc_recycle_triangles_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      const int cube = V.cube[idx] & 255;
      for (int t = kMc.ntri[cube]; t < 5; ++t) {
        const int32_t tp = V.tptr[5 * idx + t];
        if (tp == kFreePtr) continue;
        atomicSub(&V.vertices[V.triangles[3 * tp + 0]].ref_count, 1);
        atomicSub(&V.vertices[V.triangles[3 * tp + 1]].ref_count, 1);
        atomicSub(&V.vertices[V.triangles[3 * tp + 2]].ref_count, 1);
        V.triangles[3 * tp + 0] = -1;  // Triangle::Clear (triangle.h:14-16)
        V.triangles[3 * tp + 1] = -1;
        V.triangles[3 * tp + 2] = -1;
        push_heap(&V.ctr->triangle_top, V.triangle_heap, tp);
        V.tptr[5 * idx + t] = kFreePtr;
      }
    }
  }
}

// RecycleVerticesKernel (marching_cubes.cu:303-321).
__global__ __launch_bounds__(256) void mc_recycle_vertices_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      for (int ax = 0; ax < 3; ++ax) {
        const int32_t vp = V.vptr[3 * idx + ax];
        if (vp < 0) continue;
        if (V.vertices[vp].ref_count != 0) continue;
        Vertex z{};
        V.vertices[vp] = z;
        push_heap(&V.ctr->vertex_top, V.vertex_heap, vp);
        V.vptr[3 * idx + ax] = kFreePtr;
      }
    }
  }
}

void launch_marching_cubes(const View& V, const McConst& p, hipStream_t s, int grid) {
  if (V.va)
    mc_surfel_kernel<true><<<grid, 512, 0, s>>>(V, p);
  else
    mc_surfel_kernel<false><<<grid, 512, 0, s>>>(V, p);
  mc_triangle_kernel<<<grid, 512, 0, s>>>(V);
  mc_recycle_triangles_kernel<<<grid, 256, 0, s>>>(V);
  mc_recycle_vertices_kernel<<<grid, 256, 0, s>>>(V);
}

}  // namespace mh
