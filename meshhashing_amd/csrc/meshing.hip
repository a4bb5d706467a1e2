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
  uint8_t f[kShellN];  // kShOk | kShCube | kShNeg per voxel
};
// Shell voxel flags: the voxel exists (its block is allocated); it passes
// the cube tests of a corner (|sdf| <= threshold, rho >= rho_min,
// weight >= weight_min); its SDF is negative.
constexpr uint8_t kShOk = 1, kShCube = 2, kShNeg = 4;

__device__ __forceinline__ int sidx(int x, int y, int z) {
  return (x + 1) + kShell * (y + 1) + kShell * kShell * (z + 1);
}
__device__ __forceinline__ int nb27(int dx, int dy, int dz) {
  return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}

constexpr int kCfg = 9;  // local cube coordinates -1..7

// The edge with owner offset (dx, dy, dz) and axis ax (a constant for
// constant arguments).
__host__ __device__ constexpr int edge_with(int dx, int dy, int dz, int ax) {
  int found = -1;
  for (int e = 0; e < 12; ++e)
    if (edge_owner_off(e, 0) == dx && edge_owner_off(e, 1) == dy && edge_owner_off(e, 2) == dz &&
        edge_axis(e) == ax)
      found = e;
  return found;
}
static_assert(edge_with(0, 0, 0, 0) >= 0 && edge_with(0, 1, 1, 0) >= 0 && edge_with(1, 0, 1, 1) >= 0 &&
                  edge_with(1, 1, 0, 2) >= 0,
              "every owner offset across an edge's axis names an edge");

// The 27 neighbours of every worklist block (x fastest, the block itself at
// 13) with their meshed-in-this-pass marks, one lane per (item, neighbour),
// before the surfel pass: the dependent worklist -> key -> occupancy ->
// bucket -> mark reads of a lookup then run with the whole grid's
// parallelism instead of on one workgroup's critical path.  Entry: the pool
// index, with bit 30 set when the neighbour is meshed in this pass; a
// negative value for an absent block.
__device__ __forceinline__ const McConst& mc_of(const One&, uint32_t, const McConst& mc) { return mc; }
__device__ __forceinline__ const McConst& mc_of(const Many& m, uint32_t s, const McConst&) {
  return m.f[s].mc;
}
constexpr uint32_t kLookupLane = kBlockVoxels - 64;  // the surfel pass's last wave
// The surfel pass hands out its first items statically (workgroup g takes
// g, g + grid, ...) and the rest on demand from a counter, so workgroups
// whose blocks held little surface take more of them.  static_q quarters of
// the worklist (rounded down to whole rounds of the grid, at least one) are
// static; the counter (cleared by the neighbour-table pass) counts the items
// handed out after them.
__device__ __forceinline__ uint32_t mc_static_items(uint32_t n, uint32_t grid, uint32_t static_q) {
  const uint32_t k = (static_q * n / 4u) / grid;
  return (k ? k : 1u) * grid;
}
constexpr int32_t kMeshedBit = 1 << 30;
template <class Src>
__global__ __launch_bounds__(256) void mc_neighbours_kernel(Src src_, const SceneFrame* __restrict__ tab,
                                                             McConst mc1, int32_t* __restrict__ nbt) {
  const Src src = with_table(src_, tab);
  const uint32_t total = src.count() * 27u;
  // the surfel pass's counter of items handed out on demand (it runs next)
  if (blockIdx.x == 0 && threadIdx.x == 0) *src.next_ctr() = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const uint32_t ci = i / 27u, t = i % 27u;
    uint32_t scene;
    const int32_t self = src.lane_item(ci, &scene);
    const View& V = src.view(scene);
    const i3 b = unpack_key(V.block_keys[self]);
    const int dx = (int)(t % 3) - 1, dy = (int)((t / 3) % 3) - 1, dz = (int)(t / 9) - 1;
    const int32_t p = (dx | dy | dz) == 0 ? self : hash_find_eager(V, i3{b.x + dx, b.y + dy, b.z + dz});
    int32_t e = kFreePtr;
    if (p >= 0) e = p | (V.block_mark[p] == mc_of(src, scene, mc1).pass ? kMeshedBit : 0);
    nbt[i] = e;
  }
}

// nbe: the lane's neighbour-table entry for this block (lanes
// kLookupLane + t, t < 27).
template <bool PSDF>
__device__ void stage_shell(const View& V, const McConst& mc, int32_t nbe, Shell<PSDF>& S, int32_t* nb,
                            uint8_t* meshed) {
  const uint32_t t = threadIdx.x - kLookupLane;
  if (t < 27) {
    nb[t] = nbe < 0 ? nbe : (nbe & (kMeshedBit - 1));
    meshed[t] = nbe >= 0 && (nbe & kMeshedBit) != 0;
  }
  __syncthreads();
  // All of a lane's loads are issued before any is stored to LDS (one
  // memory round trip for the shell), for a 512-lane workgroup.
  constexpr int kPer = (kShellN + 511) / 512;
  float sd[kPer], wt[kPer], va[kPer], vb[kPer];
  bool ok[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int s = (int)threadIdx.x + 512 * r;
    ok[r] = false;
    sd[r] = wt[r] = va[r] = vb[r] = 0.0f;
    if (s >= kShellN) continue;
    const int lx = s % kShell - 1, ly = (s / kShell) % kShell - 1, lz = s / (kShell * kShell) - 1;
    const int dx = lx < 0 ? -1 : (lx >= 8 ? 1 : 0);
    const int dy = ly < 0 ? -1 : (ly >= 8 ? 1 : 0);
    const int dz = lz < 0 ? -1 : (lz >= 8 ? 1 : 0);
    const int32_t p = nb[nb27(dx, dy, dz)];
    if (p < 0) continue;
    const size_t idx = (size_t)p * kBlockVoxels + (size_t)((lx - 8 * dx) + 8 * (ly - 8 * dy) +
                                                           64 * (lz - 8 * dz));
    ok[r] = true;
    sd[r] = V.sdf[idx];
    wt[r] = V.weight[idx];
    if (PSDF) {
      va[r] = V.va[idx];
      vb[r] = V.vb[idx];
    }
  }
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int s = (int)threadIdx.x + 512 * r;
    if (s >= kShellN) continue;
    // the corner tests of IsValid (marching_cubes.cu:96-130), once per
    // voxel here instead of once per cube corner
    const float rho = PSDF ? va[r] / (va[r] + vb[r]) : 0.0f;
    const bool cube_ok = ok[r] && !(fabsf(sd[r]) > mc.sdf_threshold) && (!PSDF || !(rho < mc.rho_min)) &&
                         !(wt[r] < mc.weight_min);
    S.f[s] = (uint8_t)((ok[r] ? kShOk : 0) | (cube_ok ? kShCube : 0) | (sd[r] < 0.0f ? kShNeg : 0));
    S.sdf[s] = sd[r];
    S.w[s] = wt[r];
    if (PSDF) {
      S.a[s] = va[r];
      S.b[s] = vb[r];
    }
  }
  __syncthreads();
}


struct SVal {
  float sdf, a, b, radius;
};

// One axis of a GetSpatialValue sample (spatial_query.h:23-66) at world
// coordinate p: the two corner voxels (rounded from p - vs/2 and p + vs/2,
// exactly as the reference, so ties and negative coordinates pick the same
// voxels), local to the block origin base, and the ratio frac(p / vs).  A
// sample's axes are independent, so the gradient stencils p +- vs/2 e_k
// share two of their three axes with the centre sample.
struct SAxis {
  int c0, c1;
  float r;
};
__device__ __forceinline__ SAxis shell_axis(float p, float vs, int base) {
  const float pc = p - 0.5f * vs;
  const float q = p / vs;
  return SAxis{vox1(pc + 0.0f, vs) - base, vox1(pc + vs, vs) - base, q - floorf(q)};
}

// GetSpatialValue against the staged shell, from the sample's three axes.
template <bool PSDF, bool FULL = true>
__device__ __forceinline__ bool shell_value(const Shell<PSDF>& S, const SAxis& X, const SAxis& Y,
                                            const SAxis& Z, SVal* out, bool* oob) {
  const int cx[2] = {X.c0, X.c1}, cy[2] = {Y.c0, Y.c1}, cz[2] = {Z.c0, Z.c1};
  const float rx = X.r, ry = Y.r, rz = Z.r;
  // Straight-line evaluation of all 8 corners (no early exit: a missing
  // corner makes the result invalid, and the sums are then discarded).
  const bool in = cx[0] >= -1 && cx[1] <= 9 && cy[0] >= -1 && cy[1] <= 9 && cz[0] >= -1 && cz[1] <= 9 &&
                  cx[1] >= -1 && cx[0] <= 9 && cy[1] >= -1 && cy[0] <= 9 && cz[1] >= -1 && cz[0] <= 9;
  if (!in) {
    *oob = true;
    return false;
  }
  bool ok = true;
  float sdf = 0.0f, a = 0.0f, b = 0.0f, rad = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const int s = sidx(cx[mx], cy[my], cz[mz]);
    ok = ok && (S.f[s] & kShOk);
    const float w = (mx ? rx : 1.0f - rx) * (my ? ry : 1.0f - ry) * (mz ? rz : 1.0f - rz);
    sdf += w * S.sdf[s];
    if (FULL) {
      if (PSDF) {
        a += w * S.a[s];
        b += w * S.b[s];
      }
      rad += w * sqrtf(1.0f / S.w[s]);
    }
  }
  out->sdf = sdf;
  out->a = a;
  out->b = b;
  out->radius = rad;
  return ok;
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
  const float vs = g.voxel_size;
  const SAxis X = shell_axis(p.x, vs, base.x), Y = shell_axis(p.y, vs, base.y), Z = shell_axis(p.z, vs, base.z);
  const bool ok = shell_value(S, X, Y, Z, &sv, oob);
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
  // (all six samples are taken; any invalid one zeroes the normal, which is
  // what the reference's short-circuit produces)
  // (p -+ h e_k differs from p on axis k only: the other two axes are the
  // centre sample's)
  const float h = 0.5f * g.voxel_size;
  bool valid = true;
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;  // (sp[k] - sn[k]) / vs per axis, without an indexed array
  // not unrolled: the six stencils' live registers would cost mc_surfel_kernel
  // a workgroup per CU
#pragma unroll 1
  for (int k = 0; k < 3; ++k) {
    const float pk = k == 0 ? p.x : (k == 1 ? p.y : p.z);
    const int bk = k == 0 ? base.x : (k == 1 ? base.y : base.z);
    SVal q{};
    const SAxis An = shell_axis(pk - h, vs, bk);
    valid = shell_value<PSDF, false>(S, k == 0 ? An : X, k == 1 ? An : Y, k == 2 ? An : Z, &q, oob) && valid;
    const float sn = q.sdf;
    const SAxis Ap = shell_axis(pk + h, vs, bk);
    valid = shell_value<PSDF, false>(S, k == 0 ? Ap : X, k == 1 ? Ap : Y, k == 2 ? Ap : Z, &q, oob) && valid;
    const float gk = (q.sdf - sn) / vs;
    g0 = k == 0 ? gk : g0;
    g1 = k == 1 ? gk : g1;
    g2 = k == 2 ? gk : g2;
  }
  const f3 grad = mk3(g0, g1, g2);
  const float l = length(grad);
  const f3 n = (l > 0.0f && valid) ? grad / l : mk3(0.0f, 0.0f, 0.0f);
  vt->nx = n.x;
  vt->ny = n.y;
  vt->nz = n.z;
}

// ---------------------------------------------------------------- pass 1
// Corner checks of the cube at local voxel (x, y, z) (marching_cubes.cu:
// 96-130): every corner must exist, |sdf| <= threshold, rho >= rho_min and
// weight >= weight_min; returns validity and the cube index.
template <bool PSDF>
__device__ __forceinline__ bool cube_index(const Shell<PSDF>& S, int x, int y, int z, int* cube) {
  // The corners' flags (tested once per voxel when the shell was staged),
  // all eight read without early exits, so the lanes of a wave do not
  // diverge on which corner fails first.
  int c = 0;
  uint32_t all = kShCube;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t fl = S.f[sidx(x + corner_off(k, 0), y + corner_off(k, 1), z + corner_off(k, 2))];
    all &= fl;
    c |= (int)((fl >> 2) & 1u) << k;
  }
  *cube = c;
  return all != 0;
}

// Cubes at local coordinates -1..8 (10^3), for the writer rule below.
constexpr int kCubeHalo = 10;
__device__ __forceinline__ int cidx(int x, int y, int z) {
  return (x + 1) + kCubeHalo * (y + 1) + kCubeHalo * kCubeHalo * (z + 1);
}

// Payload of vertex (owner voxel at local (key & 15, key >> 4 & 15,
// key >> 8 & 15), axis key >> 12) into vertex record vp: VertexIntersection
// (marching_cubes.cu:16-28) in the canonical endpoint order, then
// AllocateVertexWithMutex's payload (:45-72).
// The surfel pass queues a block's vertex writes in LDS (a block writes ~40
// on average); past kQueue the lane writes its vertex itself.
constexpr uint32_t kQueue = 1024;
template <bool PSDF>
// Returns true when a stencil sample fell outside the staged shell.
__device__ __forceinline__ bool write_vertex(const Shell<PSDF>& S, const View& V, const McConst& mc, i3 base,
                             uint32_t key, int32_t vp) {
  bool oob = false;
  const Geo& g = V.geo;
  const int ox = (int)(key & 15u), oy = (int)((key >> 4) & 15u), oz = (int)((key >> 8) & 15u);
  const int ax = (int)(key >> 12);
  const float vA = S.sdf[sidx(ox, oy, oz)];
  const float vB = S.sdf[sidx(ox + (ax == 0), oy + (ax == 1), oz + (ax == 2))];
  const f3 pA = voxel_to_world(g, i3{base.x + ox, base.y + oy, base.z + oz});
  const f3 pB = mk3(pA.x + (ax == 0 ? g.voxel_size : 0.0f), pA.y + (ax == 1 ? g.voxel_size : 0.0f),
                    pA.z + (ax == 2 ? g.voxel_size : 0.0f));
  f3 p;
  if ((double)fabsf(vA) < mc.snap) {
    p = pA;
  } else if ((double)fabsf(vB) < mc.snap) {
    p = pB;
  } else {
    const float mu = (0.0f - vA) / (vB - vA);
    p = mk3(pA.x + mu * (pB.x - pA.x), pA.y + mu * (pB.y - pA.y), pA.z + mu * (pB.z - pA.z));
  }
  Vertex vt;
  vertex_payload<PSDF>(S, g, base, p, &vt, &oob);
  Vertex* dst = &V.vertices[vp];
  *reinterpret_cast<float4*>(&dst->px) = make_float4(vt.px, vt.py, vt.pz, vt.radius);
  dst->nx = vt.nx;
  dst->ny = vt.ny;
  dst->nz = vt.nz;
  dst->cr = vt.cr;
  dst->cg = vt.cg;
  dst->cb = vt.cb;
  return oob;
}

// The same payload spread over eight lanes per vertex, so that a vertex
// costs the latency of one trilinear sample instead of seven in a row: lanes
// 0-5 take the gradient stencils p -+ h e_k (k = sub / 2; odd: +h), lane 6
// the full sample at p (radius, colour) and stores position, radius and
// colour; lane 7 is idle.  Stencil values and validities meet in LDS
// (pay_s, pay_ok, row jl), from which lane 0 finishes the normal after the
// workgroup barrier (finish_normal).  Every sample is write_vertex's, with
// the same axes and arithmetic, so the stored bits are the same.
template <bool PSDF>
__device__ __forceinline__ bool payload_part(const Shell<PSDF>& S, const View& V, const McConst& mc,
                                             i3 base, uint32_t key, int32_t vp, uint32_t sub,
                                             float* pay_s, uint8_t* pay_ok) {
  bool oob = false;
  const Geo& g = V.geo;
  const int ox = (int)(key & 15u), oy = (int)((key >> 4) & 15u), oz = (int)((key >> 8) & 15u);
  const int ax = (int)(key >> 12);
  const float vA = S.sdf[sidx(ox, oy, oz)];
  const float vB = S.sdf[sidx(ox + (ax == 0), oy + (ax == 1), oz + (ax == 2))];
  const f3 pA = voxel_to_world(g, i3{base.x + ox, base.y + oy, base.z + oz});
  const f3 pB = mk3(pA.x + (ax == 0 ? g.voxel_size : 0.0f), pA.y + (ax == 1 ? g.voxel_size : 0.0f),
                    pA.z + (ax == 2 ? g.voxel_size : 0.0f));
  f3 p;
  if ((double)fabsf(vA) < mc.snap) {
    p = pA;
  } else if ((double)fabsf(vB) < mc.snap) {
    p = pB;
  } else {
    const float mu = (0.0f - vA) / (vB - vA);
    p = mk3(pA.x + mu * (pB.x - pA.x), pA.y + mu * (pB.y - pA.y), pA.z + mu * (pB.z - pA.z));
  }
  const float vs = g.voxel_size;
  const SAxis X = shell_axis(p.x, vs, base.x), Y = shell_axis(p.y, vs, base.y), Z = shell_axis(p.z, vs, base.z);
  if (sub < 6) {
    const int k = (int)(sub >> 1);
    const float pk = k == 0 ? p.x : (k == 1 ? p.y : p.z);
    const int bk = k == 0 ? base.x : (k == 1 ? base.y : base.z);
    const float h = 0.5f * g.voxel_size;
    const SAxis An = shell_axis((sub & 1u) ? pk + h : pk - h, vs, bk);
    SVal q{};
    const bool ok = shell_value<PSDF, false>(S, k == 0 ? An : X, k == 1 ? An : Y, k == 2 ? An : Z, &q, &oob);
    pay_s[sub] = q.sdf;
    pay_ok[sub] = ok ? 1 : 0;
  } else if (sub == 6) {
    SVal sv{};
    const bool ok = shell_value(S, X, Y, Z, &sv, &oob);
    Vertex* dst = &V.vertices[vp];
    float radius = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;  // the reference reads an uninitialised query
    if (ok) {
      const float inv_sigma2 = 1.0f / (sv.radius * sv.radius);
      radius = sqrtf(1.0f / inv_sigma2);
      const f3 c = val_to_rgb(sv.a / (sv.a + sv.b), 0.4f, 1.0f);
      cr = c.x;
      cg = c.y;
      cb = c.z;
    }
    *reinterpret_cast<float4*>(&dst->px) = make_float4(p.x, p.y, p.z, radius);
    dst->cr = cr;
    dst->cg = cg;
    dst->cb = cb;
  }
  return oob;
}

// GetSpatialSDFGradient's normal from the six stencil values of a vertex
// (pay_s: -x, +x, -y, +y, -z, +z); any invalid sample zeroes it, which is
// what the reference's short-circuit produces.
__device__ __forceinline__ void finish_normal(const View& V, int32_t vp, const float* pay_s,
                                              const uint8_t* pay_ok) {
  const float vs = V.geo.voxel_size;
  const bool valid = pay_ok[0] && pay_ok[1] && pay_ok[2] && pay_ok[3] && pay_ok[4] && pay_ok[5];
  const f3 grad = mk3((pay_s[1] - pay_s[0]) / vs, (pay_s[3] - pay_s[2]) / vs, (pay_s[5] - pay_s[4]) / vs);
  const float l = length(grad);
  const f3 n = (l > 0.0f && valid) ? grad / l : mk3(0.0f, 0.0f, 0.0f);
  Vertex* dst = &V.vertices[vp];
  dst->nx = n.x;
  dst->ny = n.y;
  dst->nz = n.z;
}

// SurfelExtractionKernel (marching_cubes.cu:75-170): cube index of every
// cube of the block, and the vertex of every active edge.  A vertex is
// shared by up to four cubes.  Its payload is written once per pass, by a
// designated cube: the first, in the order (0,0), (1,0), (0,1), (1,1) of
// the offsets across the edge, that is valid and belongs to a block meshed
// in this pass.  Every cube can evaluate that rule from the staged shell (a
// 10^3 halo of cube validities and the meshed flags of the 27 blocks), so
// no atomics are needed; the other cubes only make sure the vertex slot is
// allocated.  Any writer would store the same bits (canonical vertices).

// Work counters of one scene, flushed per wave.
struct McCounts {
  uint32_t writes = 0, assigned = 0, blocks = 0;
  int delta = 0;
  bool oob = false;
  __device__ void flush(const View& V) {
    if (__any(oob)) {
      if ((threadIdx.x & 63u) == 0) flag(V, MH_STATUS_STENCIL_RANGE);
    }
    wave_add_u64(&work_stripe(V.ctr).mc_vertex_writes, writes);
    wave_add_u64(&work_stripe(V.ctr).mc_triangle_writes, assigned);
    wave_add_u64(&work_stripe(V.ctr).mc_blocks, blocks);
    for (int off = 32; off > 0; off >>= 1) delta += __shfl_xor(delta, off);
    if ((threadIdx.x & 63u) == 0 && delta)
      atomicAdd(&work_stripe(V.ctr).tri_live, (unsigned long long)(long long)delta);
    *this = McCounts{};
  }
};

// Exclusive scan over the wave of values v whose set bits all lie in PLANES,
// one bit plane at a time (a ballot and a masked bit count per plane, no lane
// shuffles); *all gets the wave's sum.  Packed fields scan independently as
// long as no field's sum carries into the next one.
template <uint32_t PLANES>
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* all) {
  uint32_t before = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (!((PLANES >> k) & 1u)) continue;
    const uint64_t m = __ballot((v >> k) & 1u);
    before += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << k;
    sum += (uint32_t)__popcll(m) << k;
  }
  *all = sum;
  return before;
}

// Exclusive scan of one value per lane over the workgroup (up to eight
// waves), its set bits in PLANES; *total gets the sum.  Every lane must call
// it; it synchronises the workgroup twice.
template <uint32_t PLANES>
__device__ __forceinline__ uint32_t block_excl_scan_planes(uint32_t v, uint32_t* wave_sums, uint32_t* total) {
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t wsum;
  const uint32_t x = wave_excl_scan<PLANES>(v, &wsum);
  if ((threadIdx.x & 63u) == 0) wave_sums[wave] = wsum;
  __syncthreads();
  uint32_t before = 0, all = 0;
  const uint32_t nw = blockDim.x >> 6;
  for (uint32_t w = 0; w < nw; ++w) {
    const uint32_t ws = wave_sums[w];
    if (w < wave) before += ws;
    all += ws;
  }
  __syncthreads();
  *total = all;
  return before + x;
}
// The same for values v < 2^BITS.
template <int BITS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wave_sums, uint32_t* total) {
  return block_excl_scan_planes<(1u << BITS) - 1u>(v, wave_sums, total);
}

// Vertex slot of edge e of the cube at local voxel o: the slot of the edge's
// owner voxel (o + the edge's owner offset, possibly in a +x/+y/+z neighbour).
__device__ __forceinline__ int32_t* edge_slot(const View& V, const int32_t* nb, i3 o, int e) {
  const int ox = o.x + edge_owner_off(e, 0), oy = o.y + edge_owner_off(e, 1), oz = o.z + edge_owner_off(e, 2);
  const int dx = ox >> 3, dy = oy >> 3, dz = oz >> 3;
  const int32_t optr = nb[nb27(dx, dy, dz)];
  const uint32_t oi = (uint32_t)((ox - 8 * dx) + 8 * (oy - 8 * dy) + 64 * (oz - 8 * dz));
  return &V.vptr[((size_t)optr * kBlockVoxels + oi) * 3 + edge_axis(e)];
}

template <bool PSDF, class Src>
__global__ __launch_bounds__(512, 6) void mc_surfel_kernel(Src src_, const SceneFrame* __restrict__ tab, McConst mc1,
                                                           const int32_t* __restrict__ nbt, uint32_t static_q) {
  const Src src = with_table(src_, tab);
  __shared__ Shell<PSDF> S;
  __shared__ int32_t nb[27];
  __shared__ uint8_t meshed[27];
  __shared__ uint8_t cv[kCubeHalo * kCubeHalo * kCubeHalo];
  __shared__ uint8_t own_cube[kBlockVoxels];
  __shared__ int surfels[2];
  __shared__ uint32_t free_base;
  __shared__ int32_t qptr[kQueue];
  __shared__ uint16_t qkey[kQueue];
  __shared__ uint32_t qn;
  __shared__ uint32_t scan_wave[8];
  __shared__ int32_t pop_top;
  // One LDS region, three uses that never overlap in time: the
  // configurations of the cubes at local -1..7 (this block and its lower
  // neighbours) from the halo loop to the recycling test; the item list from
  // the item scan to the end of the slot claims; the payload exchange (eight
  // lanes per vertex, kPayRound vertices a round) in the payload rounds.
  // Barriers separate each use from the next.
  constexpr uint32_t kPayRound = kLookupLane / 8;
  __shared__ uint32_t region[6 * kBlockVoxels];
  uint16_t* const items = reinterpret_cast<uint16_t*>(region);  // up to 12 per cube
  uint8_t* const hcfg = reinterpret_cast<uint8_t*>(region);
  float* const pay_s = reinterpret_cast<float*>(region);
  uint8_t* const pay_ok = reinterpret_cast<uint8_t*>(region + kPayRound * 8);
  static_assert(kCfg * kCfg * kCfg <= 4 * 6 * kBlockVoxels && kPayRound * 8 * 5 <= 4 * 6 * kBlockVoxels, "region");
  __shared__ uint32_t nitems;
  // per-configuration crossed edges and triangle counts, read per lane with
  // divergent indices: from LDS rather than the constant tables
  __shared__ uint16_t etab[256];
  __shared__ uint8_t ntab[256];
  if (threadIdx.x < 256) {
    etab[threadIdx.x] = kMc.edges[threadIdx.x];
    ntab[threadIdx.x] = (uint8_t)kMc.ntri[threadIdx.x];
  }
  const uint32_t n = src.count();
  McCounts cnt;
  uint32_t cur_scene = ~0u;
  // lanes kLookupLane + t hold neighbour t's table entry of the next block
  const uint32_t nb_lane = threadIdx.x - kLookupLane;
  int32_t nbe = kFreePtr;
  if (nb_lane < 27 && blockIdx.x < n) nbe = nbt[(size_t)blockIdx.x * 27u + nb_lane];
  const uint32_t n_static = mc_static_items(n, gridDim.x, static_q);
  __shared__ uint32_t s_next;
  for (uint32_t ci = blockIdx.x; ci < n;) {
    // the next item, read by every lane after the shell's barriers and
    // rewritten only after the next barrier
    if (threadIdx.x == 0)
      s_next = ci + gridDim.x < n_static ? ci + gridDim.x : n_static + atomicAdd(src.next_ctr(), 1u);
    uint32_t scene;
    const int32_t ptr = src.item(ci, &scene);
    if (scene != cur_scene) {
      if (cur_scene != ~0u) cnt.flush(src.view(cur_scene));
      cur_scene = scene;
    }
    const View& V = src.view(scene);
    const McConst& mc = mc_of(src, scene, mc1);
    cnt.blocks += threadIdx.x == 0;
    const i3 b = unpack_key(V.block_keys[ptr]);
    const uint32_t i = threadIdx.x;
    const size_t idx = (size_t)ptr * kBlockVoxels + i;
    const uint32_t prev = V.cube[idx] & 255u;  // in flight with the shell loads
    stage_shell<PSDF>(V, mc, nbe, S, nb, meshed);
    const uint32_t next = s_next;
    int32_t vps[3];  // this voxel's vertex slots (x, y, z edges)
#pragma unroll
    for (int a = 0; a < 3; ++a) vps[a] = V.vptr[3 * idx + a];
    if (threadIdx.x == 0) {
      qn = 0;
      surfels[0] = surfels[1] = 0;
    }
    __syncthreads();
    // every cube of the 10^3 halo, the block's own among them: validity for
    // the writer rule, and the configuration of the block's own cubes
    for (int s = (int)threadIdx.x; s < kCubeHalo * kCubeHalo * kCubeHalo; s += kBlockVoxels) {
      const int x = s % kCubeHalo - 1, y = (s / kCubeHalo) % kCubeHalo - 1, z = s / (kCubeHalo * kCubeHalo) - 1;
      const int dx = x < 0 ? -1 : (x >= 8 ? 1 : 0), dy = y < 0 ? -1 : (y >= 8 ? 1 : 0),
                dz = z < 0 ? -1 : (z >= 8 ? 1 : 0);
      int c;
      const bool valid = cube_index(S, x, y, z, &c);
      const int q = nb27(dx, dy, dz);
      cv[s] = meshed[q] && valid;
      if ((dx | dy | dz) == 0) own_cube[x + 8 * y + 64 * z] = (uint8_t)(valid ? c : 0);
      if (dx <= 0 && dy <= 0 && dz <= 0) {
        // what the block's cube word holds after this pass: computed here
        // for a block meshed in it (its own pass computes the same bits from
        // the same voxels), stored for one that is not
        uint32_t hc = 0;
        if (meshed[q])
          hc = valid ? (uint32_t)c : 0u;
        else if (nb[q] >= 0)
          hc = V.cube[(size_t)nb[q] * kBlockVoxels + (x - 8 * dx) + 8 * (y - 8 * dy) + 64 * (z - 8 * dz)] & 255u;
        hcfg[(x + 1) + kCfg * (y + 1) + kCfg * kCfg * (z + 1)] = (uint8_t)hc;
      }
    }
    __syncthreads();
    // RecycleVerticesKernel on this block's own slots (marching_cubes.cu:
    // 303-321): a slot is referenced iff one of the four cubes around its edge
    // (this block and its lower neighbours, whose configurations are all
    // final above) crosses the edge.  Unreferenced allocated slots are freed
    // (to the free list; they go on the heap after the pass, so no pop of
    // this pass can take them).  The surfel counts are the slots allocated
    // once the pass's claims are done: allocated now, or referenced (a
    // referenced free slot is claimed in this pass, here or by a lower
    // neighbour).  No other workgroup touches an unreferenced slot.
    unsigned release = 0;
    {
      const i3 o = devectorize(i);
      const bool is_inner = o.x >= 1 && o.y >= 1 && o.z >= 1 && o.x < 7 && o.y < 7 && o.z < 7;
      int inner = 0, boundary = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
        bool used = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          int d[3] = {0, 0, 0};
          d[b1] = k & 1;
          d[b2] = k >> 1;
          const unsigned m = etab[hcfg[(o.x - d[0] + 1) + kCfg * (o.y - d[1] + 1) + kCfg * kCfg * (o.z - d[2] + 1)]];
          used = used || ((m >> edge_with(d[0], d[1], d[2], a)) & 1u) != 0;
        }
        if (vps[a] >= 0 || used) {
          if (is_inner) ++inner;
          else ++boundary;
        }
        if (vps[a] >= 0 && !used) release |= 1u << a;
      }
      if (inner) atomicAdd(&surfels[0], inner);
      if (boundary) atomicAdd(&surfels[1], boundary);
    }
    const i3 base = i3{b.x * kBlockSide, b.y * kBlockSide, b.z * kBlockSide};
    const int cube = own_cube[i];
    V.cube[idx] = (uint16_t)((uint32_t)cube | (prev << 8));
    // live triangles of a cube = its configuration's table row
    const int nt = ntab[cube];
    cnt.delta += nt - ntab[prev];
    if ((uint32_t)cube != prev) cnt.assigned += (uint32_t)nt;
    const unsigned em = etab[cube];  // 0 for configurations 0 and 255
    // The crossed edges of the block's cubes become a dense item list
    // (voxel | edge << 9), so the slot work below runs one item per lane
    // instead of twelve edge rounds on every cube.  The same scan ranks the
    // freed slots (bits 13 and up: the item counts sum to at most 6144).
    uint32_t free_rank;
    {
      const uint32_t mine = (uint32_t)__popc(em) | ((uint32_t)__popc(release) << 13);  // <= 12, <= 3
      uint32_t total;
      const uint32_t at2 = block_excl_scan_planes<0xFu | (0x3u << 13)>(mine, scan_wave, &total);
      uint32_t at = at2 & 8191u;
      free_rank = at2 >> 13;
      for (unsigned m = em; m; m &= m - 1u) items[at++] = (uint16_t)(i | ((uint32_t)(__ffs(m) - 1) << 9));
      if (threadIdx.x == 0) {
        nitems = total & 8191u;
        if (total >> 13) free_base = atomicAdd(&V.ctr->vfree_count, total >> 13);
      }
    }
    __syncthreads();
    const uint32_t n_items = nitems;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!((release >> a) & 1u)) continue;
      const int32_t vp = vps[a];
      Vertex zv{};
      V.vertices[vp] = zv;
      V.vptr[3 * idx + a] = kFreePtr;
      V.vfree[free_base + free_rank] = vp;
      ++free_rank;
    }
    if (threadIdx.x == 0) {
      V.surfel[2 * ptr + 0] = surfels[0];
      V.surfel[2 * ptr + 1] = surfels[1];
    }
    for (uint32_t c0 = 0; c0 < n_items; c0 += kBlockVoxels) {
      // Vertex slot of each item: read it; claim it if free (CAS to the
      // lock value); pop fresh vertices for all claims of the workgroup
      // with one heap operation; publish them.
      const uint32_t jt = c0 + threadIdx.x;
      const bool has = jt < n_items;
      i3 oo{0, 0, 0};
      int e = 0;
      int32_t* slot = nullptr;
      int32_t cur = kFreePtr;
      bool claimed = false;
      if (has) {
        const uint32_t it = items[jt];
        oo = devectorize(it & 511u);
        e = (int)(it >> 9);
        slot = edge_slot(V, nb, oo, e);
        cur = *slot;
        if (cur == kFreePtr) {
          const int32_t old = atomicCAS(slot, kFreePtr, kLockPtr);
          claimed = old == kFreePtr;
          if (!claimed) cur = old;
        }
      }
      uint32_t nclaims;
      const uint32_t rank = block_excl_scan<1>(claimed ? 1u : 0u, scan_wave, &nclaims);
      if (nclaims) {  // uniform
        if (threadIdx.x == 0) {
          const int32_t top = atomicSub(&V.ctr->vertex_top, (int32_t)nclaims);
          const int32_t avail = top + 1 > 0 ? top + 1 : 0;
          if (avail < (int32_t)nclaims) {
            atomicAdd(&V.ctr->vertex_top, (int32_t)nclaims - avail);
            atomicOr(&V.ctr->flags, MH_STATUS_VERTEX_HEAP_EXHAUSTED);
          }
          pop_top = top;
        }
        __syncthreads();
        if (claimed) {
          const int32_t hidx = pop_top - (int32_t)rank;
          const int32_t fresh = hidx >= 0 ? V.vertex_heap[hidx] : kFreePtr;
          atomicExch(slot, fresh);
          cur = fresh;
        }
      }
      if (!has || cur < 0) continue;
      const int ax = edge_axis(e);
      const int offx = edge_owner_off(e, 0), offy = edge_owner_off(e, 1), offz = edge_owner_off(e, 2);
      const int ox = oo.x + offx, oy = oo.y + offy, oz = oo.z + offz;
      // writer rule: offsets across the edge in the two other axes (b1 < b2;
      // selected without indexing, e being a run-time value here)
      const int b1 = ax == 0 ? 1 : 0, b2 = ax == 2 ? 1 : 2;
      const int own1 = b1 == 0 ? offx : offy, own2 = b2 == 1 ? offy : offz;
      const int kself = own1 + 2 * own2;
      // A slot claimed in this pass is also written by its claimer: the
      // designated cube may have read it while still locked and skipped it.
      bool writer = true;
      for (int k = 0; k < kself && writer && !claimed; ++k) {
        const int d1 = k & 1, d2 = k >> 1;
        const int qx = ox - (b1 == 0 ? d1 : 0);
        const int qy = oy - (b1 == 1 ? d1 : 0) - (b2 == 1 ? d2 : 0);
        const int qz = oz - (b2 == 2 ? d2 : 0);
        if (cv[cidx(qx, qy, qz)]) writer = false;
      }
      if (!writer) continue;
      // queue the payload; all lanes of the block work the queue below
      const uint32_t q = atomicAdd(&qn, 1u);
      const uint32_t key = (uint32_t)ox | ((uint32_t)oy << 4) | ((uint32_t)oz << 8) | ((uint32_t)ax << 12);
      if (q < kQueue) {
        qptr[q] = cur;
        qkey[q] = (uint16_t)key;
      } else {
        cnt.oob = write_vertex<PSDF>(S, V, mc, base, key, cur) || cnt.oob;
        ++cnt.writes;
      }
    }
    __syncthreads();
    if (nb_lane < 27 && next < n) nbe = nbt[(size_t)next * 27u + nb_lane];
    const uint32_t nq = qn < kQueue ? qn : kQueue;
    // payloads, eight lanes per vertex, on the waves before the last one
    // (which looks up the next block's neighbours meanwhile)
    for (uint32_t j0 = 0; j0 < nq; j0 += kPayRound) {  // uniform
      const uint32_t jl = threadIdx.x >> 3, sub = threadIdx.x & 7u;
      const uint32_t j = j0 + jl;
      const bool act = threadIdx.x < kLookupLane && j < nq;
      if (act)
        cnt.oob = payload_part<PSDF>(S, V, mc, base, qkey[j], qptr[j], sub, &pay_s[8 * jl], &pay_ok[8 * jl]) ||
                  cnt.oob;
      __syncthreads();
      if (act && sub == 0) {
        finish_normal(V, qptr[j], &pay_s[8 * jl], &pay_ok[8 * jl]);
        ++cnt.writes;
      }
      __syncthreads();
    }
    ci = next;
  }
  if (cur_scene != ~0u) cnt.flush(src.view(cur_scene));
}

// Meshed-in-this-pass marks for the writer rule.
template <class Src>
__global__ void mc_mark_kernel(Src src_, const SceneFrame* __restrict__ tab, McConst mc1) {
  const Src src = with_table(src_, tab);
  const uint32_t n = src.count();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t scene;
    const int32_t ptr = src.lane_item(i, &scene);
    src.view(scene).block_mark[ptr] = mc_of(src, scene, mc1).pass;
  }
}

// ---------------------------------------------------------------- pass 2
// The reference's TriangleExtractionKernel (marching_cubes.cu:180-275) and
// RecycleTrianglesKernel (:278-301) maintain a triangle pool: after a pass
// every processed cube holds exactly the triangles of its configuration's
// table row, over the vertices of its active edges, and keeps them until it
// is processed again.  That state is already held by the cube configuration
// and the per-voxel vertex slots, so the mesh is kept implicitly: no pool,
// no per-triangle writes, no reference-count atomics.  The triangle list
// is materialised only when the mesh is read out (dumps, compaction).  A
// vertex's reference count is likewise the sum over the four cubes around
// its edge of how many of their table triangles use that edge (edge_refs).
//
// RecycleVerticesKernel (marching_cubes.cu:303-321) frees a block's
// vertices with no references and produces its inner/boundary surfel counts
// (TriangleExtractionKernel counts the allocated vertex slots before
// recycling, :187-205).  The surfel pass does this for the blocks it meshes;
// this kernel does it for a list of other blocks (the garbage blocks, after
// their cubes lost their triangles).  Cube configurations of the block and of
// its -x/-y/-z neighbours (voxels -1..7) are staged in LDS.
template <class Src>
__global__ __launch_bounds__(256) void mc_recycle_vertices_kernel(Src src_, const SceneFrame* __restrict__ tab) {
  const Src src = with_table(src_, tab);
  __shared__ int32_t nb[8];  // blocks at offsets (dx, dy, dz) in {-1, 0}^3
  // edges used by the triangles of each staged cube's configuration
  __shared__ uint16_t cfg[kCfg * kCfg * kCfg];
  __shared__ uint16_t emask[256];
  __shared__ int counts[2];
  __shared__ uint32_t wave_sums[4];
  __shared__ int32_t push_top;
  const uint32_t n = src.count();
  static_assert(kBlockVoxels == 2 * 256, "two voxels per lane");
  // A configuration's table row uses exactly the edges of kMc.edges (its
  // crossed edges), so a vertex has references iff one of the four cubes
  // around its edge has that edge's bit set (edge_refs > 0).
  emask[threadIdx.x] = kMc.edges[threadIdx.x];
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    uint32_t scene;
    const int32_t ptr = src.item(ci, &scene);
    const View V = src.view(scene);  // a copy: read once, not reloaded after stores
    // This lane's vertex slots (voxels threadIdx.x and threadIdx.x + 256);
    // a block without any is done at once (no slot to free, no surfels).
    int32_t vps[2][3];
    bool any = false;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        vps[r][ax] = V.vptr[3 * ((size_t)ptr * kBlockVoxels + threadIdx.x + 256u * r) + ax];
        any = any || vps[r][ax] >= 0;
      }
    if (!__syncthreads_or(any)) {
      if (threadIdx.x == 0) {
        V.surfel[2 * ptr + 0] = 0;
        V.surfel[2 * ptr + 1] = 0;
      }
      continue;
    }
    if (threadIdx.x < 8) {
      const int t = (int)threadIdx.x;
      if (t == 0) {
        nb[t] = ptr;
      } else {
        const i3 b = unpack_key(V.block_keys[ptr]);
        nb[t] = hash_find(V, i3{b.x - (t & 1), b.y - ((t >> 1) & 1), b.z - ((t >> 2) & 1)});
      }
    }
    if (threadIdx.x == 0) { counts[0] = 0; counts[1] = 0; }
    __syncthreads();
    {  // all of a lane's loads in flight together (256 lanes)
      constexpr int kPer = (kCfg * kCfg * kCfg + 255) / 256;
      uint16_t cw[kPer];
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
        const int s = (int)threadIdx.x + 256 * r;
        cw[r] = 0;
        if (s >= kCfg * kCfg * kCfg) continue;
        const int x = s % kCfg - 1, y = (s / kCfg) % kCfg - 1, z = s / (kCfg * kCfg) - 1;
        const int nx = x < 0, ny = y < 0, nz = z < 0;
        const int32_t cp = nb[nx + 2 * ny + 4 * nz];
        if (cp >= 0) cw[r] = V.cube[(size_t)cp * kBlockVoxels + (x + 8 * nx) + 8 * (y + 8 * ny) + 64 * (z + 8 * nz)];
      }
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
        const int s = (int)threadIdx.x + 256 * r;
        if (s < kCfg * kCfg * kCfg) cfg[s] = emask[cw[r] & 255];
      }
    }
    __syncthreads();
    int inner = 0, boundary = 0;
    unsigned release = 0;  // bit 3r + ax: slot to free
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t i = threadIdx.x + 256u * r;
      const i3 o = devectorize(i);
      const bool is_inner = o.x >= 1 && o.y >= 1 && o.z >= 1 && o.x < 7 && o.y < 7 && o.z < 7;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const int32_t vp = vps[r][ax];
        bool used = false;
        if (vp >= 0) {
          if (is_inner) ++inner; else ++boundary;
          const int b1 = ax == 0 ? 1 : 0, b2 = ax == 2 ? 1 : 2;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            int d[3] = {0, 0, 0};
            d[b1] = k & 1;
            d[b2] = k >> 1;
            const unsigned m = cfg[(o.x - d[0] + 1) + kCfg * (o.y - d[1] + 1) + kCfg * kCfg * (o.z - d[2] + 1)];
            used |= ((m >> edge_with(d[0], d[1], d[2], ax)) & 1u) != 0;
          }
        }
        // (the reference count itself is not stored: readers derive it with
        // edge_refs)
        if (vp >= 0 && !used) release |= 1u << (3 * r + ax);
      }
    }
    // the block's freed slots go back to the heap with one counter update
    uint32_t nfree;
    uint32_t rank = block_excl_scan<3>((uint32_t)__popc(release), wave_sums, &nfree);  // <= 6
    if (nfree) {  // uniform
      if (threadIdx.x == 0) push_top = atomicAdd(&V.ctr->vertex_top, (int32_t)nfree);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          if (!((release >> (3 * r + ax)) & 1u)) continue;
          const int32_t vp = vps[r][ax];
          Vertex z{};
          V.vertices[vp] = z;
          V.vptr[3 * ((size_t)ptr * kBlockVoxels + threadIdx.x + 256u * r) + ax] = kFreePtr;
          V.vertex_heap[push_top + 1 + (int32_t)rank] = vp;
          ++rank;
        }
    }
    if (inner) atomicAdd(&counts[0], inner);
    if (boundary) atomicAdd(&counts[1], boundary);
    __syncthreads();
    if (threadIdx.x == 0) {
      V.surfel[2 * ptr + 0] = counts[0];
      V.surfel[2 * ptr + 1] = counts[1];
    }
    __syncthreads();
  }
}

// The vertex slots the surfel pass freed go on the heap once the pass is over,
// so that no claim of the pass could pop one (pops and pushes of one stack
// cannot overlap).  Workgroups (x, scene y) copy strided parts of the scene's
// free list above the top; the last to finish moves the top.
template <class Src>
__global__ __launch_bounds__(256) void mc_push_freed_kernel(Src src_, const SceneFrame* __restrict__ tab) {
  const Src src = with_table(src_, tab);
  const View& V = src.view(blockIdx.y);
  // both unchanged until every workgroup of the scene has counted itself done
  const uint32_t n = V.ctr->vfree_count;
  if (n == 0) return;  // the same for all of the scene's workgroups
  const int32_t top = V.ctr->vertex_top;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    V.vertex_heap[top + 1 + (int32_t)i] = V.vfree[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&V.ctr->vfree_done, 1u) == gridDim.x - 1) {
      V.ctr->vertex_top = top + (int32_t)n;
      V.ctr->vfree_count = 0;
      V.ctr->vfree_done = 0;
    }
  }
}
constexpr int kPushGrid = 64;  // workgroups per scene

// ---------------------------------------------------------------- recycling
// RecycleGarbageTrianglesKernel (mapping/recycle.cu:117-141): the garbage
// blocks' cubes lose their triangles (current configuration 0; the previous
// one stays in the high byte), so the reference counts of their vertices
// drop accordingly; RecycleGarbageVerticesKernel (:143-179) then frees the
// blocks' own vertex slots left without references (the MC recycle kernel
// over the garbage list).
__global__ __launch_bounds__(512) void gc_release_kernel(View V) {
  const uint32_t n = V.ctr->gc_count;
  int delta = 0;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const size_t idx = (size_t)V.gc_list[ci] * kBlockVoxels + threadIdx.x;
    const uint32_t c = V.cube[idx] & 255u;
    V.cube[idx] = (uint16_t)(c << 8);
    delta -= (c == 0 || c == 255) ? 0 : kMc.ntri[c];
  }
  for (int off = 32; off > 0; off >>= 1) delta += __shfl_xor(delta, off);
  if ((threadIdx.x & 63u) == 0 && delta)
    atomicAdd(&work_stripe(V.ctr).tri_live, (unsigned long long)(long long)delta);
}

void launch_gc_release_triangles(const View& V, hipStream_t s, int grid) {
  gc_release_kernel<<<grid, 512, 0, s>>>(V);
}

// ---------------------------------------------------------------- compaction
// CompressMesh (visualization/compress_mesh.cu:10-172) over the implicit
// mesh: the live triangles of a worklist block are its cubes' table rows;
// a triangle is kept when all three vertex slots are allocated.  Pass 1
// counts triangles per block and marks the vertex slots they use; after
// exclusive scans (triangles over the worklist, used slots over the vertex
// pool) pass 2 writes triangles in worklist x voxel x slot order with the
// vertices renumbered in pool order, and the used vertices are gathered.

// Vertex pool index of edge e of the cube at local voxel o, -1 if its slot is
// not allocated.  nb holds the blocks at offsets {0, 1}^3 (x + 2y + 4z).
__device__ __forceinline__ int32_t cube_edge_vertex(const View& V, const int32_t* nb, i3 o, int e) {
  const int ox = o.x + edge_owner_off(e, 0), oy = o.y + edge_owner_off(e, 1), oz = o.z + edge_owner_off(e, 2);
  const int dx = ox >> 3, dy = oy >> 3, dz = oz >> 3;
  const int32_t bp = nb[dx + 2 * dy + 4 * dz];
  if (bp < 0) return -1;
  const uint32_t li = (uint32_t)((ox - 8 * dx) + 8 * (oy - 8 * dy) + 64 * (oz - 8 * dz));
  return V.vptr[((size_t)bp * kBlockVoxels + li) * 3 + edge_axis(e)];
}

template <bool EMIT>
__global__ __launch_bounds__(512) void compact_tris_kernel(View V, uint32_t n, uint32_t* __restrict__ used,
                                                           uint32_t* __restrict__ blk_tris,
                                                           const uint32_t* __restrict__ blk_off,
                                                           const uint32_t* __restrict__ remap,
                                                           int32_t* __restrict__ tri_out) {
  __shared__ int32_t nb[8];
  __shared__ uint32_t wave_sums[8];
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const i3 b = unpack_key(V.block_keys[ptr]);
    if (threadIdx.x < 8) {
      const int t = (int)threadIdx.x;
      nb[t] = t == 0 ? ptr : hash_find(V, i3{b.x + (t & 1), b.y + ((t >> 1) & 1), b.z + ((t >> 2) & 1)});
    }
    __syncthreads();
    const uint32_t i = threadIdx.x;
    const i3 o = devectorize(i);
    const int c = V.cube[(size_t)ptr * kBlockVoxels + i] & 255;
    const int nt = (c == 0 || c == 255) ? 0 : kMc.ntri[c];
    uint32_t keep = 0;
    for (int t = 0; t < nt; ++t) {
      int32_t vp[3];
      bool ok = true;
      for (int k = 0; k < 3; ++k) {
        vp[k] = cube_edge_vertex(V, nb, o, kMc.tri[c][3 * t + k]);
        ok = ok && vp[k] >= 0;
      }
      if (!ok) continue;
      if (!EMIT) {
        for (int k = 0; k < 3; ++k) used[vp[k]] = 1u;
      } else {
        // slot within the block filled below after the scan
      }
      ++keep;
    }
    uint32_t total;
    const uint32_t excl = block_excl_scan<3>(keep, wave_sums, &total);  // keep <= 5
    if (!EMIT) {
      if (threadIdx.x == 0) blk_tris[ci] = total;
    } else {
      uint32_t at = blk_off[ci] + excl;
      for (int t = 0; t < nt; ++t) {
        int32_t vp[3];
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
          vp[k] = cube_edge_vertex(V, nb, o, kMc.tri[c][3 * t + k]);
          ok = ok && vp[k] >= 0;
        }
        if (!ok) continue;
        for (int k = 0; k < 3; ++k) tri_out[3 * (size_t)at + k] = (int32_t)remap[vp[k]];
        ++at;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void compact_vertices_kernel(View V, const uint32_t* __restrict__ used,
                                                               const uint32_t* __restrict__ remap,
                                                               float* __restrict__ pos, float* __restrict__ nrm,
                                                               float* __restrict__ col, uint32_t nv) {
  for (uint32_t vp = blockIdx.x * blockDim.x + threadIdx.x; vp < nv; vp += gridDim.x * blockDim.x) {
    if (!used[vp]) continue;
    const size_t j = remap[vp];
    const Vertex& x = V.vertices[vp];
    pos[3 * j + 0] = x.px; pos[3 * j + 1] = x.py; pos[3 * j + 2] = x.pz;
    nrm[3 * j + 0] = x.nx; nrm[3 * j + 1] = x.ny; nrm[3 * j + 2] = x.nz;
    col[3 * j + 0] = x.cr; col[3 * j + 1] = x.cg; col[3 * j + 2] = x.cb;
  }
}

// Exclusive scan of n uint32 (three launches: 1024-element chunks, the chunk
// sums in one workgroup, then the chunk offsets added back).
constexpr uint32_t kScanChunk = 1024;
__global__ __launch_bounds__(1024) void scan_chunks_kernel(const uint32_t* __restrict__ in,
                                                           uint32_t* __restrict__ out, uint32_t n,
                                                           uint32_t* __restrict__ chunk_sums) {
  __shared__ uint32_t ws[16];
  const uint32_t i = blockIdx.x * kScanChunk + threadIdx.x;
  const uint32_t v = i < n ? in[i] : 0u;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) ws[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < 16; ++w) {
    if (w < wave) before += ws[w];
    all += ws[w];
  }
  if (i < n) out[i] = before + x - v;
  if (threadIdx.x == 0) chunk_sums[blockIdx.x] = all;
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(uint32_t* __restrict__ sums, uint32_t nchunks,
                                                         uint32_t* __restrict__ total) {
  __shared__ uint32_t ws[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t base = 0; base < nchunks; base += kScanChunk) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nchunks ? sums[i] : 0u;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < 16; ++w) {
      if (w < wave) before += ws[w];
      all += ws[w];
    }
    const uint32_t c = carry;
    if (i < nchunks) sums[i] = c + before + x - v;
    __syncthreads();
    if (threadIdx.x == 0) carry = c + all;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void scan_add_kernel(uint32_t* __restrict__ out, uint32_t n,
                                                        const uint32_t* __restrict__ sums) {
  const uint32_t i = blockIdx.x * kScanChunk + threadIdx.x;
  if (i < n) out[i] += sums[blockIdx.x];
}

void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* chunk_sums,
                 uint32_t* total, hipStream_t s) {
  const uint32_t nchunks = n ? (n + kScanChunk - 1) / kScanChunk : 1;
  scan_chunks_kernel<<<nchunks, kScanChunk, 0, s>>>(in, out, n, chunk_sums);
  scan_sums_kernel<<<1, kScanChunk, 0, s>>>(chunk_sums, nchunks, total);
  scan_add_kernel<<<nchunks, kScanChunk, 0, s>>>(out, n, chunk_sums);
}

void launch_compact_count(const View& V, uint32_t n, uint32_t* used, uint32_t* blk_tris,
                          hipStream_t s, int grid) {
  compact_tris_kernel<false><<<grid, 512, 0, s>>>(V, n, used, blk_tris, nullptr, nullptr, nullptr);
}

void launch_compact_emit(const View& V, uint32_t n, const uint32_t* blk_off, const uint32_t* remap,
                         int32_t* tri_out, const uint32_t* used, float* pos, float* nrm, float* col,
                         hipStream_t s, int grid) {
  compact_tris_kernel<true><<<grid, 512, 0, s>>>(V, n, nullptr, nullptr, blk_off, remap, tri_out);
  compact_vertices_kernel<<<grid, 256, 0, s>>>(V, used, remap, pos, nrm, col, V.max_vertices);
}

// Workgroups of persistent kernel K (THREADS lanes) resident at once: cus x
// the occupancy the runtime reports for it (queried once per kernel).  A
// persistent grid of exactly that spreads its items over one round of
// workgroups; a larger one would run its surplus as a second, partial round.
template <auto K, int THREADS>
int resident_grid(int cus) {
  static const int per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, K, THREADS, 0) != hipSuccess || n < 1) n = 1;
    return n;
  }();
  return per_cu * cus;
}

// Surfel-pass grid: shape.grid_per_cu workgroups per CU (0: as many as are
// resident), at most the resident count.
template <auto K>
int surfel_grid(int cus, const McShape& shape) {
  const int resident = resident_grid<K, kBlockVoxels>(cus);
  const int g = shape.grid_per_cu > 0 ? shape.grid_per_cu * cus : resident;
  return g < resident ? g : resident;
}

void launch_marching_cubes(const View& V, const McConst& p, int32_t* nbt, hipStream_t s, int cus,
                           const KernelClock* kc, McShape shape) {
  const int grid = V.va ? surfel_grid<mc_surfel_kernel<true, One>>(cus, shape)
                        : surfel_grid<mc_surfel_kernel<false, One>>(cus, shape);
  const uint32_t q = (uint32_t)shape.static_quarters;
  tick(kc, s, "mc_mark_kernel");
  mc_mark_kernel<<<cus, 256, 0, s>>>(One{V}, nullptr, p);
  tick(kc, s, "mc_neighbours_kernel");
  mc_neighbours_kernel<<<cus * 8, 256, 0, s>>>(One{V}, nullptr, p, nbt);
  tick(kc, s, "mc_surfel_kernel");
  if (V.va)
    mc_surfel_kernel<true><<<grid, kBlockVoxels, 0, s>>>(One{V}, nullptr, p, nbt, q);
  else
    mc_surfel_kernel<false><<<grid, kBlockVoxels, 0, s>>>(One{V}, nullptr, p, nbt, q);
  tick(kc, s, "mc_push_freed_kernel");
  mc_push_freed_kernel<<<kPushGrid, 256, 0, s>>>(One{V}, nullptr);
  tick(kc, s, nullptr);
}

void launch_gc_recycle_vertices(const View& V, hipStream_t s, int cus) {
  mc_recycle_vertices_kernel<<<resident_grid<mc_recycle_vertices_kernel<Listed>, 256>(cus), 256, 0, s>>>(
      Listed{V, V.gc_list, &V.ctr->gc_count}, nullptr);
}

void launch_batch_marching_cubes(const BatchDev& B, int32_t* nbt, hipStream_t s, int cus, McShape shape,
                                 bool marks_stamped) {
  const McConst unused{};
  const int grid = B.psdf ? surfel_grid<mc_surfel_kernel<true, Many>>(cus, shape)
                          : surfel_grid<mc_surfel_kernel<false, Many>>(cus, shape);
  const uint32_t q = (uint32_t)shape.static_quarters;
  if (!marks_stamped) mc_mark_kernel<<<cus, 256, 0, s>>>(B.src(), B.frames, unused);
  mc_neighbours_kernel<<<cus * 8, 256, 0, s>>>(B.src(), B.frames, unused, nbt);
  if (B.psdf)
    mc_surfel_kernel<true><<<grid, kBlockVoxels, 0, s>>>(B.src(), B.frames, unused, nbt, q);
  else
    mc_surfel_kernel<false><<<grid, kBlockVoxels, 0, s>>>(B.src(), B.frames, unused, nbt, q);
  mc_push_freed_kernel<<<dim3(kPushGrid, B.nscenes), 256, 0, s>>>(B.src(), B.frames);
}

}  // namespace mh
