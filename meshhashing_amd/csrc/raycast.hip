// raycast.hip -- fixed-step ray march over the hashed TSDF with regula-falsi
// refinement.
//
// Reference: visualization/ray_caster.cu:16-150 (CastKernel),
// geometry/isosurface_intersection.h:13-55 (LinearIntersection,
// BisectionIntersection), geometry/spatial_query.h:23-97 (trilinear value and
// gradient), geometry/voxel_query.h:102-129 (world-point voxel lookup).
//
// One lane per pixel, 8x8 pixels per wave.  Samples, the crossing test and
// the three regula-falsi steps follow the reference's FP32 expressions.  The
// march keeps the eight blocks around its current lowest-corner block
// (Octant) and shares block lookups, misses included, through a per-wave LDS
// table, so runs of samples in one block, or in one unallocated block, cost
// no hash probe; the adjoint and localisation kernels keep a lane's last two
// lookups in registers (BlockCache) instead.
// The depth output is metric z (t * dir_c.z); the reference colour-maps it
// (ray_caster.cu:96-98) and keeps metric depth only in the vertex image.
#include "mh_kernels.h"

namespace mh {

// Block lookups shared by the 64 rays of a workgroup (one wave): a
// direct-mapped LDS table that a lane fills after a hash probe, so a block
// is probed about once per wave rather than once per ray.  An entry is one
// 64-bit word, read and written by single LDS instructions, so a reader
// always gets a block and its own index whatever order the wave's divergent
// paths run in: bits 0-23 hold the pool index + 3 (0 marks an empty slot),
// bits 24-63 the block coordinates minus the low bits that chose the slot
// (x, y: 3 bits, z: 2 bits), which makes (slot, entry) name one block.
// Blocks beyond +-2^15 on an axis bypass the table.  Of the lanes that miss
// on one slot in the same instruction, only the one whose id survives in
// `owner` writes it.
constexpr int kWaveTable = 256;
struct WaveTable {
  uint64_t e[kWaveTable];
  uint32_t owner[kWaveTable];
};
__device__ __forceinline__ void wave_table_init(WaveTable* t) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)kWaveTable; i += 64u) t->e[i] = 0ull;
  __syncthreads();
}

__device__ __forceinline__ int32_t table_find(const View& V, WaveTable* t, i3 b) {
  const int lim = 1 << 15;
  if (b.x < -lim || b.x >= lim || b.y < -lim || b.y >= lim || b.z < -lim || b.z >= lim)
    return coord_ok(b) ? hash_find_eager(V, b) : kFreePtr;
  const uint32_t slot = (uint32_t)(b.x & 7) | ((uint32_t)(b.y & 7) << 3) | ((uint32_t)(b.z & 3) << 6);
  const uint64_t tag = (uint64_t)((uint32_t)(b.x >> 3) & 0x1FFFu) |
                       ((uint64_t)((uint32_t)(b.y >> 3) & 0x1FFFu) << 13) |
                       ((uint64_t)((uint32_t)(b.z >> 2) & 0x3FFFu) << 26);
  const uint64_t e = t->e[slot];
  if (e != 0ull && (e >> 24) == tag) return (int32_t)(e & 0xFFFFFFu) - 3;
  const int32_t p = coord_ok(b) ? hash_find_eager(V, b) : kFreePtr;
  if (V.ray_mark && p >= 0) V.ray_mark[p] = V.ray_stamp;  // diagnostics: block sampled
  if (p >= (1 << 24) - 3) return p;  // does not fit the entry
  const uint32_t lane = threadIdx.x & 63u;
  volatile uint32_t* owner = t->owner;
  owner[slot] = lane;
  if (owner[slot] == lane) t->e[slot] = (tag << 24) | (uint64_t)(uint32_t)(p + 3);
  return p;
}

// The last two lookups of one ray (hits and misses), in registers.
struct BlockCache {
  uint64_t k0, k1;
  int32_t p0, p1;
  WaveTable* t;
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
  const int32_t p = table_find(V, c.t, b);
  c.k1 = c.k0;
  c.p1 = c.p0;
  c.k0 = k;
  c.p0 = p;
  return p;
}


// Pool indices of the blocks holding a trilinear cell's eight corners
// (corner i = (cx[i>>2 & 1], cy[i>>1 & 1], cz[i & 1])).  Each distinct
// block is looked up once, lowest corner first, from a single call site (a
// rolled loop): the probe is emitted once per caller instead of once per
// corner.  kStop: end at the first missing block (false); otherwise every
// block is looked up and missing ones are left negative.
template <bool kStop>
__device__ __forceinline__ bool corner_blocks(const View& V, BlockCache& cache, const int cx[2],
                                              const int cy[2], const int cz[2], int32_t pb[8]) {
  const int bx0 = voxel_to_block1(cx[0]), by0 = voxel_to_block1(cy[0]), bz0 = voxel_to_block1(cz[0]);
  const bool xx = voxel_to_block1(cx[1]) != bx0, yy = voxel_to_block1(cy[1]) != by0,
             zz = voxel_to_block1(cz[1]) != bz0;
  // distinct blocks: entry e = (dx << 2 | dy << 1 | dz) for block b0 + d
  uint32_t todo = 1;
#pragma unroll
  for (int e = 1; e < 8; ++e)
    if ((!(e & 4) || xx) && (!(e & 2) || yy) && (!(e & 1) || zz)) todo |= 1u << e;
  int32_t be[8];
  bool ok = true;
  while (todo) {
    const int e = __ffs(todo) - 1;
    todo &= todo - 1u;
    const int32_t q = cached_find(V, cache, i3{bx0 + ((e >> 2) & 1), by0 + ((e >> 1) & 1), bz0 + (e & 1)});
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j == e) be[j] = q;
    if (q < 0) {
      ok = false;
      if (kStop) return false;
    }
  }
  // corner i lies in entry (i & 4 if xx) | (i & 2 if yy) | (i & 1 if zz),
  // chosen by selects (constant indices keep be[] in registers)
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool sx = (i & 4) && xx, sy = (i & 2) && yy, sz = (i & 1) && zz;
    pb[i] = sx ? (sy ? (sz ? be[7] : be[6]) : (sz ? be[5] : be[4]))
               : (sy ? (sz ? be[3] : be[2]) : (sz ? be[1] : be[0]));
  }
  return ok;
}

// GetSpatialValue restricted to the SDF (the only field the march uses),
// without the empty-space test.
__device__ __forceinline__ bool ray_sdf_full(const View& V, BlockCache& cache, f3 pos, float* out) {
  const float vs = V.geo.voxel_size;
  const float hv = 0.5f * vs;
  const float pcx = pos.x - hv, pcy = pos.y - hv, pcz = pos.z - hv;
  const float qx = pos.x / vs, qy = pos.y / vs, qz = pos.z / vs;
  const float rx = qx - floorf(qx), ry = qy - floorf(qy), rz = qz - floorf(qz);
  const int cx[2] = {vox1(pcx + 0.0f, vs), vox1(pcx + vs, vs)};
  const int cy[2] = {vox1(pcy + 0.0f, vs), vox1(pcy + vs, vs)};
  const int cz[2] = {vox1(pcz + 0.0f, vs), vox1(pcz + vs, vs)};
  // The sample is invalid if any corner's block is missing; in empty space
  // the lowest corner's block already says so (one cached probe).
  int32_t pb[8];
  if (!corner_blocks<true>(V, cache, cx, cy, cz, pb)) return false;
  float sdf = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const uint32_t li = (uint32_t)((cx[mx] & 7) + 8 * (cy[my] & 7) + 64 * (cz[mz] & 7));
    const float w = (mx ? rx : 1.0f - rx) * (my ? ry : 1.0f - ry) * (mz ? rz : 1.0f - rz);
    sdf += w * V.sdf[(size_t)pb[i] * kBlockVoxels + li];
  }
  *out = sdf;
  return true;
}

// GetSpatialValue's SDF with the empty-space test first.
__device__ __forceinline__ bool ray_sdf(const View& V, BlockCache& cache, f3 pos, float inv_vs,
                                        float* out) {
  const float hv = 0.5f * V.geo.voxel_size;
  int c0x, c0y, c0z;
  if (fast_corner(pos.x - hv, inv_vs, &c0x) && fast_corner(pos.y - hv, inv_vs, &c0y) &&
      fast_corner(pos.z - hv, inv_vs, &c0z) &&
      cached_find(V, cache, i3{voxel_to_block1(c0x), voxel_to_block1(c0y), voxel_to_block1(c0z)}) < 0)
    return false;
  return ray_sdf_full(V, cache, pos, out);
}

// The eight blocks around a lowest-corner block b: b + (i>>2 & 1, i>>1 & 1,
// i & 1) for entry i, each looked up when a sample first needs it and kept
// while the march stays with b.
constexpr int32_t kNotLooked = INT32_MIN;
struct Octant {
  int bx, by, bz;
  int32_t p[8];
};
__device__ __forceinline__ Octant empty_octant() {
  return Octant{INT32_MIN, INT32_MIN, INT32_MIN,
                {kNotLooked, kNotLooked, kNotLooked, kNotLooked, kNotLooked, kNotLooked, kNotLooked, kNotLooked}};
}

// GetSpatialValue's SDF for a march sample (spatial_query.h:23-66).  The
// corner voxels come from a multiply by 1/vs when the quotient is more than
// 1e-3 voxel from a rounding tie (the same integers as the reference's
// divisions, see fast_corner; the upper corner is then the lower + 1, the
// sum pc + vs being within 1.3e-4 voxel of exact), else from the divisions;
// the ratios come from the division, as the reference computes them.  A
// missing lowest-corner block ends the sample after one compare while the
// octant is cached, which is how empty space inside the tile range costs
// little.  Octant misses go to the wave's LDS table directly (the cache's
// two-entry register cache is left out of the march: its registers cost
// more than the lookups it saves).
// The sample from its corner voxels; ratios() gives (rx, ry, rz), taken
// only once the sample is known to be valid.
template <class Ratios>
__device__ __forceinline__ bool sample_corners(const View& V, BlockCache& cache, Octant& oc, int cx0, int cx1,
                                               int cy0, int cy1, int cz0, int cz1, Ratios ratios,
                                               float* out) {
  const int bx = voxel_to_block1(cx0), by = voxel_to_block1(cy0), bz = voxel_to_block1(cz0);
  if (bx != oc.bx || by != oc.by || bz != oc.bz) {
    oc.bx = bx;
    oc.by = by;
    oc.bz = bz;
    oc.p[0] = table_find(V, cache.t, i3{bx, by, bz});
#pragma unroll
    for (int i = 1; i < 8; ++i) oc.p[i] = kNotLooked;
  }
  if (oc.p[0] < 0) return false;
  const bool xx = voxel_to_block1(cx1) != bx, yy = voxel_to_block1(cy1) != by, zz = voxel_to_block1(cz1) != bz;
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    const bool need = (!(i & 4) || xx) && (!(i & 2) || yy) && (!(i & 1) || zz);
    if (!need) continue;
    if (oc.p[i] == kNotLooked) oc.p[i] = table_find(V, cache.t, i3{bx + ((i >> 2) & 1), by + ((i >> 1) & 1), bz + (i & 1)});
    if (oc.p[i] < 0) return false;
  }
  // block of each corner (bit 2: x, bit 1: y, bit 0: z set), by selects
  const int32_t p0 = oc.p[0];
  const int32_t p1 = zz ? oc.p[1] : p0;
  const int32_t p2 = yy ? oc.p[2] : p0;
  const int32_t p3 = yy ? (zz ? oc.p[3] : oc.p[2]) : p1;
  const int32_t p4 = xx ? oc.p[4] : p0;
  const int32_t p5 = xx ? (zz ? oc.p[5] : oc.p[4]) : p1;
  const int32_t p6 = xx ? (yy ? oc.p[6] : oc.p[4]) : p2;
  const int32_t p7 = xx ? (yy ? (zz ? oc.p[7] : oc.p[6]) : (zz ? oc.p[5] : oc.p[4])) : p3;
  const int32_t pb[8] = {p0, p1, p2, p3, p4, p5, p6, p7};
  const int cx[2] = {cx0, cx1}, cy[2] = {cy0, cy1}, cz[2] = {cz0, cz1};
  const f3 r = ratios();
  const float rx = r.x, ry = r.y, rz = r.z;
  float sdf = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const uint32_t li = (uint32_t)((cx[mx] & 7) + 8 * (cy[my] & 7) + 64 * (cz[mz] & 7));
    const float w = (mx ? rx : 1.0f - rx) * (my ? ry : 1.0f - ry) * (mz ? rz : 1.0f - rz);
    sdf += w * V.sdf[(size_t)pb[i] * kBlockVoxels + li];
  }
  *out = sdf;
  return true;
}

__device__ __forceinline__ bool march_sample(const View& V, BlockCache& cache, Octant& oc, f3 pos,
                                             float inv_vs, float* out) {
  const float vs = V.geo.voxel_size;
  const float hv = 0.5f * vs;
  const float pcx = pos.x - hv, pcy = pos.y - hv, pcz = pos.z - hv;
  // per axis, so that a wave runs the divisions only for the axes on which
  // some lane is near a tie
  int cx0, cy0, cz0, cx1, cy1, cz1;
  corner_pair(pcx, vs, inv_vs, &cx0, &cx1);
  corner_pair(pcy, vs, inv_vs, &cy0, &cy1);
  corner_pair(pcz, vs, inv_vs, &cz0, &cz1);
  return sample_corners(V, cache, oc, cx0, cx1, cy0, cy1, cz0, cz1, [&] {
    const float qx = pos.x / vs, qy = pos.y / vs, qz = pos.z / vs;
    return mk3(qx - floorf(qx), qy - floorf(qy), qz - floorf(qz));
  }, out);
}

// One axis of a sample at coordinate p: its two corner voxels (as
// march_sample rounds them) and the ratio frac(p / vs).
struct RAxis {
  int c0, c1;
  float r;
};
__device__ __forceinline__ RAxis ray_axis(float p, float vs, float inv_vs) {
  RAxis a;
  corner_pair(p - 0.5f * vs, vs, inv_vs, &a.c0, &a.c1);
  const float q = p / vs;
  a.r = q - floorf(q);
  return a;
}

// march_sample from the sample's three axes (the normal's stencils share two
// of them with the crossing point).
__device__ __forceinline__ bool axes_sample(const View& V, BlockCache& cache, Octant& oc, const RAxis& X,
                                            const RAxis& Y, const RAxis& Z, float* out) {
  return sample_corners(V, cache, oc, X.c0, X.c1, Y.c0, Y.c1, Z.c0, Z.c1, [&] { return mk3(X.r, Y.r, Z.r); },
                        out);
}

// A ray parameter below which every sample position o + t dir lies more
// than 0.01 voxel inside the world box [b bs, (b + 1) bs)^3 of cell b, bs
// its side (8 voxels for a block, 32 for a super-block).  A sample's lowest
// corner voxel is vox1(pos - vs/2) per axis, so its lowest-corner block lies
// in cell b exactly on that box; the margin is 20 times the
// FP32 error of the reference's quotient (pos - vs/2) / vs at 16 m and
// covers the error of the position.  The box is shrunk by 0.02 voxel, so
// the approximate divisions (a few ulp of t, well under 1e-5 m) only
// lower the bound.
__device__ __forceinline__ float cell_exit_t(f3 o, f3 dir, int bx, int by, int bz, float bs, float vs) {
  const float eps = 0.02f * vs;
  const float ex = dir.x > 0.0f ? (float)(bx + 1) * bs - eps : (float)bx * bs + eps;
  const float ey = dir.y > 0.0f ? (float)(by + 1) * bs - eps : (float)by * bs + eps;
  const float ez = dir.z > 0.0f ? (float)(bz + 1) * bs - eps : (float)bz * bs + eps;
  float te = INFINITY;
  if (dir.x != 0.0f) te = fminf(te, __fdividef(ex - o.x, dir.x));
  if (dir.y != 0.0f) te = fminf(te, __fdividef(ey - o.y, dir.y));
  if (dir.z != 0.0f) te = fminf(te, __fdividef(ez - o.z, dir.z));
  return te;
}
__device__ __forceinline__ float block_exit_t(f3 o, f3 dir, int bx, int by, int bz, float vs) {
  return cell_exit_t(o, dir, bx, by, bz, (float)kBlockSide * vs, vs);
}
// The same for the super-block (4^3 blocks) holding block b, when no block of
// it was ever allocated (its filter bit is clear): every sample whose
// lowest-corner block lies in it is invalid.  Otherwise block_exit_t.
__device__ __forceinline__ float absent_exit_t(const View& V, f3 o, f3 dir, int bx, int by, int bz, float vs) {
  const uint32_t sb = occ2_bit(super_key(i3{bx, by, bz}));
  if ((V.occ2[sb >> 5] >> (sb & 31u)) & 1u) return block_exit_t(o, dir, bx, by, bz, vs);
  return cell_exit_t(o, dir, bx >> kSuperShift, by >> kSuperShift, bz >> kSuperShift,
                     (float)(kBlockSide << kSuperShift) * vs, vs);
}

// The first element >= target of the FP32 sequence t, t + step, ...
// (t += step, t > 0, step > 0), without taking every step.  Inside a binade
// [lo, 2 lo) every sum rounds to the grid u = lo 2^-23, so once a step has
// been taken inside the binade the increment is constant (an exact tie
// rounds to an even multiple of u, after which the tie always resolves the
// same way): with t1, t2, t3 in one binade, d = t3 - t2 repeats for as long
// as each exact sum t + step stays 2u below the binade's top.  The jump is
// computed in FP64, where t3 + k d is exact.
__device__ __forceinline__ float advance_steps(float t, float step, float target) {
  while (t < target) {
    const float t1 = t + step;
    if (!(t1 < target)) return t1;
    const float t2 = t1 + step;
    if (!(t2 < target)) return t2;
    const float t3 = t2 + step;
    if (!(t3 < target)) return t3;
    const float lo = __uint_as_float(__float_as_uint(t3) & 0x7F800000u);
    if (t1 < lo) {
      t = t3;
      continue;
    }
    const double top = 2.0 * (double)lo, u = (double)lo * 0x1p-23;
    const double d = (double)t3 - (double)t2;
    const double room = (top - 2.0 * u - (double)step - (double)t3) / d;
    if (!(room >= 1.0)) {
      t = t3;
      continue;
    }
    const double k = fmin(floor(room) + 1.0, ceil(((double)target - (double)t3) / d));
    t = (float)((double)t3 + k * d);
  }
  return t;
}

// The accepted crossing of one ray
// and (t_k, s_k), and the last regula-falsi point m.
struct Hit {
  float tp, sp, tk, sk, m;
};

// The march of CastKernel (ray_caster.cu:56-92): fixed FP32 steps, a
// crossing between the previous valid sample and the current one, three
// regula-falsi steps (isosurface_intersection.h:24-55), then the acceptance
// test.  A failed refinement or test continues the march.
//
// [t_lo, t_hi] bounds the ray parameters at which a sample can be valid (the
// tile's block range): earlier samples are stepped over without sampling,
// so the sample sequence t_min, t_min + step, ... stays the reference's, and
// the march ends after t_hi.
__device__ __forceinline__ bool find_surface(const View& V, const RayConst& R, f3 o, f3 dir, float t_min,
                                             float t_max, float t_lo, float t_hi, BlockCache& cache,
                                             Octant& oc, Hit* hit, uint32_t* nsamp) {
  float prev_sdf = 0.0f, prev_t = 0.0f;
  bool prev_ok = false;
  if (!(t_lo < t_max) || t_hi < t_lo) return false;  // no block along this ray
  float t = advance_steps(t_min, R.step, t_lo);
  // The march stops at each crossing candidate and refines it outside the
  // stepping loop, so a wave runs the refinement once for all its lanes
  // rather than once per step at which some lane crosses.
  while (true) {
    float s = 0.0f;
    bool cross = false;
    for (; t < t_max && t <= t_hi; t += R.step) {
      float sv;
      if (!march_sample(V, cache, oc, mk3(o.x + t * dir.x, o.y + t * dir.y, o.z + t * dir.z), R.inv_vs,
                        &sv)) {
        // Lowest-corner block absent: the following samples inside its box
        // (its super-block's, if that never held a block) are invalid too
        // (an invalid sample changes nothing); step over them with the same
        // FP32 additions, so the t sequence is kept.
        if (oc.p[0] < 0) {
          const float t_out = absent_exit_t(V, o, dir, oc.bx, oc.by, oc.bz, V.geo.voxel_size);
          for (float tn = t + R.step; tn < t_out && tn > t; tn = t + R.step) t = tn;
        }
        continue;
      }
      ++*nsamp;
      if (prev_ok && prev_sdf > 0.0f && sv < 0.0f) {
        s = sv;
        cross = true;
        break;
      }
      prev_sdf = sv;
      prev_t = t;
      prev_ok = true;
    }
    if (!cross) return false;
    // BisectionIntersection: three regula-falsi steps.
    float l = prev_t, r = t, m = (l + r) * 0.5f;
    float l_sdf = prev_sdf, r_sdf = s;
    bool found = true;
    for (int k = 0; k < 3; ++k) {
      m = l + (l_sdf / (l_sdf - r_sdf)) * (r - l);
      float m_sdf;
      if (!march_sample(V, cache, oc, mk3(o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z), R.inv_vs,
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
      *hit = Hit{prev_t, prev_sdf, t, s, m};
      return true;
    }
    // not accepted: the march goes on, with this sample as the previous one
    prev_sdf = s;
    prev_t = t;
    prev_ok = true;
    t += R.step;
  }
}

// The ray parameters between which pixel (x, y) can meet an allocated block:
// the tile's camera-z range over dir_c.z, widened by 0.1 %.  Without a range
// pass, the whole [t_min, t_max].
__device__ __forceinline__ void ray_bounds(const RayConst& R, uint32_t x, uint32_t y, float dcz, float t_min,
                                           float t_max, float* t_lo, float* t_hi) {
  *t_lo = t_min;
  *t_hi = t_max;
  if (!R.range || R.range[2 * R.tiles]) return;
  const uint32_t tile = (y >> 3) * R.tiles_x + (x >> 3);
  const float zmin = __uint_as_float(R.range[2 * tile]);
  const float zmax = __uint_as_float(R.range[2 * tile + 1]);
  *t_lo = fmaxf(t_min, 0.999f * (zmin / dcz));
  *t_hi = fminf(t_max, 1.001f * (zmax / dcz));
}

// March one pixel; returns the number of valid samples taken.
__device__ __forceinline__ uint32_t cast_pixel(const View& V, const RayConst& R, uint32_t x, uint32_t y,
                               WaveTable* table, float* res_d, f3* res_v, f3* res_n) {
  uint32_t nsamp = 0;
  float out_d = -INFINITY;
  f3 out_v = mk3(-INFINITY, -INFINITY, -INFINITY);
  f3 out_n = mk3(-INFINITY, -INFINITY, -INFINITY);

  const f3 dir_c = normalize(reproject(R.cam, x, y, 1.0f));
  const float t_min = R.min_depth / dir_c.z;
  const float t_max = R.max_depth / dir_c.z;
  const f3 o = xform(R.wTc, mk3(0.0f, 0.0f, 0.0f));
  const f3 dir = normalize(xform_dir(R.wTc, dir_c));

  BlockCache cache{kEmptyKey, kEmptyKey, kFreePtr, kFreePtr, table};
  Hit hit;
  float t_lo, t_hi;
  ray_bounds(R, x, y, dir_c.z, t_min, t_max, &t_lo, &t_hi);
  Octant oc = empty_octant();
  if (find_surface(V, R, o, dir, t_min, t_max, t_lo, t_hi, cache, oc, &hit, &nsamp)) {
    const float m = hit.m;
    const float d = m * dir_c.z;
    out_d = d;
    out_v = reproject(R.cam, x, y, d);
    if (R.enable_gradients) {
      const f3 iso = mk3(o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z);
      const float h = 0.5f * V.geo.voxel_size;
      bool valid = true;
      float sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
      const float vs = V.geo.voxel_size;
      // iso -+ h e_k equals iso off axis k (x - 0 = x), so the stencils take
      // the crossing point's axes there
      const RAxis X = ray_axis(iso.x, vs, R.inv_vs), Y = ray_axis(iso.y, vs, R.inv_vs),
                  Z = ray_axis(iso.z, vs, R.inv_vs);
#pragma unroll
      for (int k = 0; k < 3 && valid; ++k) {
        const float ik = k == 0 ? iso.x : (k == 1 ? iso.y : iso.z);
        const RAxis An = ray_axis(ik - h, vs, R.inv_vs);
        valid = axes_sample(V, cache, oc, k == 0 ? An : X, k == 1 ? An : Y, k == 2 ? An : Z, &sn[k]);
        if (!valid) break;
        const RAxis Ap = ray_axis(ik + h, vs, R.inv_vs);
        valid = axes_sample(V, cache, oc, k == 0 ? Ap : X, k == 1 ? Ap : Y, k == 2 ? Ap : Z, &sp[k]);
      }
      const f3 grad = mk3((sp[0] - sn[0]) / vs, (sp[1] - sn[1]) / vs, (sp[2] - sn[2]) / vs);
      const float l2 = length(grad);
      f3 nw = (l2 > 0.0f && valid) ? grad / l2 : mk3(0.0f, 0.0f, 0.0f);
      nw = -nw;
      out_n = xform_dir(R.cTw, nw);
    }
  }
  *res_d = out_d;
  *res_v = out_v;
  *res_n = out_n;
  return nsamp;
}

// ------------------------------------------------------------ adjoint
// GetSpatialValue's SDF (same corners, weights and summation order as
// ray_sdf, so the value is bit-identical) with the spatial gradient of the
// trilinear interpolant: each ratio frac(pos / vs) moves at 1 / vs.
__device__ __forceinline__ bool ray_sdf_grad(const View& V, BlockCache& cache, f3 pos, float* val,
                                             f3* grad) {
  const float vs = V.geo.voxel_size;
  const float hv = 0.5f * vs;
  const float pcx = pos.x - hv, pcy = pos.y - hv, pcz = pos.z - hv;
  const float qx = pos.x / vs, qy = pos.y / vs, qz = pos.z / vs;
  const float rx = qx - floorf(qx), ry = qy - floorf(qy), rz = qz - floorf(qz);
  const int cx[2] = {vox1(pcx + 0.0f, vs), vox1(pcx + vs, vs)};
  const int cy[2] = {vox1(pcy + 0.0f, vs), vox1(pcy + vs, vs)};
  const int cz[2] = {vox1(pcz + 0.0f, vs), vox1(pcz + vs, vs)};
  int32_t pb[8];
  if (!corner_blocks<true>(V, cache, cx, cy, cz, pb)) return false;
  float sdf = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    const uint32_t li = (uint32_t)((cx[mx] & 7) + 8 * (cy[my] & 7) + 64 * (cz[mz] & 7));
    const float c = V.sdf[(size_t)pb[i] * kBlockVoxels + li];
    const float wx = mx ? rx : 1.0f - rx, wy = my ? ry : 1.0f - ry, wz = mz ? rz : 1.0f - rz;
    sdf += wx * wy * wz * c;
    gx += (mx ? c : -c) * wy * wz;
    gy += (my ? c : -c) * wx * wz;
    gz += (mz ? c : -c) * wx * wy;
  }
  *val = sdf;
  *grad = mk3(gx / vs, gy / vs, gz / vs);
  return true;
}

// Adds coef * (trilinear weight of each corner) to the corners' gradient.
__device__ __forceinline__ void scatter_sdf(const View& V, BlockCache& cache, f3 pos, float coef,
                                            float* __restrict__ grad_sdf) {
  const float vs = V.geo.voxel_size;
  const float hv = 0.5f * vs;
  const float pcx = pos.x - hv, pcy = pos.y - hv, pcz = pos.z - hv;
  const float qx = pos.x / vs, qy = pos.y / vs, qz = pos.z / vs;
  const float rx = qx - floorf(qx), ry = qy - floorf(qy), rz = qz - floorf(qz);
  const int cx[2] = {vox1(pcx + 0.0f, vs), vox1(pcx + vs, vs)};
  const int cy[2] = {vox1(pcy + 0.0f, vs), vox1(pcy + vs, vs)};
  const int cz[2] = {vox1(pcz + 0.0f, vs), vox1(pcz + vs, vs)};
  int32_t pb[8];
  corner_blocks<false>(V, cache, cx, cy, cz, pb);
  for (int i = 0; i < 8; ++i) {
    const int mx = (i >> 2) & 1, my = (i >> 1) & 1, mz = i & 1;
    if (pb[i] < 0) continue;  // cannot happen for a sample the forward pass accepted
    const uint32_t li = (uint32_t)((cx[mx] & 7) + 8 * (cy[my] & 7) + 64 * (cz[mz] & 7));
    const float w = (mx ? rx : 1.0f - rx) * (my ? ry : 1.0f - ry) * (mz ? rz : 1.0f - rz);
    atomicAdd(&grad_sdf[(size_t)pb[i] * kBlockVoxels + li], coef * w);
  }
}

// Tangents: d/d(omega_xyz, upsilon_xyz) of a pose perturbation
// wTc <- exp(xi^) wTc (a world point x moves by omega x x + upsilon), then
// d/d(eps_0..eps_4), a perturbation of the value of each of the five SDF
// evaluations the depth depends on (the two bracketing samples and the
// three regula-falsi points).
constexpr int kTan = 11;

__device__ __forceinline__ void sample_tangent(f3 x, f3 g, int slot, float* t) {
  const f3 c = mk3(x.y * g.z - x.z * g.y, x.z * g.x - x.x * g.z, x.x * g.y - x.y * g.x);  // x cross g
  t[0] = c.x; t[1] = c.y; t[2] = c.z;
  t[3] = g.x; t[4] = g.y; t[5] = g.z;
#pragma unroll
  for (int k = 6; k < kTan; ++k) t[k] = k == slot ? 1.0f : 0.0f;
}

// Differentiates the rendered depth of one pixel: recomputes the march
// (same decisions as the forward pass), then carries tangents through the
// three regula-falsi steps.  Returns false when the pixel has no hit.
__device__ __forceinline__ bool depth_tangent(const View& V, const RayConst& R, uint32_t x, uint32_t y,
                                              BlockCache& cache, float* depth, float* dd, f3* X) {
  const f3 dir_c = normalize(reproject(R.cam, x, y, 1.0f));
  const float t_min = R.min_depth / dir_c.z;
  const float t_max = R.max_depth / dir_c.z;
  const f3 o = xform(R.wTc, mk3(0.0f, 0.0f, 0.0f));
  const f3 dir = normalize(xform_dir(R.wTc, dir_c));
  Hit hit;
  uint32_t ns = 0;
  float t_lo, t_hi;
  ray_bounds(R, x, y, dir_c.z, t_min, t_max, &t_lo, &t_hi);
  Octant oc = empty_octant();
  if (!find_surface(V, R, o, dir, t_min, t_max, t_lo, t_hi, cache, oc, &hit, &ns)) return false;
  float l = hit.tp, r = hit.tk, ls, rs;
  float dl[kTan], dr[kTan], dls[kTan], drs[kTan], dm[kTan];
  f3 g;
  X[0] = mk3(o.x + l * dir.x, o.y + l * dir.y, o.z + l * dir.z);
  X[1] = mk3(o.x + r * dir.x, o.y + r * dir.y, o.z + r * dir.z);
  ray_sdf_grad(V, cache, X[0], &ls, &g);
  sample_tangent(X[0], g, 6, dls);
  ray_sdf_grad(V, cache, X[1], &rs, &g);
  sample_tangent(X[1], g, 7, drs);
#pragma unroll
  for (int k = 0; k < kTan; ++k) dl[k] = dr[k] = 0.0f;
  float m = 0.0f;
#pragma unroll 1
  for (int j = 0; j < 3; ++j) {
    const float D = ls - rs;
    const float q = ls / D;
    m = l + (ls / (ls - rs)) * (r - l);
    const float inv_d2 = 1.0f / (D * D);
#pragma unroll
    for (int k = 0; k < kTan; ++k) {
      const float dq = (ls * drs[k] - rs * dls[k]) * inv_d2;
      dm[k] = dl[k] + dq * (r - l) + q * (dr[k] - dl[k]);
    }
    const f3 xm = mk3(o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z);
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (j == q) X[2 + q] = xm;
    float ms;
    ray_sdf_grad(V, cache, xm, &ms, &g);
    float dms[kTan];
    sample_tangent(xm, g, 8 + j, dms);
    const float gd = g.x * dir.x + g.y * dir.y + g.z * dir.z;
#pragma unroll
    for (int k = 0; k < kTan; ++k) dms[k] += gd * dm[k];
    if (ls * ms > 0.0f) {
      l = m;
      ls = ms;
#pragma unroll
      for (int k = 0; k < kTan; ++k) { dl[k] = dm[k]; dls[k] = dms[k]; }
    } else {
      r = m;
      rs = ms;
#pragma unroll
      for (int k = 0; k < kTan; ++k) { dr[k] = dm[k]; drs[k] = dms[k]; }
    }
  }
  *depth = m * dir_c.z;
#pragma unroll
  for (int k = 0; k < kTan; ++k) dd[k] = dm[k] * dir_c.z;
  return true;
}

// Per pixel: residual r = depth - observed depth (pixels with a hit, a
// finite positive observation and |r| <= max_residual; the others with a
// hit and an observation are counted as gated out), J = d depth / d xi.
// Every pixel with a finite positive observation is counted as observed,
// hit or not, so a caller can charge the pixels a pose loses.
// Pixels whose depth moves faster than max_slope times a camera translation
// (|d depth / d upsilon| on some axis; rays close to grazing the surface)
// are left out altogether.  Per workgroup (one
// wave): FP64 sums of r^2 / 2, r J, J J^T (upper triangle) and the pixel
// count, reduced in a fixed order by raycast_grad_reduce_kernel.  With
// grad_sdf, r * d depth / d sdf is scattered onto the voxels.
__global__ __launch_bounds__(64) void raycast_backward_kernel(View V, RayConst R,
                                                              const float* __restrict__ obs,
                                                              float* __restrict__ grad_sdf,
                                                              double* __restrict__ partials,
                                                              float max_residual, float max_slope) {
  __shared__ WaveTable table;
  wave_table_init(&table);
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  double acc[kGradSlots];
#pragma unroll
  for (int k = 0; k < kGradSlots; ++k) acc[k] = 0.0;
  if (x < R.cam.width && y < R.cam.height) {
    const float od = obs[y * R.cam.width + x];
    BlockCache cache{kEmptyKey, kEmptyKey, kFreePtr, kFreePtr, &table};
    float d, dd[kTan];
    f3 X[5];
    acc[30] = od > 0.0f && od < INFINITY ? 1.0 : 0.0;
    if (od > 0.0f && od < INFINITY && depth_tangent(V, R, x, y, cache, &d, dd, X) &&
        fmaxf(fabsf(dd[3]), fmaxf(fabsf(dd[4]), fabsf(dd[5]))) <= max_slope) {
      const float res = d - od;
      if (!(fabsf(res) <= max_residual)) {
        acc[29] = 1.0;
      } else {
        acc[0] = 0.5 * (double)res * (double)res;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[1 + k] = (double)res * (double)dd[k];
        int t = 7;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
          for (int j = i; j < 6; ++j) acc[t++] = (double)dd[i] * (double)dd[j];
        acc[28] = 1.0;
        if (grad_sdf)
          for (int e = 0; e < 5; ++e) scatter_sdf(V, cache, X[e], res * dd[6 + e], grad_sdf);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kGradSlots; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    acc[k] = v;
  }
  if (threadIdx.x == 0) {
    const size_t wg = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int k = 0; k < kGradSlots; ++k) partials[wg * kGradSlots + k] = acc[k];
  }
}

// Sums the per-workgroup partials in workgroup order (deterministic).
// Sums n rows of `slots` FP64 partials column by column, in row order.
__global__ void slot_reduce_kernel(const double* __restrict__ partials, uint32_t n, uint32_t slots,
                                   double* __restrict__ out) {
  const uint32_t k = threadIdx.x;
  if (k >= slots) return;
  double s = 0.0;
  for (uint32_t i = 0; i < n; ++i) s += partials[(size_t)i * slots + k];
  out[k] = s;
}

// ------------------------------------------------------------ localisation
// PointToSurfaceKernel (localizing/point_to_psdf.cu:19-82): one lane per odd
// pixel (both coordinates odd), 8 x 8 of them per 64-lane workgroup, i.e.
// a 16 x 16 pixel tile as in the reference.  A depth that is -inf, 0 or at
// least sdf_upper_bound is skipped; the world point wTc * p, the SDF
// (GetSpatialValue) and its central difference at +-vs/2
// (GetSpatialSDFGradient, spatial_query.h:71-97) are the reference's FP32
// expressions.  J = (g, p x g) for the tangent (translation, rotation);
// `corrected` = 0 keeps the reference's g.x in J's third entry (:69-74).
// The per-pixel products J J^T, d J and d^2 are FP32, as the reference's
// mat6x6 arithmetic; they are summed in FP64 per wave and then per
// workgroup in a fixed order instead of float atomics.  Slots: A's upper
// triangle (21), b (6), err, count.
__global__ __launch_bounds__(64) void point_to_psdf_kernel(View V, M4 wTc, float inv_vs, int corrected,
                                                           double* __restrict__ partials) {
  __shared__ WaveTable table;
  wave_table_init(&table);
  const uint32_t x = 2 * (blockIdx.x * 8 + (threadIdx.x & 7)) + 1;
  const uint32_t y = 2 * (blockIdx.y * 8 + (threadIdx.x >> 3)) + 1;
  double acc[kP2SSlots];
#pragma unroll
  for (int k = 0; k < kP2SSlots; ++k) acc[k] = 0.0;
  if (x < V.cam.width && y < V.cam.height) {
    const float depth = V.depth[(size_t)y * V.cam.width + x];
    if (!(depth == -INFINITY || depth == 0.0f || depth >= V.geo.sdf_upper_bound)) {
      const f3 pw = xform(wTc, reproject(V.cam, x, y, depth));
      BlockCache cache{kEmptyKey, kEmptyKey, kFreePtr, kFreePtr, &table};
      float d;
      if (ray_sdf(V, cache, pw, inv_vs, &d)) {
        const float h = 0.5f * V.geo.voxel_size;
        bool ok = true;
        float sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
        // x-, x+, y-, y+, z-, z+ from one (rolled) sample site
#pragma unroll 1
        for (int j = 0; j < 6 && ok; ++j) {
          const int k = j >> 1;
          const f3 dp = mk3(k == 0 ? h : 0.0f, k == 1 ? h : 0.0f, k == 2 ? h : 0.0f);
          float v;
          ok = ray_sdf(V, cache, (j & 1) ? pw + dp : pw - dp, inv_vs, &v);
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            if (j == 2 * q) sn[q] = v;
            if (j == 2 * q + 1) sp[q] = v;
          }
        }
        if (ok) {
          const float vs = V.geo.voxel_size;
          const f3 g = mk3((sp[0] - sn[0]) / vs, (sp[1] - sn[1]) / vs, (sp[2] - sn[2]) / vs);
          const float J[6] = {g.x, g.y, corrected ? g.z : g.x,
                              -g.y * pw.z + g.z * pw.y,
                              -g.z * pw.x + g.x * pw.z,
                              -g.x * pw.y + g.y * pw.x};
          int t = 0;
#pragma unroll
          for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = i; j < 6; ++j) acc[t++] = (double)(J[i] * J[j]);
#pragma unroll
          for (int i = 0; i < 6; ++i) acc[21 + i] = (double)(d * J[i]);
          acc[27] = (double)(d * d);
          acc[28] = 1.0;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kP2SSlots; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    acc[k] = v;
  }
  if (threadIdx.x == 0) {
    const size_t wg = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int k = 0; k < kP2SSlots; ++k) partials[wg * kP2SSlots + k] = acc[k];
  }
}

uint32_t point_to_psdf_partials(const View& V) {
  return ((V.cam.width / 2 + 7) / 8) * ((V.cam.height / 2 + 7) / 8);
}

void launch_point_to_psdf(const View& V, const M4& wTc, int corrected, double* partials, double* out,
                          hipStream_t s) {
  const dim3 grid((V.cam.width / 2 + 7) / 8, (V.cam.height / 2 + 7) / 8);
  point_to_psdf_kernel<<<grid, 64, 0, s>>>(V, wTc, 1.0f / V.geo.voxel_size, corrected, partials);
  slot_reduce_kernel<<<1, 64, 0, s>>>(partials, grid.x * grid.y, kP2SSlots, out);
}

__device__ __forceinline__ const RayConst& ray_of(const One&, uint32_t, const RayConst& r) { return r; }
__device__ __forceinline__ const RayConst& ray_of(const Many& m, uint32_t s, const RayConst&) {
  return m.f[s].ray;
}

// Scene = blockIdx.z; a batch's outputs are scene-major (W*H per scene).
// At most 96 registers: five waves per SIMD hold a 640x480 frame's 4800
// tiles at once on 256 CUs (four would leave a partial second round).
// TW: tile width in pixels (the tile is TW x 64 / TW; every shape gives the
// same per-pixel results, the autotuner picks the fastest).
template <class Src, int TW = 8>
__global__ __launch_bounds__(64, kRayWavesPerSimd) void raycast_kernel(Src src_, const SceneFrame* __restrict__ tab, RayConst R1, float* __restrict__ depth_out,
                                                     float* __restrict__ vertex_out,
                                                     float* __restrict__ normal_out) {
  const Src src = with_table(src_, tab);
  __shared__ WaveTable table;
  wave_table_init(&table);
  const uint32_t scene = blockIdx.z;
  // copies, not references into the frame table: read once, and not
  // reloaded after the stores of the march (LDS table, outputs)
  const View V = src.view(scene);
  const RayConst R = ray_of(src, scene, R1);
  if (R.boost) __builtin_amdgcn_s_setprio(2);
  constexpr int TH = 64 / TW;
  const uint32_t x = blockIdx.x * TW + (threadIdx.x % TW);
  const uint32_t y = blockIdx.y * TH + (threadIdx.x / TW);
  uint32_t ns = 0;
  if (x < R.cam.width && y < R.cam.height) {
    const size_t pix = (size_t)scene * R.cam.width * R.cam.height + y * R.cam.width + x;
    float d;
    f3 vv, nn;
    ns = cast_pixel(V, R, x, y, &table, &d, &vv, &nn);
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
  wave_add_u64(&work_stripe(V.ctr).ray_samples, ns);
}

// ------------------------------------------------------------ block ranges
// Every valid sample lies in the box of the block holding its lowest
// trilinear corner, [8b, 8b + 8] voxels per axis; with one more voxel on each
// side the box bounds it beyond any rounding.  Each allocated block's box is
// projected (8 corners, all at z >= min_depth / 2, else the range test is
// switched off for the image) and its camera-z extent merged into the 8 x 8
// pixel tiles its projected rectangle, widened by a pixel, overlaps.  A ray
// then meets no valid sample outside its tile's [zmin, zmax].
template <class Src>
__global__ __launch_bounds__(256) void ray_range_clear_kernel(Src src_, const SceneFrame* __restrict__ tab, RayConst R1) {
  const Src src = with_table(src_, tab);
  const RayConst& R = ray_of(src, blockIdx.y, R1);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= 2 * R.tiles; i += gridDim.x * blockDim.x)
    R.range[i] = i == 2 * R.tiles ? 0u : ((i & 1u) ? 0u : 0x7f800000u);  // zmin +inf, zmax 0
}

// kLds: a workgroup first merges its blocks' ranges in an LDS copy of the
// tile table (up to kRangeLdsTiles tiles) and then sends each tile it
// touched to the global table with one atomic per bound, instead of one per
// block and tile.
constexpr uint32_t kRangeLdsTiles = 4800;  // 640 x 480 pixels
template <class Src, bool kLds>
__global__ __launch_bounds__(256) void ray_range_kernel(Src src_, const SceneFrame* __restrict__ tab, RayConst R1) {
  const Src src = with_table(src_, tab);
  const uint32_t scene = blockIdx.y;
  const View& V = src.view(scene);
  const RayConst& R = ray_of(src, scene, R1);
  const uint32_t n = V.ctr->block_hwm;
  const float vs = V.geo.voxel_size;
  const float zcut = 0.5f * R.min_depth;
  __shared__ uint32_t lds_range[kLds ? 2 * kRangeLdsTiles : 1];
  uint32_t* const rng = kLds ? lds_range : R.range;
  if (kLds) {
    if (R.tiles > kRangeLdsTiles) {  // not launched so; the range test is then switched off
      if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&R.range[2 * R.tiles], 1u);
      return;
    }
    for (uint32_t t = threadIdx.x; t < 2 * R.tiles; t += blockDim.x) lds_range[t] = (t & 1u) ? 0u : 0x7f800000u;
    __syncthreads();
  }
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const uint64_t key = V.block_keys[p];
    if (key == kEmptyKey) continue;
    const i3 b = unpack_key(key);
    const float lo[3] = {(float)(8 * b.x - 1) * vs, (float)(8 * b.y - 1) * vs, (float)(8 * b.z - 1) * vs};
    const float hi[3] = {(float)(8 * b.x + 9) * vs, (float)(8 * b.y + 9) * vs, (float)(8 * b.z + 9) * vs};
    float zmin = INFINITY, zmax = -INFINITY, umin = INFINITY, umax = -INFINITY, vmin = INFINITY,
          vmax = -INFINITY;
    bool near = false;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const f3 pc = xform(R.cTw, mk3((c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]));
      near = near || !(pc.z >= zcut);
      const float z = fmaxf(pc.z, zcut);
      const float u = R.cam.fx * pc.x / z + R.cam.cx, v = R.cam.fy * pc.y / z + R.cam.cy;
      zmin = fminf(zmin, pc.z);
      zmax = fmaxf(zmax, pc.z);
      umin = fminf(umin, u);
      umax = fmaxf(umax, u);
      vmin = fminf(vmin, v);
      vmax = fmaxf(vmax, v);
    }
    if (zmax < R.min_depth || zmin > R.max_depth) continue;  // no ray sample reaches it
    if (near) {
      atomicOr(&R.range[2 * R.tiles], 1u);
      continue;
    }
    const float W = (float)R.cam.width, H = (float)R.cam.height;
    if (!(umax >= -1.0f && umin <= W && vmax >= -1.0f && vmin <= H)) continue;  // off the image
    const int x0 = (int)fmaxf(floorf(umin) - 1.0f, 0.0f), x1 = (int)fminf(ceilf(umax) + 1.0f, W - 1.0f);
    const int y0 = (int)fmaxf(floorf(vmin) - 1.0f, 0.0f), y1 = (int)fminf(ceilf(vmax) + 1.0f, H - 1.0f);
    const uint32_t zlo = __float_as_uint(fmaxf(zmin, 0.0f)), zhi = __float_as_uint(zmax);
    for (int ty = y0 >> 3; ty <= (y1 >> 3); ++ty)
      for (int tx = x0 >> 3; tx <= (x1 >> 3); ++tx) {
        const uint32_t t = (uint32_t)ty * R.tiles_x + (uint32_t)tx;
        atomicMin(&rng[2 * t], zlo);
        atomicMax(&rng[2 * t + 1], zhi);
      }
  }
  if (kLds) {
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < R.tiles; t += blockDim.x) {
      const uint32_t lo = lds_range[2 * t], hi = lds_range[2 * t + 1];
      if (lo != 0x7f800000u) atomicMin(&R.range[2 * t], lo);
      if (hi != 0u) atomicMax(&R.range[2 * t + 1], hi);
    }
  }
}

template <class Src>
void launch_ray_range(Src src, const SceneFrame* tab, const RayConst& R, uint32_t tiles, uint32_t nscenes,
                      hipStream_t s) {
  ray_range_clear_kernel<<<dim3(8, nscenes), 256, 0, s>>>(src, tab, R);
  if (tiles <= kRangeLdsTiles) {
    ray_range_kernel<Src, true><<<dim3(nscenes > 1 ? 64 : 128, nscenes), 256, 0, s>>>(src, tab, R);
  } else {
    // one lane per pool slot of a 118 k-block volume over the whole chip
    // (the projections are independent; 128 workgroups had left most CUs idle)
    ray_range_kernel<Src, false><<<dim3(nscenes > 1 ? 256 : 1024, nscenes), 256, 0, s>>>(src, tab, R);
  }
}

template <class Src>
void raycast_shaped(Src src, const SceneFrame* tab, const RayConst& p, uint32_t w, uint32_t h, uint32_t nscenes,
                    int shape, float* depth, float* vertex, float* normal, hipStream_t s) {
  const int tw = kRayTileWidths[shape], th = 64 / tw;
  const dim3 grid((w + tw - 1) / tw, (h + th - 1) / th, nscenes);
  switch (tw) {
    case 16: raycast_kernel<Src, 16><<<grid, 64, 0, s>>>(src, tab, p, depth, vertex, normal); break;
    case 4: raycast_kernel<Src, 4><<<grid, 64, 0, s>>>(src, tab, p, depth, vertex, normal); break;
    case 32: raycast_kernel<Src, 32><<<grid, 64, 0, s>>>(src, tab, p, depth, vertex, normal); break;
    default: raycast_kernel<Src, 8><<<grid, 64, 0, s>>>(src, tab, p, depth, vertex, normal); break;
  }
}

void launch_ray_range_one(const View& V, const RayConst& p, hipStream_t s) {
  launch_ray_range(One{V}, nullptr, p, p.tiles, 1, s);
}

void launch_raycast_march(const View& V, const RayConst& p, float* depth, float* vertex, float* normal,
                          hipStream_t s, int shape) {
  raycast_shaped(One{V}, nullptr, p, p.cam.width, p.cam.height, 1, shape, depth, vertex, normal, s);
}

void launch_raycast(const View& V, const RayConst& p, float* depth, float* vertex, float* normal,
                    hipStream_t s, const KernelClock* kc, int shape) {
  if (p.range) {
    tick(kc, s, "ray_range_kernel");
    launch_ray_range(One{V}, nullptr, p, p.tiles, 1, s);
  }
  tick(kc, s, "raycast_kernel");
  raycast_shaped(One{V}, nullptr, p, p.cam.width, p.cam.height, 1, shape, depth, vertex, normal, s);
  tick(kc, s, nullptr);
}

// Diagnostics: the distinct blocks the last raycast's lookups found (their
// ray_mark equals its stamp), added to the work counters.
__global__ __launch_bounds__(256) void ray_mark_count_kernel(View V) {
  const uint32_t n = V.ctr->block_hwm;
  uint32_t c = 0;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    c += V.ray_mark[p] == V.ray_stamp;
  wave_add_u64(&work_stripe(V.ctr).ray_blocks, c);
}

void launch_ray_mark_count(const View& V, hipStream_t s, int grid) {
  ray_mark_count_kernel<<<grid, 256, 0, s>>>(V);
}

uint32_t raycast_grad_partials(const RayConst& p) {
  return ((p.cam.width + 7) / 8) * ((p.cam.height + 7) / 8);
}

void launch_raycast_backward(const View& V, const RayConst& p, const float* obs, float* grad_sdf,
                             double* partials, double* out, float max_residual, float max_slope,
                             hipStream_t s) {
  dim3 grid((p.cam.width + 7) / 8, (p.cam.height + 7) / 8);
  if (p.range) launch_ray_range(One{V}, nullptr, p, p.tiles, 1, s);
  raycast_backward_kernel<<<grid, 64, 0, s>>>(V, p, obs, grad_sdf, partials, max_residual,
                                              max_slope);
  slot_reduce_kernel<<<1, 64, 0, s>>>(partials, grid.x * grid.y, kGradSlots, out);
}

void launch_batch_raycast(const BatchDev& B, float* depth, float* vertex, float* normal,
                          hipStream_t s, int shape) {
  launch_ray_range(B.src(), B.frames, RayConst{}, ((B.width + 7) / 8) * ((B.height + 7) / 8), B.nscenes, s);
  raycast_shaped(B.src(), B.frames, RayConst{}, B.width, B.height, B.nscenes, shape, depth, vertex, normal, s);
}

}  // namespace mh
