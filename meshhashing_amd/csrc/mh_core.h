// mh_core.h -- data layout, FP32 decision arithmetic and hash-table access
// shared by every kernel of the library.
//
// Arithmetic convention.  The whole library is compiled with
// -ffp-contract=off and IEEE division / square root.  Every expression that
// decides something discrete (which pixel a voxel projects to, which block a
// DDA step enters, which voxel a world point rounds to, the sign of an
// interpolated SDF) is written out with the reference's operation order, so
// that the FP64 oracle, which evaluates the same decision expressions in
// FP32, makes bit-identical decisions.  Values that are only accumulated
// (transforms feeding an SDF value, fusion updates) are free to use FP64 or
// explicit fma.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/meshhashing.h"

#define MH_HD __host__ __device__ __forceinline__

namespace mh {

constexpr int kBlockSide = 8;                 // core/common.h:46-47
constexpr int kBlockVoxels = 512;
constexpr int32_t kFreePtr = -2;              // core/common.h:54-58
constexpr int32_t kLockPtr = -1;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kMaxDdaSteps = 1024;            // mapping/allocate.cu:72
constexpr uint32_t kOverflowQueue = 1u << 16; // list inserts per frame

// ---------------------------------------------------------------- vectors
struct f3 { float x, y, z; };
struct i3 { int x, y, z; };
struct d3 { double x, y, z; };

MH_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
MH_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MH_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MH_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
MH_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
MH_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
MH_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
MH_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
MH_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MH_HD float length(f3 a) { return sqrtf(dot(a, a)); }
// v * (1/|v|): the reference multiplies by rsqrtf (helper_math normalize);
// here the reciprocal square root is the correctly rounded 1/sqrtf so that
// host and device agree bit for bit.
MH_HD f3 normalize(f3 v) { return v * (1.0f / sqrtf(dot(v, v))); }

// sign(s) = s < 0 ? -1 : +1 (sign(0) = +1), helper_math.h:1464-1467.
MH_HD int sgn(float s) { return s < 0.0f ? -1 : 1; }

// float -> int conversion with the GPU's saturating semantics (NaN -> 0),
// written out so host code (the oracle, the loaders) behaves identically.
MH_HD int f2i(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  return (int)f;
}

// ---------------------------------------------------------------- poses
// Row-major 4x4, m[r*4+c]; affine transform of a point (w = 1) as
// matrix.h:919-926: ((m11 x + m12 y) + m13 z) + m14.
struct M4 { float m[16]; };
MH_HD f3 xform(const M4& T, f3 p) {
  return f3{T.m[0] * p.x + T.m[1] * p.y + T.m[2] * p.z + T.m[3] * 1.0f,
            T.m[4] * p.x + T.m[5] * p.y + T.m[6] * p.z + T.m[7] * 1.0f,
            T.m[8] * p.x + T.m[9] * p.y + T.m[10] * p.z + T.m[11] * 1.0f};
}
// Direction (w = 0), matrix.h:908-916 with v.w = 0.
MH_HD f3 xform_dir(const M4& T, f3 d) {
  return f3{T.m[0] * d.x + T.m[1] * d.y + T.m[2] * d.z + T.m[3] * 0.0f,
            T.m[4] * d.x + T.m[5] * d.y + T.m[6] * d.z + T.m[7] * 0.0f,
            T.m[8] * d.x + T.m[9] * d.y + T.m[10] * d.z + T.m[11] * 0.0f};
}
// FP64 rows of a pose, used only for SDF values.
struct D34 { double m[12]; };

// ---------------------------------------------------------------- geometry
// GeometryHelper, geometry/geometry_helper.h:36-217.
struct Geo {
  float voxel_size;
  float truncation_distance;
  float truncation_distance_scale;
  float sdf_upper_bound;
  float weight_sample;
};
struct Cam {  // SensorParams / RayCasterParams intrinsics
  float fx, fy, cx, cy;
  uint32_t width, height;
  float min_depth, max_depth;
};

MH_HD float truncate_distance(const Geo& g, float z) {
  return g.truncation_distance + g.truncation_distance_scale * z;
}
// WorldToVoxeli: int(p/vs + sign(p)*0.5), half away from zero.
MH_HD i3 world_to_voxel(const Geo& g, f3 w) {
  f3 p = w / g.voxel_size;
  return i3{f2i(p.x + (float)sgn(p.x) * 0.5f), f2i(p.y + (float)sgn(p.y) * 0.5f),
            f2i(p.z + (float)sgn(p.z) * 0.5f)};
}
// WorldToVoxeli along one axis.
__device__ __forceinline__ int vox1(float w, float vs) {
  const float p = w / vs;
  return f2i(p + (float)sgn(p) * 0.5f);
}

// vox1(pc, vs) from a multiply by 1/vs instead of the IEEE division;
// accepted only when the quotient is more than 1e-3 voxel away from every
// rounding tie and from 0 (and |q| < 2048, where the multiply is within 2 ulp
// of the quotient), so it equals the exact result.
__device__ __forceinline__ bool fast_corner(float pc, float inv_vs, int* c) {
  const float q = pc * inv_vs;
  const float aq = fabsf(q);
  if (!(aq < 2048.0f) || aq < 1e-3f) return false;
  const float f = aq - floorf(aq);
  if (fabsf(f - 0.5f) < 1e-3f) return false;
  *c = f2i(q + (float)sgn(q) * 0.5f);
  return true;
}

// The two corner voxels vox1(pc, vs) and vox1(pc + vs, vs) of a trilinear
// sample along one axis (spatial_query.h:23-66 rounds both).  Clear of a
// tie, the upper one is the lower + 1: pc + vs is within 1.3e-4 voxel of
// exact, so its quotient stays clear of the next tie.
__device__ __forceinline__ void corner_pair(float pc, float vs, float inv_vs, int* c0, int* c1) {
  if (fast_corner(pc, inv_vs, c0)) {
    *c1 = *c0 + 1;
  } else {
    *c0 = vox1(pc + 0.0f, vs);
    *c1 = vox1(pc + vs, vs);
  }
}

MH_HD int voxel_to_block1(int v) { return (v - (v < 0) * (kBlockSide - 1)) / kBlockSide; }
MH_HD i3 voxel_to_block(i3 v) {
  return i3{voxel_to_block1(v.x), voxel_to_block1(v.y), voxel_to_block1(v.z)};
}
MH_HD f3 voxel_to_world(const Geo& g, i3 v) {
  return f3{(float)v.x * g.voxel_size, (float)v.y * g.voxel_size, (float)v.z * g.voxel_size};
}
MH_HD f3 block_to_world(const Geo& g, i3 b) {
  return voxel_to_world(g, i3{b.x * kBlockSide, b.y * kBlockSide, b.z * kBlockSide});
}
MH_HD i3 world_to_block(const Geo& g, f3 w) { return voxel_to_block(world_to_voxel(g, w)); }

// CameraProjectToImagef / ImageReprojectToCamera, geometry_helper.h:141-168.
MH_HD float proj_u(const Cam& c, f3 p) { return p.x * c.fx / p.z + c.cx; }
MH_HD float proj_v(const Cam& c, f3 p) { return p.y * c.fy / p.z + c.cy; }
MH_HD f3 reproject(const Cam& c, uint32_t ux, uint32_t uy, float depth) {
  const float x = ((float)ux - c.cx) / c.fx;
  const float y = ((float)uy - c.cy) / c.fy;
  return f3{depth * x, depth * y, depth};
}
MH_HD float normalize_depth(float z, float mn, float mx) { return (z - mn) / (mx - mn); }

// IsPointInCameraFrustum / IsBlockInCameraFrustum, geometry_helper.h:186-216.
MH_HD bool point_in_frustum(const M4& cTw, f3 w, const Cam& c) {
  f3 p = xform(cTw, w);
  float u = proj_u(c, p), v = proj_v(c, p);
  float wm1 = (float)c.width - 1.0f, hm1 = (float)c.height - 1.0f;
  float nx = (2.0f * u - wm1) / wm1;
  float ny = (hm1 - 2.0f * v) / hm1;
  float nz = normalize_depth(p.z, c.min_depth, c.max_depth);
  nx *= 0.95f; ny *= 0.95f; nz *= 0.95f;
  return !(nx < -1.0f || nx > 1.0f || ny < -1.0f || ny > 1.0f || nz < 0.0f || nz > 1.0f);
}
MH_HD bool block_in_frustum(const Geo& g, const M4& cTw, i3 b, const Cam& c) {
  f3 w = block_to_world(g, b);
  float h = g.voxel_size * 0.5f * ((float)kBlockSide - 1.0f);
  return point_in_frustum(cTw, f3{w.x + h, w.y + h, w.z + h}, c);
}

// ---------------------------------------------------------------- hash table
// Entry: 16 bytes.  The block position is packed into a 64-bit key
// (21 bits per axis, two's complement) so that a slot is claimed with one
// 64-bit compare-and-swap; ptr is the block-pool index, offset the list link
// relative to the bucket's last slot (core/hash_entry.h:11-29).
struct alignas(16) Entry {
  uint64_t key;
  int32_t ptr;
  uint32_t offset;
};
constexpr int kCoordBits = 21;
constexpr int kCoordMax = (1 << (kCoordBits - 1)) - 1;
MH_HD bool coord_ok(i3 p) {
  return p.x >= -kCoordMax && p.x <= kCoordMax && p.y >= -kCoordMax && p.y <= kCoordMax &&
         p.z >= -kCoordMax && p.z <= kCoordMax;
}
MH_HD uint64_t pack_key(i3 p) {
  const uint64_t m = (1ull << kCoordBits) - 1;
  return (((uint64_t)(uint32_t)p.x & m) << (2 * kCoordBits)) |
         (((uint64_t)(uint32_t)p.y & m) << kCoordBits) | ((uint64_t)(uint32_t)p.z & m);
}
MH_HD int sext21(uint64_t v) {
  int x = (int)(v & ((1u << kCoordBits) - 1));
  return (x ^ (1 << (kCoordBits - 1))) - (1 << (kCoordBits - 1));
}
MH_HD i3 unpack_key(uint64_t k) {
  return i3{sext21(k >> (2 * kCoordBits)), sext21(k >> kCoordBits), sext21(k)};
}
// HashBucketForBlockPos, core/hash_table.h:259-269: 32-bit wrapping products,
// xor, then an *unsigned* modulo (bucket_count is uint, :19).
MH_HD uint32_t hash_bucket(i3 p, uint32_t bucket_count) {
  uint32_t h = ((uint32_t)p.x * 73856093u) ^ ((uint32_t)p.y * 19349669u) ^
               ((uint32_t)p.z * 83492791u);
  return h % bucket_count;
}

// Work counters (accumulated since reset), for throughput and the roofline.
// They are summed by one atomic per wave; every wave of a launch adding to
// one address serialises those atomics device-wide (they cost more than a
// whole fusion pass at high occupancy), so the counters are striped over
// kWorkStripes 256-byte lines, picked by workgroup, and the host adds the
// stripes up.
constexpr uint32_t kWorkStripes = 64;
struct alignas(256) WorkStripe {
  unsigned long long voxel_updates;      // voxels whose fused state changed
  unsigned long long blocks_integrated;  // worklist blocks visited by fusion
  unsigned long long mc_vertex_writes;   // vertex records written by MC pass 1
  unsigned long long mc_triangle_writes; // triangles of cubes whose configuration changed
  unsigned long long mc_blocks;          // worklist blocks meshed
  unsigned long long ray_samples;        // valid trilinear samples of the raycaster
  unsigned long long tri_live;           // live triangles (sum of table rows; deltas wrap)
  unsigned long long ray_blocks;         // distinct blocks each diagnostic raycast sampled
};

// Device-wide counters and status, one instance per volume.
struct Counters {
  int32_t block_top;       // heap counter (hash_table.h:279-286)
  uint32_t block_hwm;      // 1 + largest block index ever handed out
  uint32_t cand_count;     // worklist length
  uint32_t ovf_count;      // keys queued for list insertion this frame
  int32_t vertex_top;      // vertex heap counter (mesh.h:53-59)
  uint32_t gc_count;       // blocks on the garbage list (RecycleGarbageBlockArray)
  uint32_t flags;          // MH_STATUS_* bits
  uint32_t dropped;
  uint32_t dda_cap;
  uint32_t list_inserts;
  uint32_t gc_free_count;  // garbage blocks left with no vertex (to be freed)
  uint32_t mc_next;        // marching cubes: worklist items handed out on demand (after the static ones)
  uint32_t vfree_count;    // entries of vfree
  uint32_t vfree_done;     // workgroups done pushing vfree on the vertex heap
  WorkStripe work[kWorkStripes];
};

// The stripe this workgroup adds its work counts to.
__device__ __forceinline__ WorkStripe& work_stripe(Counters* c) {
  return c->work[blockIdx.x % kWorkStripes];
}

// Sum of the stripes (host side).
inline WorkStripe work_totals(const Counters& c) {
  WorkStripe t{};
  for (uint32_t i = 0; i < kWorkStripes; ++i) {
    t.voxel_updates += c.work[i].voxel_updates;
    t.blocks_integrated += c.work[i].blocks_integrated;
    t.mc_vertex_writes += c.work[i].mc_vertex_writes;
    t.mc_triangle_writes += c.work[i].mc_triangle_writes;
    t.mc_blocks += c.work[i].mc_blocks;
    t.ray_samples += c.work[i].ray_samples;
    t.tri_live += c.work[i].tri_live;
    t.ray_blocks += c.work[i].ray_blocks;
  }
  return t;
}

// Sum v over the wave and add it to *dst once.  Every lane of the wave must
// reach the call.
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63u) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// Vertex record (core/vertex.h:11-26), 48 bytes: three 16-byte stores.
struct alignas(16) Vertex {
  float px, py, pz, radius;
  float nx, ny, nz;
  int32_t ref_count;
  float cr, cg, cb, pad;
};

// Everything a kernel needs, passed by value.
struct View {
  // hash table
  Entry* entries;
  uint32_t bucket_count, bucket_size, entry_count, linked_list_size;
  uint64_t* block_keys;  // key of the block at each pool index (kEmptyKey: free)
  int32_t* block_heap;
  uint32_t capacity;
  uint32_t* occ;         // occupancy filter, 2^kOccBits bits
  uint32_t* occ2;        // super-block (4^3 blocks) occupancy filter, 2^kOcc2Bits bits
  Counters* ctr;
  uint64_t* ovf;         // overflow queue
  // voxels, pool-index-major, 512 per block
  float* sdf;
  float* weight;         // inv_sigma2 of the reference's Voxel
  uint32_t* color;       // RGBA8
  float* va;             // Beta a (PSDF only)
  float* vb;             // Beta b (PSDF only)
  // mesh units (core/voxel.h:27-60), 512 per block
  int32_t* vptr;         // 3 per voxel: x, y, z edge vertex
  int32_t* vfree;        // vertex slots freed by the surfel pass, pushed on the heap after it
  uint16_t* cube;        // curr | prev << 8
  int32_t* surfel;       // per block: inner, boundary surfel counts (block.h:17-18)
  uint32_t* block_mark;  // per block: last meshing pass that processed it
  // mesh pools
  Vertex* vertices;
  int32_t* vertex_heap;
  uint32_t max_vertices;
  uint32_t max_triangles;  // bound on live triangles (MeshParams)
  // worklist
  int32_t* cand;
  uint8_t* cand_flag;    // per worklist entry: garbage flag (recycle.cu)
  int32_t* life;         // per block: life_count_down (block.h:18)
  int32_t* gc_list;      // garbage blocks (pool indices)
  int32_t* gc_free;      // garbage blocks with no vertex left
  // primal-dual state (lazily allocated), 512 per block
  float* pd_sdf0;
  float* pd_sdf_bar;
  float* pd_w;
  float* pd_p;           // 3 per voxel, interleaved
  uint8_t* pd_mask;
  // geometry and frame
  Geo geo;
  Cam cam;
  float range_factor;
  float* depth;          // W*H metric depth of the current frame
  const uint16_t* raw;   // W*H raw depth not yet converted (alloc converts it), or null
  const uint8_t* raw_color;  // W*H BGRA colour not yet converted (alloc converts it), or null
  uint32_t* color_in;    // W*H fusion colour of the current frame, R | G << 8 | B << 16
  int has_color;         // the current frame has a colour image
  float* inlier;         // W*H PSDF inlier ratio
  M4 wTc, cTw;           // FP32 pose (decisions)
  D34 cTw64;             // FP64 world->camera (values)
  // raycast diagnostics: per block, the stamp of the last raycast whose
  // block lookups found it (null: not recorded)
  uint32_t* ray_mark;
  uint32_t ray_stamp;
};

// Occupancy filter: one bit per hashed block key (2^kOccBits bits, 2 MiB,
// L2-resident).  Set on every insertion, never cleared within a volume's
// life, so a clear bit proves a block is absent; a set bit falls through to
// the hash table.  Most lookups of the raycaster and of the stencil kernels
// are for empty space and end here with one cached load.
constexpr int kOccBits = 24;
MH_HD uint32_t occ_bit(uint64_t key) {
  return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - kOccBits));
}
// Super-block filter: one bit per hashed key of a 4 x 4 x 4-block cell, set
// when any block of the cell is inserted and never cleared within a
// volume's life, so a clear bit proves the cell holds no block (the
// raycast steps over such a cell at once).
constexpr int kOcc2Bits = 20;
constexpr int kSuperShift = 2;  // blocks per super-block axis: 1 << kSuperShift
MH_HD uint64_t super_key(i3 b) {
  return pack_key(i3{b.x >> kSuperShift, b.y >> kSuperShift, b.z >> kSuperShift});
}
MH_HD uint32_t occ2_bit(uint64_t skey) {
  return (uint32_t)((skey * 0xC2B2AE3D27D4EB4Full) >> (64 - kOcc2Bits));
}
__device__ __forceinline__ void occ_mark(const View& V, uint64_t key, i3 b) {
  const uint32_t ob = occ_bit(key);
  atomicOr(&V.occ[ob >> 5], 1u << (ob & 31u));
  const uint32_t sb = occ2_bit(super_key(b));
  atomicOr(&V.occ2[sb >> 5], 1u << (sb & 31u));
}

__device__ __forceinline__ bool occ_test(const uint32_t* occ, uint64_t key) {
  const uint32_t b = occ_bit(key);
  return (occ[b >> 5] >> (b & 31u)) & 1u;
}

// GetEntry, hash_table.h:57-93: bucket scan, then the list that starts at
// the bucket's last slot.  Returns the pool index or kFreePtr.  A key is
// placed in the first free slot of its bucket and the tables run at a few
// keys per thousand buckets, so slot 0 is tried alone first (one 16-byte
// load); the other nine slots are then loaded together (one round trip).
// kSlot0Done: the caller has compared slot 0 already.
template <bool kSlot0Done = false>
__device__ __forceinline__ int32_t hash_find_key(const View& V, i3 p, uint64_t key) {
  const uint32_t first = hash_bucket(p, V.bucket_count) * V.bucket_size;
  const uint32_t last = first + V.bucket_size - 1;
  uint32_t last_offset;
  if (V.bucket_size == 10) {
    if (!kSlot0Done) {
      const Entry e0 = V.entries[first];
      if (e0.key == key) return e0.ptr;
    }
    Entry e[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) e[j] = V.entries[first + 1 + j];
    int32_t found = kFreePtr;
#pragma unroll
    for (int j = 0; j < 9; ++j)
      if (e[j].key == key) found = e[j].ptr;
    if (e[0].key == key || e[1].key == key || e[2].key == key || e[3].key == key ||
        e[4].key == key || e[5].key == key || e[6].key == key || e[7].key == key ||
        e[8].key == key)
      return found;
    last_offset = e[8].offset;
  } else {
    for (uint32_t j = kSlot0Done ? 1u : 0u; j < V.bucket_size; ++j) {
      const Entry e = V.entries[first + j];
      if (e.key == key) return e.ptr;
    }
    last_offset = V.entries[last].offset;
  }
  // list: its head is the last slot (already compared above)
  if (V.linked_list_size == 0 || last_offset == 0) return kFreePtr;
  uint32_t i = (last + last_offset) % V.entry_count;
  for (uint32_t it = 1; it < V.linked_list_size; ++it) {
    const Entry e = V.entries[i];
    if (e.key == key) return e.ptr;
    if (e.offset == 0) break;
    i = (last + e.offset) % V.entry_count;
  }
  return kFreePtr;
}

__device__ __forceinline__ int32_t hash_find(const View& V, i3 p) {
  const uint64_t key = pack_key(p);
  if (!occ_test(V.occ, key)) return kFreePtr;
  return hash_find_key(V, p, key);
}

// hash_find for latency-bound callers: the occupancy word and the bucket's
// first slot are loaded together (one round trip instead of two for a key
// in slot 0, the usual place), at the cost of a wasted slot load for an
// absent key.  Same result as hash_find.
__device__ __forceinline__ int32_t hash_find_eager(const View& V, i3 p) {
  const uint64_t key = pack_key(p);
  const uint32_t bit = occ_bit(key);
  const uint32_t first = hash_bucket(p, V.bucket_count) * V.bucket_size;
  const uint32_t word = V.occ[bit >> 5];
  const Entry e0 = V.entries[first];
  if (!((word >> (bit & 31u)) & 1u)) return kFreePtr;
  if (e0.key == key) return e0.ptr;
  return hash_find_key<true>(V, p, key);
}

__device__ __forceinline__ void flag(const View& V, uint32_t bit) {
  atomicOr(&V.ctr->flags, bit);
}

// Block-pool pop (hash_table.h:279-286): the heap holds free indices,
// initialised so that pops return 0, 1, 2, ...
__device__ __forceinline__ int32_t alloc_block(const View& V) {
  const int32_t addr = atomicSub(&V.ctr->block_top, 1);
  if (addr < 0) {
    atomicAdd(&V.ctr->block_top, 1);  // undo: the top never stays below -1
    flag(V, MH_STATUS_BLOCK_HEAP_EXHAUSTED);
    return kFreePtr;
  }
  const int32_t p = V.block_heap[addr];
  atomicMax(&V.ctr->block_hwm, (uint32_t)p + 1u);
  return p;
}

// Local voxel index <-> offset, geometry_helper.h:111-129.
MH_HD i3 devectorize(uint32_t i) {
  return i3{(int)(i % 8u), (int)((i % 64u) / 8u), (int)(i / 64u)};
}

}  // namespace mh
