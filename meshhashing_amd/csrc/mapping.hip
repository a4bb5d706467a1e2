// mapping.hip -- frame input, block allocation, worklist compaction and
// TSDF / PSDF fusion kernels for gfx950.
//
// Reference behaviour followed here (file:line in theNded/MeshHashing):
//   depth conversion        sensor/preprocess.cu:25-42
//   inlier-ratio reset      sensor/preprocess.cu:10-23
//   block allocation        mapping/allocate.cu:8-99, core/hash_table.h:96-179
//   visible-block compaction core/collect_block_array.cu:25-91
//   TSDF update             mapping/update_simple.cu:13-77, core/voxel.h:107-115
//   PSDF inlier prediction  mapping/update_bayesian.cu:16-101
//   PSDF update             mapping/update_bayesian.cu:103-185
#include "mh_kernels.h"

namespace mh {

// ------------------------------------------------------------------ resets
__global__ void reset_entries_kernel(Entry* e, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    e[i] = Entry{kEmptyKey, kFreePtr, 0u};
}

__global__ void reset_pool_kernel(View V) {
  const uint32_t cap = V.capacity;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
    V.block_heap[i] = (int32_t)(cap - 1u - i);
    V.block_keys[i] = kEmptyKey;
    V.block_mark[i] = 0u;
    V.life[i] = kBlockLife;
    V.surfel[2 * i] = 0;
    V.surfel[2 * i + 1] = 0;
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (1u << (kOccBits - 5));
       i += gridDim.x * blockDim.x)
    V.occ[i] = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (1u << (kOcc2Bits - 5));
       i += gridDim.x * blockDim.x)
    V.occ2[i] = 0u;
  const uint32_t nv = V.max_vertices;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    V.vertex_heap[i] = (int32_t)(nv - 1u - i);
    Vertex z{};
    V.vertices[i] = z;
  }
}

// Voxel / mesh-unit state of a block, as Block::Clear (core/block.h:27-39).
__global__ void reset_voxels_kernel(View V, size_t nvox) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox;
       i += (size_t)gridDim.x * blockDim.x) {
    V.sdf[i] = 0.0f;
    V.weight[i] = 0.0f;
    V.color[i] = 0u;
    if (V.va) { V.va[i] = 0.0f; V.vb[i] = 0.0f; }
    V.vptr[3 * i + 0] = kFreePtr;
    V.vptr[3 * i + 1] = kFreePtr;
    V.vptr[3 * i + 2] = kFreePtr;
    V.cube[i] = 0;
  }
}

// Counters.  The vertex pool reserves index 0: the reference's
// PredictOutlierRatio skips vertex pointers <= 0 (update_bayesian.cu:84),
// i.e. whichever vertex happened to be allocated first; reserving the slot
// keeps that expression while excluding no real vertex.
__global__ void reset_counters_kernel(View V) {
  uint32_t* w = reinterpret_cast<uint32_t*>(V.ctr);
  for (uint32_t i = threadIdx.x; i < sizeof(Counters) / 4; i += blockDim.x) w[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) {
    V.ctr->block_top = (int32_t)V.capacity - 1;
    V.ctr->vertex_top = (int32_t)V.max_vertices - 2;
  }
}

__global__ void set_u32_kernel(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ------------------------------------------------------------ frame input
__global__ void convert_depth_kernel(const uint16_t* __restrict__ raw, float* __restrict__ depth,
                                     float* __restrict__ inlier, uint32_t n, float range_factor,
                                     float dmin, float dmax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = range_factor * (float)(int16_t)raw[i];
  depth[i] = (d >= dmin && d <= dmax) ? d : -INFINITY;
  if (inlier) inlier[i] = 0.1f;
}

// ConvertColorFormatKernel (sensor/preprocess.cu:44-58) followed by the
// float -> uchar conversion of UpdateBlocksSimpleKernel (update_simple.cu:
// 68-70): valid iff B, G, R are all non-zero; channel = uchar(255 * (c /
// 255.0f)), a truncation that clamps to [0, 255] (an invalid pixel's -inf
// becomes 0).
__device__ __forceinline__ uint32_t fuse_channel(uint32_t c) {
  const float f = 255.0f * ((float)c / 255.0f);
  const int i = f2i(f);
  return (uint32_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
}
__device__ __forceinline__ uint32_t fuse_color(uint32_t c) {
  const uint32_t b = c & 255u, g = (c >> 8) & 255u, r = (c >> 16) & 255u;
  return (b && g && r) ? (fuse_channel(r) | (fuse_channel(g) << 8) | (fuse_channel(b) << 16)) : 0u;
}
// ConvertDepthFormatKernel (sensor/preprocess.cu:25-42).
__device__ __forceinline__ float convert_depth(uint16_t raw, float range_factor, float dmin, float dmax) {
  const float d = range_factor * (float)(int16_t)raw;
  return (d >= dmin && d <= dmax) ? d : -INFINITY;
}

__global__ void convert_color_kernel(const uint32_t* __restrict__ bgra, uint32_t* __restrict__ out,
                                     uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = fuse_color(bgra[i]);
}

// ------------------------------------------------------------- allocation
// Plain-load presence test (bucket, then list).  Keys are never removed
// inside a frame, so a hit is final; a miss is re-checked coherently.
// The occupancy filter first (every insertion sets the key's bit, so a clear
// bit means absent); then the bucket, its slots loaded together (one memory
// round trip for the usual ten), then the list.
__device__ __forceinline__ bool key_present(const View& V, uint64_t key, uint32_t first) {
  if (!occ_test(V.occ, key)) return false;
  const uint32_t last = first + V.bucket_size - 1;
  if (V.bucket_size == 10) {
    uint64_t k[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) k[j] = V.entries[first + j].key;
    bool hit = false;
#pragma unroll
    for (int j = 0; j < 10; ++j) hit = hit || k[j] == key;
    if (hit) return true;
  } else {
    bool hit = false;
    for (uint32_t j = 0; j < V.bucket_size; ++j) hit = hit || V.entries[first + j].key == key;
    if (hit) return true;
  }
  uint32_t i = last;
  for (uint32_t it = 0; it < V.linked_list_size; ++it) {
    const Entry e = V.entries[i];
    if (e.key == key) return true;
    if (e.offset == 0) break;
    i = (last + e.offset) % V.entry_count;
  }
  return false;
}

// Lossless insert.  A free bucket slot is claimed by a 64-bit CAS on its key;
// losing a race re-scans the bucket with device-coherent loads (each lost
// CAS means one more slot is taken, so this ends within bucket_size + 1
// rounds).  A full bucket defers the key to the serial list-insert pass.
__device__ void insert_block(const View& V, i3 p) {
  if (!coord_ok(p)) {
    flag(V, MH_STATUS_COORD_RANGE);
    return;
  }
  const uint64_t key = pack_key(p);
  const uint32_t first = hash_bucket(p, V.bucket_count) * V.bucket_size;
  if (key_present(V, key, first)) return;
  for (uint32_t round = 0; round <= V.bucket_size; ++round) {
    // every slot's key read before any is tested (no early exit between the
    // loads, so they are in flight together)
    int empty = -1;
    bool seen = false;
    if (V.bucket_size == 10) {
      uint64_t k[10];
#pragma unroll
      for (int j = 0; j < 10; ++j)
        k[j] = __hip_atomic_load(&V.entries[first + j].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int j = 9; j >= 0; --j) {
        seen = seen || k[j] == key;
        if (k[j] == kEmptyKey) empty = j;
      }
    } else {
      for (uint32_t j = 0; j < V.bucket_size; ++j) {
        const uint64_t k = __hip_atomic_load(&V.entries[first + j].key, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        seen = seen || k == key;
        if (k == kEmptyKey && empty < 0) empty = (int)j;
      }
    }
    if (seen) return;
    if (empty < 0) break;
    Entry* e = &V.entries[first + (uint32_t)empty];
    const uint64_t old = atomicCAS((unsigned long long*)&e->key, (unsigned long long)kEmptyKey,
                                   (unsigned long long)key);
    if (old == kEmptyKey) {
      occ_mark(V, key, p);
      const int32_t ptr = alloc_block(V);
      if (ptr < 0) {
        // Heap exhausted (flagged by alloc_block): give the slot back, so
        // the key is not taken for allocated and a later frame retries it
        // once recycling has returned pool slots.
        __hip_atomic_store(&e->key, kEmptyKey, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      e->ptr = ptr;
      V.block_keys[ptr] = key;
      return;
    }
    if (old == key) return;
  }
  const uint32_t q = atomicAdd(&V.ctr->ovf_count, 1u);
  if (q < kOverflowQueue)
    V.ovf[q] = key;
  else
    flag(V, MH_STATUS_OVERFLOW_QUEUE_FULL);
}

// AllocBlockArrayKernel, allocate.cu:8-99: per pixel, walk the blocks between
// depth - trunc and depth + trunc with a 3-D DDA (Amanatides-Woo in block
// units) and insert every block whose centre lies in the frustum.
template <class Src>
__global__ __launch_bounds__(256) void alloc_kernel(Src src_, const SceneFrame* __restrict__ tab) {
  const Src src = with_table(src_, tab);
  const View V = src.view(blockIdx.z);  // a copy: read once, not reloaded after stores
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15);
  const uint32_t y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= V.cam.width || y >= V.cam.height) return;
  const Geo& g = V.geo;
  const uint32_t pix = y * V.cam.width + x;
  float depth;
  if (V.raw) {  // sensor preprocessing fused in: this pixel's conversion
    depth = convert_depth(V.raw[pix], V.range_factor, V.cam.min_depth, V.cam.max_depth);
    V.depth[pix] = depth;
    if (V.inlier) V.inlier[pix] = 0.1f;
    if (V.raw_color) V.color_in[pix] = fuse_color(reinterpret_cast<const uint32_t*>(V.raw_color)[pix]);
  } else {
    depth = V.depth[pix];
  }
  if (depth == -INFINITY || depth == 0.0f || depth >= g.sdf_upper_bound) return;

  const float tr = truncate_distance(g, depth);
  const float near_d = fminf(g.sdf_upper_bound, depth - tr);
  const float far_d = fminf(g.sdf_upper_bound, depth + tr);
  if (near_d >= far_d) return;

  const f3 cn = reproject(V.cam, x, y, near_d);
  const f3 cf = reproject(V.cam, x, y, far_d);
  const f3 wn = xform(V.wTc, cn);
  const f3 wf = xform(V.wTc, cf);
  const f3 dir = normalize(wf - wn);

  const i3 bn = world_to_block(g, wn);
  const i3 bf = world_to_block(g, wf);
  const f3 step = mk3((float)sgn(dir.x), (float)sgn(dir.y), (float)sgn(dir.z));

  // BlockToWorld(b_near + clamp(step, 0, 1)) - 0.5 vs
  const i3 bo = i3{bn.x + f2i(fminf(fmaxf(step.x, 0.0f), 1.0f)),
                   bn.y + f2i(fminf(fmaxf(step.y, 0.0f), 1.0f)),
                   bn.z + f2i(fminf(fmaxf(step.z, 0.0f), 1.0f))};
  const f3 bw = block_to_world(g, bo);
  const float hv = 0.5f * g.voxel_size;
  const f3 nvc = mk3(bw.x - hv, bw.y - hv, bw.z - hv);
  f3 t = (nvc - wn) / dir;
  const float bsz = (float)kBlockSide;
  f3 dt = mk3(step.x * bsz * g.voxel_size, step.y * bsz * g.voxel_size,
              step.z * bsz * g.voxel_size) / dir;
  const i3 bound = i3{f2i((float)bf.x + step.x), f2i((float)bf.y + step.y),
                      f2i((float)bf.z + step.z)};
  if (dir.x == 0.0f) { t.x = INFINITY; dt.x = INFINITY; }
  if (dir.y == 0.0f) { t.y = INFINITY; dt.y = INFINITY; }
  if (dir.z == 0.0f) { t.z = INFINITY; dt.z = INFINITY; }

  const int sx = (int)step.x, sy = (int)step.y, sz = (int)step.z;
  i3 cur = bn;
  for (int it = 0; it < kMaxDdaSteps; ++it) {
    if (block_in_frustum(g, V.cTw, cur, V.cam)) insert_block(V, cur);
    if (t.x < t.y && t.x < t.z) {
      cur.x += sx;
      if (cur.x == bound.x) return;
      t.x += dt.x;
    } else if (t.y < t.z) {
      cur.y += sy;
      if (cur.y == bound.y) return;
      t.y += dt.y;
    } else {
      cur.z += sz;
      if (cur.z == bound.z) return;
      t.z += dt.z;
    }
  }
  atomicAdd(&V.ctr->dda_cap, 1u);
  flag(V, MH_STATUS_DDA_CAP);
}

// Overflow keys (their bucket was full when the allocation pass met them):
// the reference's list insertion (hash_table.h:145-177): probe offsets
// 1..linked_list_size from the bucket's last slot, skipping multiples of
// bucket_size, and link the first free slot in as the list's head.  One
// 1024-lane workgroup per scene (blockIdx.x) works the queue in chunks of
// 1024 keys: a chunk's duplicates (one key is queued by every ray that met
// it) are dropped through an LDS set, keys already in the table are
// skipped, and the rest are inserted in parallel.  A slot is claimed by a
// CAS on its key and gets a pool index before it is linked; the link is a
// CAS on the head offset with the slot's own offset written first, so a
// concurrent walk of the list never reads a half-linked slot.  Slots of one
// list can come out in any order (the reference's order is a race too).
// After a frame's allocation (kClearWork) the kernel also clears the
// worklist counter for the collect pass that follows; loading blocks leaves
// the worklist as it is.
constexpr uint32_t kListChunk = 1024;
constexpr uint32_t kListSet = 2 * kListChunk;  // LDS set of a chunk's keys

// key_present with device-coherent loads (slots are changed by other lanes
// of the same pass).
__device__ __forceinline__ bool key_present_coherent(const View& V, uint64_t key, uint32_t first) {
  for (uint32_t j = 0; j < V.bucket_size; ++j)
    if (__hip_atomic_load(&V.entries[first + j].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == key) return true;
  const uint32_t last = first + V.bucket_size - 1;
  uint32_t i = last;
  for (uint32_t it = 0; it < V.linked_list_size; ++it) {
    if (__hip_atomic_load(&V.entries[i].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == key) return true;
    const uint32_t o = __hip_atomic_load(&V.entries[i].offset, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (o == 0) break;
    i = (last + o) % V.entry_count;
  }
  return false;
}

// Link slot i (offset o from the bucket's last slot) in as the list's head.
__device__ __forceinline__ void list_link(const View& V, uint32_t last, uint32_t i, uint32_t o) {
  uint32_t* head = &V.entries[last].offset;
  uint32_t old = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    __hip_atomic_store(&V.entries[i].offset, old, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();  // the slot's offset before the head that leads to it
    const uint32_t seen = atomicCAS(head, old, o);
    if (seen == old) return;
    old = seen;
  }
}

// work_count (a batch's concatenated worklist length, or null) is cleared
// with the worklist counters.
template <class Src, bool kClearWork = true>
__global__ __launch_bounds__(kListChunk) void list_insert_kernel(Src src_, const SceneFrame* __restrict__ tab,
                                                                 uint32_t* work_count = nullptr) {
  const Src src = with_table(src_, tab);
  const View& V = src.view(blockIdx.x);
  __shared__ unsigned long long set[kListSet];
  const uint32_t n = min(V.ctr->ovf_count, kOverflowQueue);
  for (uint32_t base = 0; base < n; base += kListChunk) {
    for (uint32_t k = threadIdx.x; k < kListSet; k += blockDim.x) set[k] = kEmptyKey;
    __syncthreads();
    const uint32_t q = base + threadIdx.x;
    bool mine = false;
    uint64_t key = kEmptyKey;
    if (q < n) {
      key = V.ovf[q];
      // first of its chunk to claim the key in the set
      uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 53) & (kListSet - 1);
      while (true) {
        const unsigned long long old = atomicCAS(&set[h], (unsigned long long)kEmptyKey, (unsigned long long)key);
        if (old == kEmptyKey) {
          mine = true;
          break;
        }
        if (old == key) break;
        h = (h + 1) & (kListSet - 1);
      }
    }
    // presence tests, then (after a barrier) the inserts: a test never runs
    // beside an insert of the same chunk
    const i3 p = unpack_key(key);
    const uint32_t first = mine ? hash_bucket(p, V.bucket_count) * V.bucket_size : 0;
    const bool need = mine && !key_present_coherent(V, key, first);
    __syncthreads();
    if (need) {
      {
        occ_mark(V, key, p);
        bool done = false;
        // a bucket slot freed since the allocation pass
        for (uint32_t j = 0; j < V.bucket_size && !done; ++j) {
          Entry* e = &V.entries[first + j];
          if (__hip_atomic_load(&e->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kEmptyKey) continue;
          if (atomicCAS((unsigned long long*)&e->key, (unsigned long long)kEmptyKey, (unsigned long long)key) !=
              kEmptyKey)
            continue;
          const int32_t ptr = alloc_block(V);  // an exhausted heap is flagged there
          if (ptr >= 0) {
            e->ptr = ptr;
            V.block_keys[ptr] = key;
          } else {
            __hip_atomic_store(&e->key, kEmptyKey, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          }
          done = true;
        }
        const uint32_t last = first + V.bucket_size - 1;
        uint32_t offset = 0;
        for (uint32_t it = 0; it < V.linked_list_size && !done; ++it) {
          ++offset;
          if (offset % V.bucket_size == 0) continue;
          const uint32_t i = (last + offset) % V.entry_count;
          Entry* e = &V.entries[i];
          if (__hip_atomic_load(&e->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kEmptyKey) continue;
          if (atomicCAS((unsigned long long*)&e->key, (unsigned long long)kEmptyKey, (unsigned long long)key) !=
              kEmptyKey)
            continue;
          const int32_t ptr = alloc_block(V);
          if (ptr >= 0) {
            e->ptr = ptr;
            V.block_keys[ptr] = key;
            list_link(V, last, i, offset);
            atomicAdd(&V.ctr->list_inserts, 1u);
          } else {
            __hip_atomic_store(&e->key, kEmptyKey, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          }
          done = true;
        }
        if (!done) {
          atomicAdd(&V.ctr->dropped, 1u);
          atomicOr(&V.ctr->flags, MH_STATUS_INSERT_DROPPED);
        }
      }
    }
    __threadfence();
    __syncthreads();  // the chunk's inserts before the next chunk's presence tests
  }
  if (threadIdx.x == 0) {
    V.ctr->ovf_count = 0;
    if (kClearWork) V.ctr->cand_count = 0;
    if (kClearWork && work_count && blockIdx.x == 0) *work_count = 0;
  }
}

// --------------------------------------------------------------- compaction
// Stream compaction of allocated blocks (optionally in-frustum) over the
// block pool, not over the whole entry table: the pool is dense up to its
// high-water mark, which is ~100x smaller than entry_count.  One atomic per
// wave; the worklist order is not deterministic (neither is the reference's).
// Scene = blockIdx.y.  A batch also appends (scene, block) items to the
// concatenated worklist that the batched block kernels run over.
// stamp (batches): also give each collected block the frame's marching-cubes
// pass stamp (what mc_mark_kernel would write), so the meshing pass that
// follows in the chain needs no marking launch.
template <class Src>
__global__ __launch_bounds__(256) void collect_kernel(Src src_, const SceneFrame* __restrict__ tab, int frustum, uint32_t* work,
                                                      uint32_t* work_count, bool stamp) {
  const Src src = with_table(src_, tab);
  const uint32_t scene = blockIdx.y;
  const View& V = src.view(scene);
  const uint32_t n = V.ctr->block_hwm;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t p = base + threadIdx.x;
    bool take = false;
    if (p < n) {
      const uint64_t k = V.block_keys[p];
      if (k != kEmptyKey)
        take = !frustum || block_in_frustum(V.geo, V.cTw, unpack_key(k), V.cam);
    }
    const uint64_t mask = __ballot(take);
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t out = 0, gout = 0;
    if (lane == 0 && cnt) {
      out = atomicAdd(&V.ctr->cand_count, cnt);
      if (work) gout = atomicAdd(work_count, cnt);
    }
    out = __shfl(out, 0);
    if (take) {
      V.cand[out + prefix] = (int32_t)p;
      V.cand_flag[out + prefix] = 0;  // no recycling flag outlives its worklist
      if (stamp) V.block_mark[p] = tab[scene].mc.pass;
    }
    if (work) {
      gout = __shfl(gout, 0);
      if (take) work[gout + prefix] = (scene << kSceneShift) | p;
    }
  }
}

// ------------------------------------------------------------- TSDF fusion
// UpdateBlocksSimpleKernel (update_simple.cu:13-77) + Voxel::Update
// (voxel.h:107-115).  Persistent workgroups over the worklist, a tiling of
// lanes to blocks chosen by the autotuner, coalesced SoA state.  The pixel
// choice and the skip tests are FP32 decisions; the observed SDF value uses
// an FP64 camera z.
__device__ __forceinline__ uint32_t blend_color(uint32_t c, uint32_t d) {
  uint32_t out = 0;
  for (int k = 0; k < 3; ++k) {
    const float cp = (float)((c >> (8 * k)) & 255u);
    const float cd = (float)((d >> (8 * k)) & 255u);
    const float cc = 0.5f * cp + 0.5f * cd;
    out |= ((uint32_t)(uint8_t)(cc + 0.5f)) << (8 * k);
  }
  return out;
}

// Fuse the observation into the K voxels lt, lt + L, ... of worklist block
// ptr held by one lane; returns how many changed.  Three rounds, so that the
// lane's loads are in flight together: projections and depth reads; the
// update tests and voxel reads; the updates and stores.
template <int K, uint32_t L>
__device__ __forceinline__ uint32_t integrate_voxels(const View& V, const i3 b, int32_t ptr, uint32_t lt) {
  const Geo& g = V.geo;
  const Cam& c = V.cam;
  uint32_t pix[K];
  bool live[K];
  float depth[K], zc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const i3 o = devectorize(lt + (uint32_t)k * L);
    const i3 vp = i3{b.x * kBlockSide + o.x, b.y * kBlockSide + o.y, b.z * kBlockSide + o.z};
    const f3 cp = xform(V.cTw, voxel_to_world(g, vp));
    const uint32_t u = (uint32_t)f2i(proj_u(c, cp) + 0.5f);
    const uint32_t v = (uint32_t)f2i(proj_v(c, cp) + 0.5f);
    live[k] = u < c.width && v < c.height;
    pix[k] = live[k] ? v * c.width + u : 0u;
    zc[k] = cp.z;
    depth[k] = live[k] ? V.depth[pix[k]] : 0.0f;
  }
  float so[K], wo[K], wgt[K], sdf[K];
  uint32_t co[K], dc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float d = depth[k];
    live[k] = live[k] && !(d == -INFINITY || d == 0.0f || d >= g.sdf_upper_bound);
    const float sdf_dec = d - zc[k];
    wgt[k] = fmaxf(10.0f * g.weight_sample * (1.0f - normalize_depth(d, c.min_depth, c.max_depth)), 1.0f);
    const float tr = truncate_distance(g, d);
    live[k] = live[k] && !(sdf_dec <= -tr);
    sdf[k] = 0.0f;
    so[k] = wo[k] = 0.0f;
    co[k] = 0u;
    dc[k] = 255u << 8;  // (0, 255, 0)
    if (!live[k]) continue;
    const i3 o = devectorize(lt + (uint32_t)k * L);
    const i3 vp = i3{b.x * kBlockSide + o.x, b.y * kBlockSide + o.y, b.z * kBlockSide + o.z};
    const double vsd = (double)g.voxel_size;
    const double X = (double)vp.x * vsd, Y = (double)vp.y * vsd, Z = (double)vp.z * vsd;
    const double z64 = fma(V.cTw64.m[8], X, fma(V.cTw64.m[9], Y, fma(V.cTw64.m[10], Z, V.cTw64.m[11])));
    double sv = (double)d - z64;
    const double trd = (double)tr;
    sv = sv > trd ? trd : (sv < -trd ? -trd : sv);
    sdf[k] = (float)sv;
    const size_t idx = (size_t)ptr * kBlockVoxels + lt + (uint32_t)k * L;
    so[k] = V.sdf[idx];
    wo[k] = V.weight[idx];
    co[k] = V.color[idx];
    if (V.has_color) dc[k] = V.color_in[pix[k]];
  }
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!live[k]) continue;
    const size_t idx = (size_t)ptr * kBlockVoxels + lt + (uint32_t)k * L;
    V.sdf[idx] = (so[k] * wo[k] + sdf[k] * wgt[k]) / (wo[k] + wgt[k]);
    V.weight[idx] = wo[k] + wgt[k];
    const uint32_t nc = blend_color(co[k], dc[k]);
    if (nc != co[k]) V.color[idx] = nc;  // a converged colour is left alone (same bits)
    ++n;
  }
  return n;
}

// One workgroup of NT lanes fuses BPW worklist blocks at a time (NT / BPW
// lanes per block, a multiple of the wave size, so every wave works on one
// block of one scene).  Counters are flushed per wave whenever the scene
// changes, so a batch's per-scene counters equal those of separate runs.
template <class Src, int NT, int BPW>
__global__ __launch_bounds__(NT) void integrate_tsdf_kernel(Src src_, const SceneFrame* __restrict__ tab) {
  const Src src = with_table(src_, tab);
  constexpr uint32_t kLanes = NT / BPW;
  static_assert(kLanes % 64 == 0 && kBlockVoxels % kLanes == 0, "tiling");
  const uint32_t n = src.count();
  // kLanes is a multiple of the wave size, so sub is uniform per wave
  const uint32_t sub = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes), lt = threadIdx.x % kLanes;
  uint32_t nupd = 0, nblk = 0, cur = ~0u;
  for (uint32_t ci = blockIdx.x * BPW + sub; ci < n; ci += gridDim.x * BPW) {
    uint32_t scene;
    const int32_t ptr = src.item(ci, &scene);
    if (scene != cur) {
      if (cur != ~0u) {
        const View& P = src.view(cur);
        wave_add_u64(&work_stripe(P.ctr).voxel_updates, nupd);
        wave_add_u64(&work_stripe(P.ctr).blocks_integrated, nblk);
      }
      nupd = 0;
      nblk = 0;
      cur = scene;
    }
    const View V = src.view(scene);  // a copy: read once, not reloaded after stores
    const i3 b = unpack_key(V.block_keys[ptr]);
    nblk += lt == 0;
    nupd += integrate_voxels<kBlockVoxels / kLanes, kLanes>(V, b, ptr, lt);
  }
  if (cur != ~0u) {
    const View& P = src.view(cur);
    wave_add_u64(&work_stripe(P.ctr).voxel_updates, nupd);
    wave_add_u64(&work_stripe(P.ctr).blocks_integrated, nblk);
  }
}

// The same fusion with the latency chain cut short.  A lane's voxel state
// is loaded before the depth test, together with the projection's depth
// gathers, so a block costs one memory round trip instead of two (the state
// of a voxel that then fails the test is read and left alone: its cache
// line is mostly fetched anyway, since the test splits lines).  Each wave
// fetches the worklist items and block keys of its next 64 blocks with one
// vector load each, so no block waits on its own key.  Same arithmetic and
// the same stores as integrate_voxels, so the results are bit-identical.
template <int K, uint32_t L>
__device__ __forceinline__ uint32_t integrate_voxels_early(const View& V, const i3 b, int32_t ptr,
                                                           uint32_t lt) {
  const Geo& g = V.geo;
  const Cam& c = V.cam;
  const size_t base = (size_t)ptr * kBlockVoxels + lt;
  float so[K], wo[K];
  uint32_t co[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    so[k] = V.sdf[base + (uint32_t)k * L];
    wo[k] = V.weight[base + (uint32_t)k * L];
    co[k] = V.color[base + (uint32_t)k * L];
  }
  uint32_t pix[K];
  bool live[K];
  float depth[K], zc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const i3 o = devectorize(lt + (uint32_t)k * L);
    const i3 vp = i3{b.x * kBlockSide + o.x, b.y * kBlockSide + o.y, b.z * kBlockSide + o.z};
    const f3 cp = xform(V.cTw, voxel_to_world(g, vp));
    const uint32_t u = (uint32_t)f2i(proj_u(c, cp) + 0.5f);
    const uint32_t v = (uint32_t)f2i(proj_v(c, cp) + 0.5f);
    live[k] = u < c.width && v < c.height;
    pix[k] = live[k] ? v * c.width + u : 0u;
    zc[k] = cp.z;
    depth[k] = live[k] ? V.depth[pix[k]] : 0.0f;
  }
  float wgt[K], sdf[K];
  uint32_t dc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float d = depth[k];
    live[k] = live[k] && !(d == -INFINITY || d == 0.0f || d >= g.sdf_upper_bound);
    const float sdf_dec = d - zc[k];
    wgt[k] = fmaxf(10.0f * g.weight_sample * (1.0f - normalize_depth(d, c.min_depth, c.max_depth)), 1.0f);
    const float tr = truncate_distance(g, d);
    live[k] = live[k] && !(sdf_dec <= -tr);
    sdf[k] = 0.0f;
    dc[k] = 255u << 8;  // (0, 255, 0)
    if (!live[k]) continue;
    const i3 o = devectorize(lt + (uint32_t)k * L);
    const i3 vp = i3{b.x * kBlockSide + o.x, b.y * kBlockSide + o.y, b.z * kBlockSide + o.z};
    const double vsd = (double)g.voxel_size;
    const double X = (double)vp.x * vsd, Y = (double)vp.y * vsd, Z = (double)vp.z * vsd;
    const double z64 = fma(V.cTw64.m[8], X, fma(V.cTw64.m[9], Y, fma(V.cTw64.m[10], Z, V.cTw64.m[11])));
    double sv = (double)d - z64;
    const double trd = (double)tr;
    sv = sv > trd ? trd : (sv < -trd ? -trd : sv);
    sdf[k] = (float)sv;
    if (V.has_color) dc[k] = V.color_in[pix[k]];
  }
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!live[k]) continue;
    const size_t idx = base + (uint32_t)k * L;
    V.sdf[idx] = (so[k] * wo[k] + sdf[k] * wgt[k]) / (wo[k] + wgt[k]);
    V.weight[idx] = wo[k] + wgt[k];
    const uint32_t nc = blend_color(co[k], dc[k]);
    if (nc != co[k]) V.color[idx] = nc;  // a converged colour is left alone (same bits)
    ++n;
  }
  return n;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, uint32_t j) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, j);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), j);
  return ((uint64_t)hi << 32) | lo;
}

template <class Src, int NT, int BPW>
__global__ __launch_bounds__(NT) void integrate_tsdf_early_kernel(Src src_, const SceneFrame* __restrict__ tab) {
  const Src src = with_table(src_, tab);
  constexpr uint32_t kLanes = NT / BPW;
  static_assert(kLanes % 64 == 0 && kBlockVoxels % kLanes == 0, "tiling");
  const uint32_t n = src.count();
  const uint32_t sub = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes), lt = threadIdx.x % kLanes;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t stride = gridDim.x * BPW;
  uint32_t nupd = 0, nblk = 0, cur = ~0u;
  for (uint32_t c0 = blockIdx.x * BPW + sub; c0 < n; c0 += 64u * stride) {
    // lane j holds this wave's j-th next item: scene, pool index and key
    const uint32_t cj = c0 + lane * stride;
    uint32_t sj = 0;
    int32_t pj = 0;
    uint64_t kj = 0;
    if (cj < n) {
      pj = src.lane_item(cj, &sj);
      kj = src.view(sj).block_keys[pj];
    }
    const uint32_t m = min(64u, (n - c0 + stride - 1u) / stride);
    for (uint32_t j = 0; j < m; ++j) {
      const uint32_t scene = __builtin_amdgcn_readlane(sj, j);
      const int32_t ptr = __builtin_amdgcn_readlane(pj, j);
      const i3 b = unpack_key(readlane_u64(kj, j));
      if (scene != cur) {
        if (cur != ~0u) {
          const View& P = src.view(cur);
          wave_add_u64(&work_stripe(P.ctr).voxel_updates, nupd);
          wave_add_u64(&work_stripe(P.ctr).blocks_integrated, nblk);
        }
        nupd = 0;
        nblk = 0;
        cur = scene;
      }
      const View V = src.view(scene);
      nblk += lt == 0;
      nupd += integrate_voxels_early<kBlockVoxels / kLanes, kLanes>(V, b, ptr, lt);
    }
  }
  if (cur != ~0u) {
    const View& P = src.view(cur);
    wave_add_u64(&work_stripe(P.ctr).voxel_updates, nupd);
    wave_add_u64(&work_stripe(P.ctr).blocks_integrated, nblk);
  }
}

// Save / restore the fused state of the worklist blocks (dense, worklist order).
__global__ __launch_bounds__(256) void state_copy_kernel(View V, float* sdf, float* weight,
                                                         uint32_t* color, int restore) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const size_t dst = (size_t)V.cand[ci] * kBlockVoxels, buf = (size_t)ci * kBlockVoxels;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      if (restore) {
        V.sdf[dst + i] = sdf[buf + i];
        V.weight[dst + i] = weight[buf + i];
        V.color[dst + i] = color[buf + i];
      } else {
        sdf[buf + i] = V.sdf[dst + i];
        weight[buf + i] = V.weight[dst + i];
        color[buf + i] = V.color[dst + i];
      }
    }
  }
}

// ------------------------------------------------------------- PSDF fusion
// gaussian(x, u, s2) = exp(-(x-u)^2/s2) / (sqrt(2 pi) sqrt(s2)), note: no
// factor 1/2 in the exponent (geometry/geometry_helper.h:30-34).
__device__ __forceinline__ float gaussian_psdf(float x, float u, float s2) {
  const float d = x - u;
  return expf(-(d * d) / s2) / (2.5066282746310002f * sqrtf(s2));
}

// PredictOutlierRatioKernel, update_bayesian.cu:36-101.  Every voxel of the
// worklist that owns a vertex of the previous mesh scores the pixel it
// projects to; the pixel keeps the maximum.  Scores are positive floats, so
// an integer atomicMax on their bits is an exact, order-free float max.
__global__ __launch_bounds__(256) void predict_inlier_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  const Geo g = V.geo;
  const Cam c = V.cam;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const i3 b = unpack_key(V.block_keys[ptr]);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      const int32_t p0 = V.vptr[3 * idx + 0], p1 = V.vptr[3 * idx + 1], p2 = V.vptr[3 * idx + 2];
      if (p0 <= 0 && p1 <= 0 && p2 <= 0) continue;
      const i3 o = devectorize(i);
      const i3 vp = i3{b.x * kBlockSide + o.x, b.y * kBlockSide + o.y, b.z * kBlockSide + o.z};
      const f3 cp = xform(V.cTw, voxel_to_world(g, vp));
      const uint32_t u = (uint32_t)f2i(proj_u(c, cp) + 0.5f);
      const uint32_t v = (uint32_t)f2i(proj_v(c, cp) + 0.5f);
      if (u >= c.width || v >= c.height) continue;
      const uint32_t pix = u + v * c.width;
      const float depth = V.depth[pix];
      if (depth == -INFINITY || depth == 0.0f || depth >= g.sdf_upper_bound) continue;
      const f3 pc = reproject(c, u, v, depth);
      const float vsdf = V.sdf[idx];
      const float vis2 = V.weight[idx];
      const int32_t ps[3] = {p0, p1, p2};
      for (int k = 0; k < 3; ++k) {
        if (ps[k] <= 0) continue;
        const Vertex vt = V.vertices[ps[k]];
        const f3 nrm = mk3(vt.nx, vt.ny, vt.nz);
        const f3 x = mk3(vt.px, vt.py, vt.pz);
        // The normal goes through the affine transform, translation included
        // (update_bayesian.cu:78,84).
        const f3 nc = normalize(xform(V.cTw, nrm));
        const float cos_a = dot(normalize(-pc), nc);
        const f3 vec = pc - xform(V.cTw, x);
        const float pd = dot(vec, nc);
        const float disk = sqrtf(dot(vec, vec) - pd * pd);
        const float dd = pd - vsdf;
        const float w_dist = expf(-(dd * dd) * vis2);
        const float w_disk = 0.5f + 0.5f / (1.0f + expf(-fabsf(disk / vt.radius)));
        const float w_ang = fmaxf((cos_a - 0.1736f) / (1.0f - 0.1736f), 0.1f);
        const float r = fmaxf(w_disk * w_dist * w_ang, 0.01f);
        atomicMax((int*)&V.inlier[pix], __float_as_int(r));
      }
    }
  }
}

// UpdateBlocksBayesianKernel, update_bayesian.cu:103-185.
__global__ __launch_bounds__(512) void integrate_psdf_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  const Geo g = V.geo;
  const Cam c = V.cam;
  const double vsd = (double)g.voxel_size;
  uint32_t nupd = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&work_stripe(V.ctr).blocks_integrated, (unsigned long long)n);
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t ptr = V.cand[ci];
    const i3 b = unpack_key(V.block_keys[ptr]);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const i3 o = devectorize(i);
      const i3 vp = i3{b.x * kBlockSide + o.x, b.y * kBlockSide + o.y, b.z * kBlockSide + o.z};
      const f3 cp = xform(V.cTw, voxel_to_world(g, vp));
      const uint32_t u = (uint32_t)f2i(proj_u(c, cp) + 0.5f);
      const uint32_t v = (uint32_t)f2i(proj_v(c, cp) + 0.5f);
      if (u >= c.width || v >= c.height) continue;
      const uint32_t pix = u + v * c.width;
      const float depth = V.depth[pix];
      if (depth == -INFINITY || depth == 0.0f || depth >= g.sdf_upper_bound) continue;
      const float x_dec = depth - cp.z;
      const float rho = V.inlier[pix];
      const float tr = truncate_distance(g, depth);
      if (x_dec <= -tr) continue;
      ++nupd;
      const double X = (double)vp.x * vsd, Y = (double)vp.y * vsd, Z = (double)vp.z * vsd;
      const double z64 = fma(V.cTw64.m[8], X, fma(V.cTw64.m[9], Y, fma(V.cTw64.m[10], Z, V.cTw64.m[11])));
      double s = (double)depth - z64;
      const double trd = (double)tr;
      s = s > trd ? trd : (s < -trd ? -trd : s);
      const float x = (float)s;

      const float tau = (depth - 0.4f) * 0.012f + 0.019f;
      const size_t idx = (size_t)ptr * kBlockVoxels + i;
      const float is2 = V.weight[idx];
      if (is2 == 0.0f) {
        V.sdf[idx] = x;
        V.weight[idx] = 1.0f / (tau * tau);
        V.va[idx] = rho * 10.0f;
        V.vb[idx] = (1.0f - rho) * 10.0f;
        continue;
      }
      const float mu = V.sdf[idx];
      const float s2 = 1.0f / is2;
      const float t2 = tau * tau;
      const float ss2 = 1.0f / (1.0f / s2 + 1.0f / t2);
      const float m = ss2 * (mu / s2 + x / t2);
      float C1 = rho * gaussian_psdf(x, mu, s2 + t2);
      float C2 = (depth < c.max_depth && depth >= c.min_depth) ? (1.0f - rho) / (5.0f - 0.1f) : 1.0f;
      const float csum = C1 + C2;
      C1 /= csum;
      C2 /= csum;
      const float a = V.va[idx], bb = V.vb[idx];
      const float ab1 = a + bb + 1.0f, ab2 = a + bb + 2.0f;
      const float f = C1 * (a + 1.0f) / ab1 + C2 * a / ab1;
      const float e = C1 * (a + 1.0f) * (a + 2.0f) / (ab1 * ab2) + C2 * a * (a + 1.0f) / (ab1 * ab2);
      const float nsdf = C1 * m + C2 * mu;
      V.sdf[idx] = nsdf;
      V.weight[idx] = 1.0f / (C1 * (ss2 + m * m) + C2 * (s2 + mu * mu) - nsdf * nsdf);
      const float na = (e - f) / (f - e / f);
      V.va[idx] = na;
      V.vb[idx] = na * (1.0f - f) / f;
    }
  }
  wave_add_u64(&work_stripe(V.ctr).voxel_updates, nupd);
}

// ------------------------------------------------------------ block restore
// Insert explicit block keys with the allocation kernel's insertion rule,
// then copy their voxel fields in (checkpoint restore, parity fixtures).
__global__ void load_insert_kernel(View V, const uint64_t* keys, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    insert_block(V, unpack_key(keys[i]));
}

__global__ __launch_bounds__(256) void load_scatter_kernel(View V, const uint64_t* keys, uint32_t n,
                                                           const float* sdf, const float* weight,
                                                           const uint32_t* color, const float* a,
                                                           const float* b) {
  for (uint32_t blk = blockIdx.x; blk < n; blk += gridDim.x) {
    const int32_t p = hash_find(V, unpack_key(keys[blk]));
    if (p < 0) continue;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const size_t src = (size_t)blk * kBlockVoxels + i;
      const size_t dst = (size_t)p * kBlockVoxels + i;
      V.sdf[dst] = sdf[src];
      V.weight[dst] = weight[src];
      if (color) V.color[dst] = color[src];
      if (a && V.va) V.va[dst] = a[src];
      if (b && V.vb) V.vb[dst] = b[src];
    }
  }
}

void launch_load_insert(const View& V, const uint64_t* keys, uint32_t n, hipStream_t s) {
  load_insert_kernel<<<(n + 255) / 256, 256, 0, s>>>(V, keys, n);
  list_insert_kernel<One, false><<<1, kListChunk, 0, s>>>(One{V}, nullptr);
}

void launch_load_scatter(const View& V, const uint64_t* keys, uint32_t n, const float* sdf,
                         const float* weight, const uint32_t* color, const float* a,
                         const float* b, hipStream_t s) {
  load_scatter_kernel<<<n < 4096u ? n : 4096u, 256, 0, s>>>(V, keys, n, sdf, weight, color, a, b);
}

void launch_state_save(const View& V, float* sdf, float* weight, uint32_t* color, hipStream_t s,
                       int grid) {
  state_copy_kernel<<<grid, 256, 0, s>>>(V, sdf, weight, color, 0);
}

void launch_state_restore(const View& V, const float* sdf, const float* weight,
                          const uint32_t* color, hipStream_t s, int grid) {
  state_copy_kernel<<<grid, 256, 0, s>>>(V, const_cast<float*>(sdf), const_cast<float*>(weight),
                                         const_cast<uint32_t*>(color), 1);
}

// ------------------------------------------------------------ block recycling
// CollectLowSurfelBlocksKernel (mapping/recycle.cu:81-115) in two passes: a
// block with fewer than 10 inner and no boundary surfels ages by one frame,
// any other is renewed; then a block whose life is over is flagged when no
// allocated face neighbour is still alive.  (The reference ages and reads
// the neighbours in one launch, so a neighbour may be read before or after
// its own update; here every block reads the updated values.)
__global__ void gc_life_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t p = V.cand[i];
    const bool low = V.surfel[2 * p] < 10 && V.surfel[2 * p + 1] == 0;
    V.life[p] = low ? V.life[p] - 1 : kBlockLife;
  }
}

__global__ void gc_low_surfel_flag_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t p = V.cand[i];
    uint8_t f = 0;
    if (V.life[p] <= 0) {
      f = 1;
      const i3 b = unpack_key(V.block_keys[p]);
      const i3 off[6] = {{0, 0, 1}, {0, 0, -1}, {0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}};
      for (int j = 0; j < 6; ++j) {
        const int32_t q = hash_find(V, i3{b.x + off[j].x, b.y + off[j].y, b.z + off[j].z});
        if (q >= 0 && V.life[q] != 0) f = 0;
      }
    }
    V.cand_flag[i] = f;
  }
}

// StarveOccupiedBlocksKernel (recycle.cu:15-24): weight -= 1, floored at 0.
__global__ __launch_bounds__(256) void gc_starve_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const size_t base = (size_t)V.cand[ci] * kBlockVoxels;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x)
      V.weight[base + i] = fmaxf(0.0f, V.weight[base + i] - 1.0f);
  }
}

// CollectGarbageBlockArrayKernel (recycle.cu:27-77): garbage if every
// observed voxel (weight >= 1e-6, the double EPSILON) has |sdf| >= the
// truncation at 5 m, or no voxel is observed.
__global__ __launch_bounds__(256) void gc_flag_garbage_kernel(View V) {
  __shared__ float smin[4], smax[4];
  const uint32_t n = V.ctr->cand_count;
  const float t = truncate_distance(V.geo, 5.0f);
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const size_t base = (size_t)V.cand[ci] * kBlockVoxels;
    float mn = INFINITY, mx = 0.0f;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      const float w = V.weight[base + i];
      const float a = (double)w < 1e-6 ? INFINITY : fabsf(V.sdf[base + i]);
      mn = fminf(mn, a);
      mx = fmaxf(mx, w);
    }
    for (int off = 32; off > 0; off >>= 1) {
      mn = fminf(mn, __shfl_xor(mn, off));
      mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63u) == 0) {
      smin[threadIdx.x >> 6] = mn;
      smax[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float a = smin[0], b = smax[0];
      for (int w = 1; w < 4; ++w) {
        a = fminf(a, smin[w]);
        b = fmaxf(b, smax[w]);
      }
      V.cand_flag[ci] = (a >= t || (double)b < 1e-6) ? 1 : 0;
    }
    __syncthreads();
  }
}

__global__ void gc_begin_kernel(View V) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    V.ctr->gc_count = 0;
    V.ctr->gc_free_count = 0;
  }
}

// Flagged worklist entries -> the garbage list.
__global__ __launch_bounds__(256) void gc_list_kernel(View V) {
  const uint32_t n = V.ctr->cand_count;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool take = i < n && V.cand_flag[i];
    const uint64_t mask = __ballot(take);
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t out = 0;
    if (lane == 0 && cnt) out = atomicAdd(&V.ctr->gc_count, cnt);
    out = __shfl(out, 0);
    if (take) V.gc_list[out + prefix] = V.cand[i];
  }
}

// Garbage blocks left with no allocated vertex slot are to be freed
// (RecycleGarbageVerticesKernel's valid_vertex_count == 0, recycle.cu:169-178).
__global__ __launch_bounds__(256) void gc_count_valid_kernel(View V) {
  __shared__ int any;
  const uint32_t n = V.ctr->gc_count;
  for (uint32_t ci = blockIdx.x; ci < n; ci += gridDim.x) {
    const int32_t p = V.gc_list[ci];
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    bool has = false;
    for (uint32_t k = threadIdx.x; k < 3u * kBlockVoxels; k += blockDim.x)
      has = has || V.vptr[(size_t)p * 3 * kBlockVoxels + k] >= 0;
    if (has) any = 1;
    __syncthreads();
    if (threadIdx.x == 0 && !any) V.gc_free[atomicAdd(&V.ctr->gc_free_count, 1u)] = p;
    __syncthreads();
  }
}

// HashTable::FreeEntry (core/hash_table.h:183-255, the linked-list variant):
// a bucket slot heading a list takes over the list's first node; a list
// node is unlinked from its predecessor.  Runs on one lane, so no bucket
// lock is needed.
__device__ bool free_entry(const View& V, uint64_t key) {
  if (key == kEmptyKey) return false;
  const i3 p = unpack_key(key);
  const uint32_t first = hash_bucket(p, V.bucket_count) * V.bucket_size;
  const uint32_t last = first + V.bucket_size - 1;
  const Entry empty{kEmptyKey, kFreePtr, 0u};
  for (uint32_t j = 0; j < V.bucket_size; ++j) {
    const uint32_t i = first + j;
    const Entry e = V.entries[i];
    if (e.key != key) continue;
    if (e.offset != 0) {
      const uint32_t nx = (i + e.offset) % V.entry_count;
      V.entries[i] = V.entries[nx];
      V.entries[nx] = empty;
    } else {
      V.entries[i] = empty;
    }
    return true;
  }
  if (V.entries[last].offset == 0) return false;
  uint32_t prev = last, i = (last + V.entries[last].offset) % V.entry_count;
  for (uint32_t it = 0; it < V.linked_list_size; ++it) {
    const Entry e = V.entries[i];
    if (e.key == key) {
      V.entries[i] = empty;
      V.entries[prev].offset = e.offset;
      return true;
    }
    if (e.offset == 0) return false;
    prev = i;
    i = (last + e.offset) % V.entry_count;
  }
  return false;
}

// Frees the entries of the garbage list's empty blocks in list order, returns
// their pool slots to the heap (Free, hash_table.h:288-292) and keeps only
// the freed blocks on gc_free.  A slot already free (its key cleared) is
// skipped, so the heap never takes a slot twice.
__global__ void gc_free_entries_kernel(View V) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t n = V.ctr->gc_free_count;
  uint32_t kept = 0;
  for (uint32_t q = 0; q < n; ++q) {
    const int32_t p = V.gc_free[q];
    if (V.ctr->block_top + 1 >= (int32_t)V.capacity) break;
    if (!free_entry(V, V.block_keys[p])) continue;
    V.block_keys[p] = kEmptyKey;
    const int32_t top = V.ctr->block_top + 1;
    V.ctr->block_top = top;
    V.block_heap[top] = p;
    V.gc_free[kept++] = p;
  }
  V.ctr->gc_free_count = kept;
}

// Block::Clear (core/block.h:24-39) for the freed blocks.
__global__ __launch_bounds__(256) void gc_clear_kernel(View V) {
  const uint32_t n = V.ctr->gc_free_count;
  for (uint32_t q = blockIdx.x; q < n; q += gridDim.x) {
    const int32_t p = V.gc_free[q];
    const size_t base = (size_t)p * kBlockVoxels;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBlockVoxels; i += blockDim.x) {
      V.sdf[base + i] = 0.0f;
      V.weight[base + i] = 0.0f;
      V.color[base + i] = 0u;
      if (V.va) {
        V.va[base + i] = 0.0f;
        V.vb[base + i] = 0.0f;
      }
      V.vptr[3 * (base + i) + 0] = kFreePtr;
      V.vptr[3 * (base + i) + 1] = kFreePtr;
      V.vptr[3 * (base + i) + 2] = kFreePtr;
      V.cube[base + i] = 0;
    }
    if (threadIdx.x == 0) {
      V.surfel[2 * p] = 0;
      V.surfel[2 * p + 1] = 0;
      V.life[p] = kBlockLife;
      V.block_mark[p] = 0u;
    }
  }
}

void launch_gc_low_surfel(const View& V, hipStream_t s, int grid) {
  gc_life_kernel<<<grid, 256, 0, s>>>(V);
  gc_low_surfel_flag_kernel<<<grid, 256, 0, s>>>(V);
}

void launch_gc_starve(const View& V, hipStream_t s, int grid) {
  gc_starve_kernel<<<grid, 256, 0, s>>>(V);
}

void launch_gc_flag_garbage(const View& V, hipStream_t s, int grid) {
  gc_flag_garbage_kernel<<<grid, 256, 0, s>>>(V);
}

// RecycleGarbageBlockArray (recycle.cu:231-260) over the flagged worklist.
void launch_gc_recycle(const View& V, hipStream_t s, int cus) {
  const int grid = cus * 8;
  gc_begin_kernel<<<1, 64, 0, s>>>(V);
  gc_list_kernel<<<grid, 256, 0, s>>>(V);
  launch_gc_release_triangles(V, s, grid);
  launch_gc_recycle_vertices(V, s, cus);
  gc_count_valid_kernel<<<grid, 256, 0, s>>>(V);
  gc_free_entries_kernel<<<1, 64, 0, s>>>(V);
  gc_clear_kernel<<<grid, 256, 0, s>>>(V);
}

// ------------------------------------------------------------------ launchers
void launch_reset(const View& V, hipStream_t s, int grid) {
  reset_entries_kernel<<<grid, 256, 0, s>>>(V.entries, V.entry_count);
  reset_pool_kernel<<<grid, 256, 0, s>>>(V);
  reset_voxels_kernel<<<grid * 4, 256, 0, s>>>(V, (size_t)V.capacity * kBlockVoxels);
  reset_counters_kernel<<<1, 256, 0, s>>>(V);
}

void launch_convert_depth(const View& V, const uint16_t* raw, hipStream_t s) {
  const uint32_t n = V.cam.width * V.cam.height;
  convert_depth_kernel<<<(n + 255) / 256, 256, 0, s>>>(raw, V.depth, V.inlier, n, V.range_factor,
                                                       V.cam.min_depth, V.cam.max_depth);
}

void launch_convert_color(const View& V, const uint8_t* bgra, hipStream_t s) {
  const uint32_t n = V.cam.width * V.cam.height;
  convert_color_kernel<<<(n + 255) / 256, 256, 0, s>>>(reinterpret_cast<const uint32_t*>(bgra),
                                                       V.color_in, n);
}

void launch_alloc(const View& V, hipStream_t s, const KernelClock* kc) {
  dim3 grid((V.cam.width + 15) / 16, (V.cam.height + 15) / 16);
  tick(kc, s, "alloc_kernel");
  alloc_kernel<<<grid, 256, 0, s>>>(One{V}, nullptr);
  tick(kc, s, "list_insert_kernel");
  list_insert_kernel<<<1, kListChunk, 0, s>>>(One{V}, nullptr);
  tick(kc, s, nullptr);
}

void launch_collect(const View& V, int frustum, hipStream_t s, int grid, const KernelClock* kc,
                    bool count_cleared) {
  if (!count_cleared) {  // the allocation pass clears it; otherwise here
    tick(kc, s, "set_u32_kernel");
    set_u32_kernel<<<1, 64, 0, s>>>(&V.ctr->cand_count, 0u);
  }
  tick(kc, s, "collect_kernel");
  collect_kernel<<<grid, 256, 0, s>>>(One{V}, nullptr, frustum, nullptr, nullptr, false);
  tick(kc, s, nullptr);
}

namespace {
template <class Src, int T = 0>
void integrate_dispatch(const Src& src, hipStream_t s, int grid, int tiling) {
  if constexpr (T < kNumIntegrateTilings) {
    constexpr int NT = kIntegrateTilings[T].nt, BPW = kIntegrateTilings[T].bpw;
    if (tiling == T) {
      if constexpr (kIntegrateTilings[T].early)
        integrate_tsdf_early_kernel<Src, NT, BPW><<<grid, NT, 0, s>>>(src, table_of(src));
      else
        integrate_tsdf_kernel<Src, NT, BPW><<<grid, NT, 0, s>>>(src, table_of(src));
    } else
      integrate_dispatch<Src, T + 1>(src, s, grid, tiling);
  }
}
}  // namespace

void launch_integrate_tsdf(const View& V, hipStream_t s, int grid, int tiling, const KernelClock* kc) {
  tick(kc, s, "integrate_tsdf_kernel");
  integrate_dispatch(One{V}, s, grid, tiling);
  tick(kc, s, nullptr);
}

void launch_predict_inlier(const View& V, hipStream_t s, int grid) {
  predict_inlier_kernel<<<grid, 256, 0, s>>>(V);
}

void launch_integrate_psdf(const View& V, hipStream_t s, int grid, int lanes) {
  integrate_psdf_kernel<<<grid, lanes, 0, s>>>(V);
}

// ------------------------------------------------------------ batched launchers
// Per-scene worklist counters, the concatenated worklist's count and the
// overflow queues are cleared by kernels (a captured graph holds no memsets).
__global__ void batch_reset_counts_kernel(Many src_, const SceneFrame* __restrict__ tab, uint32_t nscenes, uint32_t* work_count) {
  const Many src = with_table(src_, tab);
  for (uint32_t t = threadIdx.x; t < nscenes; t += blockDim.x) src.view(t).ctr->cand_count = 0u;
  if (threadIdx.x == 0) *work_count = 0u;
}

void launch_batch_alloc(const BatchDev& B, hipStream_t s) {
  dim3 grid((B.width + 15) / 16, (B.height + 15) / 16, B.nscenes);
  alloc_kernel<<<grid, 256, 0, s>>>(B.src(), B.frames);
  list_insert_kernel<<<B.nscenes, kListChunk, 0, s>>>(B.src(), B.frames, B.work_count);
}

void launch_batch_collect(const BatchDev& B, int frustum, hipStream_t s, int grid, bool counts_cleared,
                          bool stamp_mc) {
  if (!counts_cleared) batch_reset_counts_kernel<<<1, 256, 0, s>>>(B.src(), B.frames, B.nscenes, B.work_count);
  const uint32_t gx = (uint32_t)grid / B.nscenes > 0 ? (uint32_t)grid / B.nscenes : 1u;
  collect_kernel<<<dim3(gx, B.nscenes), 256, 0, s>>>(B.src(), B.frames, frustum, B.work, B.work_count, stamp_mc);
}

void launch_batch_integrate_tsdf(const BatchDev& B, hipStream_t s, int grid, int tiling) {
  integrate_dispatch(B.src(), s, grid, tiling);
}

}  // namespace mh
