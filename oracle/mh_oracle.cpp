// mh_oracle.cpp -- scalar CPU restatement of MeshHashing's per-frame hot
// path, used ONLY as the checker of the HIP library (tests, smoke(), and the
// cpu_baseline leg of bench.py).  It is never linked into or called by the
// product path.
//
// Arithmetic: every discrete decision (pixel choice, DDA step, frustum test,
// voxel rounding, cube sign, regula-falsi branch) is evaluated in FP32 with
// the reference's operation order and no FMA contraction (build with
// -ffp-contract=off).  Every accumulated value (SDF, weights, PSDF
// parameters, inlier ratios, normals, primal-dual state) is FP64.  The
// raycaster, whose output is mostly decisions, runs entirely in FP32 on the
// FP32-rounded SDF.
//
// Parity pinning of each part is listed in oracle/README.md; functions are
// annotated with the reference file:line they restate (paths relative to the
// reference repository root, src/ omitted).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/meshhashing.h"

namespace {

constexpr int kFree = -2;

// ---------------------------------------------------------------- FP32 math
struct F3 {
  float x, y, z;
};
struct I3 {
  int x, y, z;
  bool operator<(const I3& o) const {
    if (x != o.x) return x < o.x;
    if (y != o.y) return y < o.y;
    return z < o.z;
  }
  bool operator==(const I3& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct I3Hash {
  size_t operator()(const I3& p) const {
    return ((size_t)(uint32_t)p.x * 1000003u) ^ ((size_t)(uint32_t)p.y * 998244353u) ^
           ((size_t)(uint32_t)p.z * 19260817u);
  }
};

// helper_math.h:1464-1467: sign(s) = s < 0 ? -1 : 1.
int sgn(float s) { return s < 0.0f ? -1 : 1; }
// CUDA / AMDGPU float->int: truncation, saturating, NaN -> 0.
int f2i(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return -2147483647 - 1;
  return (int)f;
}
float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// helper_math normalize = v * rsqrt(dot(v,v)); rsqrt is taken as the
// correctly rounded 1/sqrtf here (the library does the same).
F3 normalizef(F3 v) {
  const float s = 1.0f / std::sqrt(dotf(v, v));
  return F3{v.x * s, v.y * s, v.z * s};
}
// matrix.h:919-926: float4x4 * float3 (w = 1), row-major m11..m44.
F3 xf(const float* m, F3 p) {
  return F3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3] * 1.0f,
            m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7] * 1.0f,
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11] * 1.0f};
}
// matrix.h:908-916 with w = 0.
F3 xfd(const float* m, F3 d) {
  return F3{m[0] * d.x + m[1] * d.y + m[2] * d.z + m[3] * 0.0f,
            m[4] * d.x + m[5] * d.y + m[6] * d.z + m[7] * 0.0f,
            m[8] * d.x + m[9] * d.y + m[10] * d.z + m[11] * 0.0f};
}

// General 4x4 inverse through 2x2 minors; FP32 for decisions, FP64 for
// values (the reference: matrix.h:964-1089, rgbd_sensor.h:51-54).
template <typename T>
void inv4(const T* a, T* r) {
  const T s0 = a[0] * a[5] - a[4] * a[1];
  const T s1 = a[0] * a[6] - a[4] * a[2];
  const T s2 = a[0] * a[7] - a[4] * a[3];
  const T s3 = a[1] * a[6] - a[5] * a[2];
  const T s4 = a[1] * a[7] - a[5] * a[3];
  const T s5 = a[2] * a[7] - a[6] * a[3];
  const T c5 = a[10] * a[15] - a[14] * a[11];
  const T c4 = a[9] * a[15] - a[13] * a[11];
  const T c3 = a[9] * a[14] - a[13] * a[10];
  const T c2 = a[8] * a[15] - a[12] * a[11];
  const T c1 = a[8] * a[14] - a[12] * a[10];
  const T c0 = a[8] * a[13] - a[12] * a[9];
  const T det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  const T id = T(1) / det;
  r[0] = (a[5] * c5 - a[6] * c4 + a[7] * c3) * id;
  r[1] = (-a[1] * c5 + a[2] * c4 - a[3] * c3) * id;
  r[2] = (a[13] * s5 - a[14] * s4 + a[15] * s3) * id;
  r[3] = (-a[9] * s5 + a[10] * s4 - a[11] * s3) * id;
  r[4] = (-a[4] * c5 + a[6] * c2 - a[7] * c1) * id;
  r[5] = (a[0] * c5 - a[2] * c2 + a[3] * c1) * id;
  r[6] = (-a[12] * s5 + a[14] * s2 - a[15] * s1) * id;
  r[7] = (a[8] * s5 - a[10] * s2 + a[11] * s1) * id;
  r[8] = (a[4] * c4 - a[5] * c2 + a[7] * c0) * id;
  r[9] = (-a[0] * c4 + a[1] * c2 - a[3] * c0) * id;
  r[10] = (a[12] * s4 - a[13] * s2 + a[15] * s0) * id;
  r[11] = (-a[8] * s4 + a[9] * s2 - a[11] * s0) * id;
  r[12] = (-a[4] * c3 + a[5] * c1 - a[6] * c0) * id;
  r[13] = (a[0] * c3 - a[1] * c1 + a[2] * c0) * id;
  r[14] = (-a[12] * s3 + a[13] * s1 - a[14] * s0) * id;
  r[15] = (a[8] * s3 - a[9] * s1 + a[10] * s0) * id;
}

// ------------------------------------------------------------ geometry (FP32)
// geometry/geometry_helper.h:36-217.
struct Geo {
  float vs, tr, tr_scale, ub, ws;
  float truncate(float z) const { return tr + tr_scale * z; }
  int w2v1(float w) const {
    const float p = w / vs;
    return f2i(p + (float)sgn(p) * 0.5f);
  }
  I3 w2v(F3 w) const { return I3{w2v1(w.x), w2v1(w.y), w2v1(w.z)}; }
  static int v2b1(int v) { return (v - (v < 0) * 7) / 8; }
  static I3 v2b(I3 v) { return I3{v2b1(v.x), v2b1(v.y), v2b1(v.z)}; }
  F3 v2w(I3 v) const { return F3{(float)v.x * vs, (float)v.y * vs, (float)v.z * vs}; }
  F3 b2w(I3 b) const { return v2w(I3{b.x * 8, b.y * 8, b.z * 8}); }
};
struct Cam {
  float fx, fy, cx, cy;
  uint32_t w, h;
  float dmin, dmax;
  float pu(F3 p) const { return p.x * fx / p.z + cx; }
  float pv(F3 p) const { return p.y * fy / p.z + cy; }
  F3 reproj(uint32_t ux, uint32_t uy, float d) const {
    const float x = ((float)ux - cx) / fx;
    const float y = ((float)uy - cy) / fy;
    return F3{d * x, d * y, d};
  }
};
float normdepth(float z, float mn, float mx) { return (z - mn) / (mx - mn); }

// IsPointInCameraFrustum / IsBlockInCameraFrustum (geometry_helper.h:186-216).
bool block_in_frustum(const Geo& g, const float* cTw, I3 b, const Cam& c) {
  const F3 w0 = g.b2w(b);
  const float h = g.vs * 0.5f * (8.0f - 1.0f);
  const F3 p = xf(cTw, F3{w0.x + h, w0.y + h, w0.z + h});
  const float u = c.pu(p), v = c.pv(p);
  const float wm1 = (float)c.w - 1.0f, hm1 = (float)c.h - 1.0f;
  float nx = (2.0f * u - wm1) / wm1, ny = (hm1 - 2.0f * v) / hm1, nz = normdepth(p.z, c.dmin, c.dmax);
  nx *= 0.95f;
  ny *= 0.95f;
  nz *= 0.95f;
  return !(nx < -1.0f || nx > 1.0f || ny < -1.0f || ny > 1.0f || nz < 0.0f || nz > 1.0f);
}

uint32_t bucket_of(I3 p, uint32_t n) {
  // hash_table.h:259-269: int products wrap in 32 bits; the modulo is unsigned.
  const uint32_t h = ((uint32_t)p.x * 73856093u) ^ ((uint32_t)p.y * 19349669u) ^
                     ((uint32_t)p.z * 83492791u);
  return h % n;
}

// ---------------------------------------------------------------- MC tables
// Bourke's triangulation table (the table of meshing/mc_tables.h:16-332),
// one string per configuration, each character an edge index.
const char* const kTri[256] = {
#include "mc_table.inc"
};
// Corner numbering (mc_tables.h:334-344) as per-axis bit masks.
int corner(int i, int ax) { return ((ax == 0 ? 0x66 : ax == 1 ? 0x0F : 0x33) >> i) & 1; }
// Edge endpoints (mc_tables.h:375-389).
int ea(int e) { return e < 8 ? e : e - 4; }
int eb(int e) { return e < 4 ? (e + 1) & 3 : (e < 8 ? 4 + ((e - 3) & 3) : e - 8); }
int hexv(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }
int ntri(int cube) { return (int)std::strlen(kTri[cube]) / 3; }
int tri_edge(int cube, int k) { return hexv(kTri[cube][k]); }
unsigned edge_mask(int cube) {
  unsigned m = 0;
  for (const char* s = kTri[cube]; *s; ++s) m |= 1u << hexv(*s);
  return m;
}

// ---------------------------------------------------------------- state
struct Block {
  I3 pos;
  double sdf[512];
  double w[512];
  double a[512];
  double b[512];
  uint8_t col[512][3];
  uint16_t cube[512];  // curr | prev << 8
  // primal-dual (core/voxel.h:62-82)
  double sdf0[512], sdf_bar[512], wpd[512], p[512][3];
  uint8_t mask[512];
  // recycling (core/block.h:16-18)
  int inner = 0, boundary = 0, life = 3;
  bool freed = false;
  void clear() {
    inner = boundary = 0;
    life = 3;
    for (int i = 0; i < 512; ++i) {
      sdf[i] = w[i] = a[i] = b[i] = 0.0;
      col[i][0] = col[i][1] = col[i][2] = 0;
      cube[i] = 0;
      sdf0[i] = sdf_bar[i] = wpd[i] = 0.0;
      p[i][0] = p[i][1] = p[i][2] = 0.0;
      mask[i] = 0;
    }
  }
};

struct Entry {
  I3 pos{0, 0, 0};
  int ptr = kFree;
  uint32_t offset = 0;
};

struct VKey {
  int k[4];  // owner voxel x, y, z, axis
  bool operator<(const VKey& o) const { return std::lexicographical_compare(k, k + 4, o.k, o.k + 4); }
};
struct Vtx {
  float pos[3];
  double normal[3], color[3], radius;
  int ref = 0;
};
struct TKey {
  int k[4];  // cube voxel x, y, z, slot
  bool operator<(const TKey& o) const { return std::lexicographical_compare(k, k + 4, o.k, o.k + 4); }
};

}  // namespace

// Threads of the stages that are parallel over blocks or pixels (integrate,
// marching cubes pass 1, raycast); results do not depend on it.  Allocation,
// triangle bookkeeping and vertex recycling stay serial.
static int g_threads = 1;

struct orc_volume {
  mh_hash_params hp;
  Geo g;
  Cam cam;
  float range_factor;
  bool psdf;
  std::vector<Entry> entries;
  std::vector<Block> blocks;  // block pool, index = ptr
  std::unordered_map<I3, int, I3Hash> where;
  std::vector<int> cand;  // worklist (pool indices)
  std::vector<uint8_t> cand_flag;  // garbage flag per worklist entry
  std::vector<int> free_ptrs;      // pool slots returned by FreeEntry (LIFO)
  // frame
  std::vector<float> depth;
  std::vector<double> inlier;
  std::vector<uint32_t> color_in;  // fusion colour per pixel, R | G << 8 | B << 16
  bool has_color = false;
  float wTc[16], cTw[16];
  double cTw64[16];
  // mesh
  std::map<VKey, Vtx> verts;
  std::map<TKey, std::array<VKey, 3>> tris;
  uint32_t dropped = 0;
  uint32_t list_inserts = 0;
  uint64_t voxel_updates = 0;
};

namespace {

int find_block(const orc_volume* v, I3 p) {
  auto it = v->where.find(p);
  return it == v->where.end() ? kFree : it->second;
}

// --------------------------------------------------------------- hash table
// HashTable::AllocEntry (hash_table.h:96-179) without contention: a key goes
// to the first free slot of its bucket; a full bucket links it into the list
// that starts at the bucket's last slot.  Bucket placements of a frame are
// made before its list placements (the library's policy; with fewer than
// bucket_size new keys per bucket and frame -- every config here -- the
// order is immaterial).
bool in_table(const orc_volume* v, I3 p) {
  const uint32_t first = bucket_of(p, v->hp.bucket_count) * v->hp.bucket_size;
  for (uint32_t j = 0; j < v->hp.bucket_size; ++j) {
    const Entry& e = v->entries[first + j];
    if (e.ptr != kFree && e.pos == p) return true;
  }
  const uint32_t last = first + v->hp.bucket_size - 1;
  uint32_t i = last;
  for (uint32_t it = 0; it < v->hp.linked_list_size; ++it) {
    const Entry& e = v->entries[i];
    if (e.ptr != kFree && e.pos == p) return true;
    if (e.offset == 0) break;
    i = (last + e.offset) % v->hp.entry_count;
  }
  return false;
}

int new_block(orc_volume* v, I3 p) {
  if (!v->free_ptrs.empty()) {  // Free / Alloc share one LIFO stack (hash_table.h:279-292)
    const int ptr = v->free_ptrs.back();
    v->free_ptrs.pop_back();
    Block& B = v->blocks[ptr];
    B.clear();
    B.freed = false;
    B.pos = p;
    v->where[p] = ptr;
    return ptr;
  }
  const int ptr = (int)v->blocks.size();
  if ((uint32_t)ptr >= v->hp.value_capacity) return kFree;
  v->blocks.emplace_back();
  v->blocks.back().clear();
  v->blocks.back().pos = p;
  v->where[p] = ptr;
  return ptr;
}

// HashTable::FreeEntry (hash_table.h:183-255, linked-list variant).
bool free_entry(orc_volume* v, I3 p) {
  const uint32_t N = v->hp.entry_count;
  const uint32_t first = bucket_of(p, v->hp.bucket_count) * v->hp.bucket_size;
  const uint32_t last = first + v->hp.bucket_size - 1;
  for (uint32_t j = 0; j < v->hp.bucket_size; ++j) {
    const uint32_t i = first + j;
    const Entry e = v->entries[i];
    if (!(e.ptr != kFree && e.pos == p)) continue;
    if (e.offset != 0) {
      const uint32_t nx = (i + e.offset) % N;
      v->entries[i] = v->entries[nx];
      v->entries[nx] = Entry{};
    } else {
      v->entries[i] = Entry{};
    }
    return true;
  }
  if (v->entries[last].offset == 0) return false;
  uint32_t prev = last, i = (last + v->entries[last].offset) % N;
  for (uint32_t it = 0; it < v->hp.linked_list_size; ++it) {
    const Entry e = v->entries[i];
    if (e.ptr != kFree && e.pos == p) {
      v->entries[i] = Entry{};
      v->entries[prev].offset = e.offset;
      return true;
    }
    if (e.offset == 0) return false;
    prev = i;
    i = (last + e.offset) % N;
  }
  return false;
}

// Returns false when the bucket is full (caller defers to list insertion).
bool bucket_insert(orc_volume* v, I3 p) {
  if (in_table(v, p)) return true;
  const uint32_t first = bucket_of(p, v->hp.bucket_count) * v->hp.bucket_size;
  for (uint32_t j = 0; j < v->hp.bucket_size; ++j) {
    Entry& e = v->entries[first + j];
    if (e.ptr == kFree) {
      e.pos = p;
      e.ptr = new_block(v, p);
      return true;
    }
  }
  return false;
}

void list_insert(orc_volume* v, I3 p) {
  if (in_table(v, p)) return;
  const uint32_t first = bucket_of(p, v->hp.bucket_count) * v->hp.bucket_size;
  for (uint32_t j = 0; j < v->hp.bucket_size; ++j) {
    Entry& e = v->entries[first + j];
    if (e.ptr == kFree) {
      e.pos = p;
      e.ptr = new_block(v, p);
      return;
    }
  }
  const uint32_t last = first + v->hp.bucket_size - 1;
  uint32_t offset = 0;
  for (uint32_t it = 0; it < v->hp.linked_list_size; ++it) {
    ++offset;
    if (offset % v->hp.bucket_size == 0) continue;
    const uint32_t i = (last + offset) % v->hp.entry_count;
    Entry& e = v->entries[i];
    if (e.ptr == kFree) {
      e.pos = p;
      e.offset = v->entries[last].offset;
      e.ptr = new_block(v, p);
      v->entries[last].offset = offset;
      ++v->list_inserts;
      return;
    }
  }
  ++v->dropped;
}

bool coord_ok(I3 p) {
  const int m = (1 << 20) - 1;
  return p.x >= -m && p.x <= m && p.y >= -m && p.y <= m && p.z >= -m && p.z <= m;
}

// ---------------------------------------------------------------- trilinear
// GetSpatialValue (spatial_query.h:23-66).  Corner voxels are chosen by
// FP32 rounding of the corner's world position; the interpolation itself
// is FP64 (or FP32 when fp32 is set, for the raycaster).
struct SV {
  double sdf, a, b, radius;
};
bool voxel_at(const orc_volume* v, I3 vox, int* bi, int* li) {
  const I3 b = Geo::v2b(vox);
  const int p = find_block(v, b);
  if (p < 0) return false;
  *bi = p;
  *li = (vox.x - b.x * 8) + 8 * (vox.y - b.y * 8) + 64 * (vox.z - b.z * 8);
  return true;
}

bool spatial64(const orc_volume* v, F3 pos, SV* out) {
  const Geo& g = v->g;
  const float hv = 0.5f * g.vs;
  const float pc[3] = {pos.x - hv, pos.y - hv, pos.z - hv};
  const float pp[3] = {pos.x, pos.y, pos.z};
  int cv[3][2];
  double r[3];
  for (int k = 0; k < 3; ++k) {
    cv[k][0] = g.w2v1(pc[k] + 0.0f);
    cv[k][1] = g.w2v1(pc[k] + g.vs);
    // The ratio is an FP32 quantity in the reference (frac(pos / vs)), and it
    // is decision-like: within an ulp of a voxel plane it puts the whole
    // weight on one corner layer or the other, independently of the corner
    // choice above.  MC vertices lie on voxel planes in two axes, so it is
    // taken in FP32 exactly as the reference and widened afterwards.
    const float q = pp[k] / g.vs;
    r[k] = (double)(q - std::floor(q));
  }
  double sdf = 0, a = 0, b = 0, rad = 0;
  for (int i = 0; i < 8; ++i) {
    const int m[3] = {(i >> 2) & 1, (i >> 1) & 1, i & 1};
    int bi, li;
    if (!voxel_at(v, I3{cv[0][m[0]], cv[1][m[1]], cv[2][m[2]]}, &bi, &li)) return false;
    const Block& B = v->blocks[bi];
    const double w = (m[0] ? r[0] : 1.0 - r[0]) * (m[1] ? r[1] : 1.0 - r[1]) * (m[2] ? r[2] : 1.0 - r[2]);
    sdf += w * B.sdf[li];
    a += w * B.a[li];
    b += w * B.b[li];
    rad += w * std::sqrt(1.0 / B.w[li]);
  }
  *out = SV{sdf, a, b, rad};
  return true;
}

bool spatial32(const orc_volume* v, F3 pos, float* out) {
  const Geo& g = v->g;
  const float hv = 0.5f * g.vs;
  const float pc[3] = {pos.x - hv, pos.y - hv, pos.z - hv};
  const float pp[3] = {pos.x, pos.y, pos.z};
  int cv[3][2];
  float r[3];
  for (int k = 0; k < 3; ++k) {
    cv[k][0] = g.w2v1(pc[k] + 0.0f);
    cv[k][1] = g.w2v1(pc[k] + g.vs);
    const float q = pp[k] / g.vs;
    r[k] = q - std::floor(q);
  }
  float sdf = 0.0f;
  for (int i = 0; i < 8; ++i) {
    const int m[3] = {(i >> 2) & 1, (i >> 1) & 1, i & 1};
    int bi, li;
    if (!voxel_at(v, I3{cv[0][m[0]], cv[1][m[1]], cv[2][m[2]]}, &bi, &li)) return false;
    const float w = (m[0] ? r[0] : 1.0f - r[0]) * (m[1] ? r[1] : 1.0f - r[1]) * (m[2] ? r[2] : 1.0f - r[2]);
    sdf += w * (float)v->blocks[bi].sdf[li];
  }
  *out = sdf;
  return true;
}

// ValToRGB (visualization/color_util.h:13-37), FP64.
void val_to_rgb(double v, double vmin, double vmax, double* c) {
  c[0] = c[1] = c[2] = 1.0;
  if (v < vmin) v = vmin;
  if (v > vmax) v = vmax;
  const double dv = vmax - vmin;
  if (v < vmin + 0.25 * dv) {
    c[0] = 0.0;
    c[1] = 4.0 * (v - vmin) / dv;
  } else if (v < vmin + 0.5 * dv) {
    c[0] = 0.0;
    c[2] = 1.0 + 4.0 * (vmin + 0.25 * dv - v) / dv;
  } else if (v < vmin + 0.75 * dv) {
    c[0] = 4.0 * (v - vmin - 0.5 * dv) / dv;
    c[2] = 0.0;
  } else {
    c[1] = 1.0 + 4.0 * (vmin + 0.75 * dv - v) / dv;
    c[2] = 0.0;
  }
}

void set_pose(orc_volume* v, const float* wTc) {
  std::memcpy(v->wTc, wTc, sizeof(float) * 16);
  inv4<float>(wTc, v->cTw);
  double w64[16];
  for (int i = 0; i < 16; ++i) w64[i] = (double)wTc[i];
  inv4<double>(w64, v->cTw64);
}

bool depth_ok(float d, float ub) { return !(d == -INFINITY || d == 0.0f || d >= ub); }

// Pixel of voxel vp (FP32 decision), or false when it falls outside.
bool project(const orc_volume* v, I3 vp, uint32_t* u, uint32_t* vv, F3* cp) {
  *cp = xf(v->cTw, v->g.v2w(vp));
  *u = (uint32_t)f2i(v->cam.pu(*cp) + 0.5f);
  *vv = (uint32_t)f2i(v->cam.pv(*cp) + 0.5f);
  return *u < v->cam.w && *vv < v->cam.h;
}

double cam_z64(const orc_volume* v, I3 vp) {
  const double vs = (double)v->g.vs;
  const double X = (double)vp.x * vs, Y = (double)vp.y * vs, Z = (double)vp.z * vs;
  return v->cTw64[8] * X + v->cTw64[9] * Y + v->cTw64[10] * Z + v->cTw64[11];
}

I3 voxel_of(const Block& B, int i) {
  return I3{B.pos.x * 8 + i % 8, B.pos.y * 8 + (i % 64) / 8, B.pos.z * 8 + i / 64};
}

}  // namespace

extern "C" {

orc_volume* orc_create(const mh_hash_params* hp, const mh_volume_params* vp,
                       const mh_sensor_params* sp, int psdf) {
  orc_volume* v = new orc_volume();
  v->hp = *hp;
  v->g = Geo{vp->voxel_size, vp->truncation_distance, vp->truncation_distance_scale,
             vp->sdf_upper_bound, (float)vp->weight_sample};
  v->cam = Cam{sp->fx, sp->fy, sp->cx, sp->cy, sp->width, sp->height, sp->min_depth_range,
               sp->max_depth_range};
  v->range_factor = sp->range_factor;
  v->psdf = psdf != 0;
  v->entries.assign(hp->entry_count, Entry{});
  v->blocks.reserve(std::min<uint32_t>(hp->value_capacity, 1u << 16));
  v->depth.assign((size_t)sp->width * sp->height, -INFINITY);
  v->inlier.assign((size_t)sp->width * sp->height, 0.1);
  const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  set_pose(v, id);
  return v;
}

void orc_destroy(orc_volume* v) { delete v; }

uint32_t orc_hash_bucket(int32_t x, int32_t y, int32_t z, uint32_t bucket_count) {
  return bucket_of(I3{x, y, z}, bucket_count);
}

void orc_invert4f(const float* m, float* out) { inv4<float>(m, out); }
void orc_invert4d(const double* m, double* out) { inv4<double>(m, out); }

// Sensor::Process (sensor/rgbd_sensor.cu:118-137): ConvertDepthFormatKernel
// (sensor/preprocess.cu:25-42) and ResetInlierRatioKernel (:10-23).
int orc_set_frame(orc_volume* v, const uint16_t* raw, const float* wTc) {
  set_pose(v, wTc);
  v->has_color = false;
  for (size_t i = 0; i < v->depth.size(); ++i) {
    const float d = v->range_factor * (float)(int16_t)raw[i];
    v->depth[i] = (d >= v->cam.dmin && d <= v->cam.dmax) ? d : -INFINITY;
    v->inlier[i] = 0.1;
  }
  return 0;
}

// ConvertColorFormatKernel (sensor/preprocess.cu:44-58): BGRA, valid iff B,
// G and R are non-zero, channel c -> c / 255.0f (FP32); then the fusion
// colour of UpdateBlocksSimpleKernel (update_simple.cu:68-70),
// uchar(255 * x) with truncation clamped to [0, 255] (-inf -> 0).
int orc_set_color(orc_volume* v, const uint8_t* bgra) {
  v->color_in.assign(v->depth.size(), 0u);
  auto conv = [](uint8_t c) -> uint32_t {
    const float x = (float)c / 255.0f;
    const int i = f2i(255.0f * x);
    return (uint32_t)std::max(0, std::min(255, i));
  };
  for (size_t i = 0; i < v->depth.size(); ++i) {
    const uint8_t b = bgra[4 * i], g = bgra[4 * i + 1], r = bgra[4 * i + 2];
    v->color_in[i] = (b && g && r) ? (conv(r) | (conv(g) << 8) | (conv(b) << 16)) : 0u;
  }
  v->has_color = true;
  return 0;
}

// AllocBlockArrayKernel (mapping/allocate.cu:8-99), pixels in row-major order.
int orc_alloc(orc_volume* v) {
  const Geo& g = v->g;
  const Cam& c = v->cam;
  std::vector<I3> deferred;
  for (uint32_t y = 0; y < c.h; ++y) {
    for (uint32_t x = 0; x < c.w; ++x) {
      const float depth = v->depth[y * c.w + x];
      if (!depth_ok(depth, g.ub)) continue;
      const float tr = g.truncate(depth);
      const float nd = std::min(g.ub, depth - tr);
      const float fd = std::min(g.ub, depth + tr);
      if (nd >= fd) continue;
      const F3 wn = xf(v->wTc, c.reproj(x, y, nd));
      const F3 wf = xf(v->wTc, c.reproj(x, y, fd));
      const F3 dir = normalizef(F3{wf.x - wn.x, wf.y - wn.y, wf.z - wn.z});
      const I3 bn = Geo::v2b(g.w2v(wn));
      const I3 bf = Geo::v2b(g.w2v(wf));
      const float st[3] = {(float)sgn(dir.x), (float)sgn(dir.y), (float)sgn(dir.z)};
      const float d3[3] = {dir.x, dir.y, dir.z};
      const float wn3[3] = {wn.x, wn.y, wn.z};
      const int bn3[3] = {bn.x, bn.y, bn.z};
      const int bf3[3] = {bf.x, bf.y, bf.z};
      float t[3], dt[3];
      int bound[3], cur[3] = {bn.x, bn.y, bn.z};
      for (int k = 0; k < 3; ++k) {
        const int off = f2i(std::min(std::max(st[k], 0.0f), 1.0f));
        const float bw = (float)((bn3[k] + off) * 8) * g.vs;
        const float nvc = bw - 0.5f * g.vs;
        t[k] = (nvc - wn3[k]) / d3[k];
        dt[k] = (st[k] * 8.0f * g.vs) / d3[k];
        bound[k] = f2i((float)bf3[k] + st[k]);
        if (d3[k] == 0.0f) {
          t[k] = INFINITY;
          dt[k] = INFINITY;
        }
      }
      for (int it = 0; it < 1024; ++it) {
        const I3 b{cur[0], cur[1], cur[2]};
        if (block_in_frustum(g, v->cTw, b, c) && coord_ok(b))
          if (!bucket_insert(v, b)) deferred.push_back(b);
        int k;
        if (t[0] < t[1] && t[0] < t[2]) k = 0;
        else if (t[1] < t[2]) k = 1;
        else k = 2;
        cur[k] += (int)st[k];
        if (cur[k] == bound[k]) break;
        t[k] += dt[k];
      }
    }
  }
  for (const I3& b : deferred) list_insert(v, b);
  return 0;
}

// CollectBlocksInFrustumKernel / CollectAllBlocksKernel
// (core/collect_block_array.cu:25-91).
uint32_t orc_collect(orc_volume* v, int frustum) {
  v->cand.clear();
  for (size_t p = 0; p < v->blocks.size(); ++p)
    if (v->blocks[p].freed) continue;
    else if (!frustum || block_in_frustum(v->g, v->cTw, v->blocks[p].pos, v->cam)) v->cand.push_back((int)p);
  v->cand_flag.assign(v->cand.size(), 0);
  return (uint32_t)v->cand.size();
}

// UpdateBlocksSimpleKernel (mapping/update_simple.cu:13-77), Voxel::Update
// (core/voxel.h:107-115).  Colour stream absent: delta colour (0, 255, 0).
int orc_integrate_tsdf(orc_volume* v) {
  const Geo& g = v->g;
  const Cam& c = v->cam;
  uint64_t updates = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : updates) num_threads(g_threads)
  for (size_t ci = 0; ci < v->cand.size(); ++ci) {
    Block& B = v->blocks[v->cand[ci]];
    for (int i = 0; i < 512; ++i) {
      const I3 vp = voxel_of(B, i);
      uint32_t u, vv;
      F3 cp;
      if (!project(v, vp, &u, &vv, &cp)) continue;
      const float depth = v->depth[vv * c.w + u];
      if (!depth_ok(depth, g.ub)) continue;
      const float sdf_dec = depth - cp.z;
      const float tr = g.truncate(depth);
      if (sdf_dec <= -tr) continue;
      // values
      double s = (double)depth - cam_z64(v, vp);
      const double trd = (double)g.tr + (double)g.tr_scale * (double)depth;
      s = std::max(-trd, std::min(trd, s));
      const double nd = ((double)depth - c.dmin) / ((double)c.dmax - c.dmin);
      const double w = std::max(10.0 * (double)g.ws * (1.0 - nd), 1.0);
      B.sdf[i] = (B.sdf[i] * B.w[i] + s * w) / (B.w[i] + w);
      B.w[i] += w;
      ++updates;
      uint8_t dcol[3] = {0, 255, 0};  // no colour image (update_simple.cu:71-73)
      if (v->has_color) {
        const uint32_t cin = v->color_in[vv * c.w + u];
        dcol[0] = (uint8_t)(cin & 255u);
        dcol[1] = (uint8_t)((cin >> 8) & 255u);
        dcol[2] = (uint8_t)((cin >> 16) & 255u);
      }
      for (int k = 0; k < 3; ++k) {
        const float cc = 0.5f * (float)B.col[i][k] + 0.5f * (float)dcol[k];
        B.col[i][k] = (uint8_t)(cc + 0.5f);
      }
    }
  }
  v->voxel_updates += updates;
  return 0;
}

// PredictOutlierRatioKernel (mapping/update_bayesian.cu:36-101).  Vertices
// come from the previous meshing pass.  Note the normal is transformed as
// a point (translation included), as in the reference.
int orc_predict_inlier(orc_volume* v) {
  const Geo& g = v->g;
  const Cam& c = v->cam;
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      const I3 vp = voxel_of(B, i);
      uint32_t u, vv;
      F3 cp;
      if (!project(v, vp, &u, &vv, &cp)) continue;
      const uint32_t pix = u + vv * c.w;
      const float depth = v->depth[pix];
      if (!depth_ok(depth, g.ub)) continue;
      // FP32 in the reference's operation order: proj_disk is the square
      // root of a difference that cancels when vec runs along the normal, so
      // the ratio is decision-like there (a negative difference gives NaN,
      // which fmaxf(., 0.01) turns into 0.01), as on the GPU
      const F3 pc = c.reproj(u, vv, depth);
      const float sdf = (float)B.sdf[i], is2 = (float)B.w[i];
      for (int ax = 0; ax < 3; ++ax) {
        auto it = v->verts.find(VKey{{vp.x, vp.y, vp.z, ax}});
        if (it == v->verts.end()) continue;
        const Vtx& vt = it->second;
        const F3 n{(float)vt.normal[0], (float)vt.normal[1], (float)vt.normal[2]};
        const F3 x{vt.pos[0], vt.pos[1], vt.pos[2]};
        const F3 nc = normalizef(xf(v->cTw, n));  // the normal as a point
        const float cos_a = dotf(normalizef(F3{-pc.x, -pc.y, -pc.z}), nc);
        const F3 xc = xf(v->cTw, x);
        const F3 vec{pc.x - xc.x, pc.y - xc.y, pc.z - xc.z};
        const float pd = dotf(vec, nc);
        const float disk = std::sqrt(dotf(vec, vec) - pd * pd);
        const float dd = pd - sdf;
        const float w_dist = std::exp(-(dd * dd) * is2);
        const float w_disk = 0.5f + 0.5f / (1.0f + std::exp(-std::fabs(disk / (float)vt.radius)));
        const float w_ang = std::fmax((cos_a - 0.1736f) / (1.0f - 0.1736f), 0.1f);
        const float r = std::fmax(w_disk * w_dist * w_ang, 0.01f);
        v->inlier[pix] = std::max(v->inlier[pix], (double)r);
      }
    }
  }
  return 0;
}

// UpdateBlocksBayesianKernel (mapping/update_bayesian.cu:103-185) with
// gaussian() of geometry_helper.h:30-34 (no 1/2 in the exponent).
int orc_integrate_psdf(orc_volume* v) {
  const Geo& g = v->g;
  const Cam& c = v->cam;
  const double sqrt2pi = std::sqrt(2.0 * 3.14159265358979323846);
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      const I3 vp = voxel_of(B, i);
      uint32_t u, vv;
      F3 cp;
      if (!project(v, vp, &u, &vv, &cp)) continue;
      const uint32_t pix = u + vv * c.w;
      const float depth = v->depth[pix];
      if (!depth_ok(depth, g.ub)) continue;
      const float x_dec = depth - cp.z;
      const float tr = g.truncate(depth);
      if (x_dec <= -tr) continue;
      double x = (double)depth - cam_z64(v, vp);
      const double trd = (double)g.tr + (double)g.tr_scale * (double)depth;
      x = std::max(-trd, std::min(trd, x));
      const double rho = v->inlier[pix];
      const double tau = ((double)depth - 0.4) * 0.012 + 0.019;
      if (B.w[i] == 0.0) {
        B.sdf[i] = x;
        B.w[i] = 1.0 / (tau * tau);
        B.a[i] = rho * 10.0;
        B.b[i] = (1.0 - rho) * 10.0;
        continue;
      }
      const double mu = B.sdf[i], s2 = 1.0 / B.w[i], t2 = tau * tau;
      const double ss2 = 1.0 / (1.0 / s2 + 1.0 / t2);
      const double m = ss2 * (mu / s2 + x / t2);
      const double var = s2 + t2;
      double C1 = rho * std::exp(-(x - mu) * (x - mu) / var) / (sqrt2pi * std::sqrt(var));
      double C2 = (depth < c.dmax && depth >= c.dmin) ? (1.0 - rho) / (5.0 - 0.1) : 1.0;
      const double cs = C1 + C2;
      C1 /= cs;
      C2 /= cs;
      const double a = B.a[i], b = B.b[i];
      const double f = C1 * (a + 1) / (a + b + 1) + C2 * a / (a + b + 1);
      const double e = C1 * (a + 1) * (a + 2) / ((a + b + 1) * (a + b + 2)) +
                       C2 * a * (a + 1) / ((a + b + 1) * (a + b + 2));
      const double ns = C1 * m + C2 * mu;
      B.sdf[i] = ns;
      B.w[i] = 1.0 / (C1 * (ss2 + m * m) + C2 * (s2 + mu * mu) - ns * ns);
      B.a[i] = (e - f) / (f - e / f);
      B.b[i] = B.a[i] * (1.0 - f) / f;
    }
  }
  return 0;
}

// Marching cubes (meshing/marching_cubes.cu:16-321) over the worklist, with
// canonical vertices: (owner voxel, axis), endpoints owner -> owner + e_axis.
// Releases the unreferenced vertex slots owned by the given blocks, after
// counting each block's allocated slots as inner (every local coordinate in
// 1..6) or boundary surfels.
static void recycle_vertices(orc_volume* v, const std::vector<int>& ptrs) {
  for (int ptr : ptrs) {
    Block& B = v->blocks[ptr];
    B.inner = B.boundary = 0;
    for (int i = 0; i < 512; ++i) {
      const I3 vp = voxel_of(B, i);
      const int lx = i % 8, ly = (i / 8) % 8, lz = i / 64;
      const bool inner = lx >= 1 && ly >= 1 && lz >= 1 && lx < 7 && ly < 7 && lz < 7;
      for (int ax = 0; ax < 3; ++ax) {
        auto it = v->verts.find(VKey{{vp.x, vp.y, vp.z, ax}});
        if (it == v->verts.end()) continue;
        if (inner) ++B.inner; else ++B.boundary;
        if (it->second.ref == 0) v->verts.erase(it);
      }
    }
  }
}

int orc_marching_cubes(orc_volume* v, const mh_mc_params* mp) {
  const Geo& g = v->g;
  const float inv = 1.0f / g.vs;
  const float wmin = mp->weight_min < 0.0f ? inv * inv : mp->weight_min;
  struct Corner {
    bool ok;
    float d;
  };
  auto corner_val = [&](I3 vox, float* d, float* w, float* a, float* b) {
    int bi, li;
    if (!voxel_at(v, vox, &bi, &li)) return false;
    const Block& B = v->blocks[bi];
    *d = (float)B.sdf[li];
    *w = (float)B.w[li];
    *a = (float)B.a[li];
    *b = (float)B.b[li];
    return true;
  };
  // pass 1, parallel over the worklist: each block writes only its own cube
  // words; vertices go to per-thread lists and are merged afterwards (every
  // writer of a vertex computes the same values, so the merge order is free)
  std::vector<std::vector<std::pair<VKey, Vtx>>> found((size_t)std::max(g_threads, 1));
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (size_t ci = 0; ci < v->cand.size(); ++ci) {
    const int ptr = v->cand[ci];
#ifdef _OPENMP
    auto& mine = found[(size_t)omp_get_thread_num()];
#else
    auto& mine = found[0];
#endif
    for (int i = 0; i < 512; ++i) {
      Block& B = v->blocks[ptr];
      const I3 vp = voxel_of(B, i);
      const uint16_t prev = B.cube[i] & 255;
      int cube = 0;
      bool valid = true;
      for (int k = 0; k < 8 && valid; ++k) {
        float d, w, a, b;
        if (!corner_val(I3{vp.x + corner(k, 0), vp.y + corner(k, 1), vp.z + corner(k, 2)}, &d, &w, &a, &b)) {
          valid = false;
          break;
        }
        if (std::fabs(d) > mp->sdf_threshold) { valid = false; break; }
        const float rho = a / (a + b);
        if (rho < mp->rho_min || w < wmin) { valid = false; break; }
        if (d < 0.0f) cube |= 1 << k;
      }
      if (!valid) cube = 0;
      v->blocks[ptr].cube[i] = (uint16_t)(cube | (prev << 8));
      if (cube == 0 || cube == 255) continue;
      const unsigned em = edge_mask(cube);
      for (int e = 0; e < 12; ++e) {
        if (!(em & (1u << e))) continue;
        int ax = 0;
        for (int q = 0; q < 3; ++q)
          if (corner(ea(e), q) != corner(eb(e), q)) ax = q;
        const I3 o{vp.x + std::min(corner(ea(e), 0), corner(eb(e), 0)),
                   vp.y + std::min(corner(ea(e), 1), corner(eb(e), 1)),
                   vp.z + std::min(corner(ea(e), 2), corner(eb(e), 2))};
        const I3 o2{o.x + (ax == 0), o.y + (ax == 1), o.z + (ax == 2)};
        float vA, vB, t0, t1, t2;
        corner_val(o, &vA, &t0, &t1, &t2);
        corner_val(o2, &vB, &t0, &t1, &t2);
        const F3 pA = g.v2w(o);
        const F3 pB{pA.x + (ax == 0 ? g.vs : 0.0f), pA.y + (ax == 1 ? g.vs : 0.0f),
                    pA.z + (ax == 2 ? g.vs : 0.0f)};
        // VertexIntersection (marching_cubes.cu:16-28); 0.008 is a double.
        F3 p;
        if ((double)std::fabs(vA) < mp->snap_distance) p = pA;
        else if ((double)std::fabs(vB) < mp->snap_distance) p = pB;
        else {
          const float mu = (0.0f - vA) / (vB - vA);
          p = F3{pA.x + mu * (pB.x - pA.x), pA.y + mu * (pB.y - pA.y), pA.z + mu * (pB.z - pA.z)};
        }
        Vtx vt;
        vt.pos[0] = p.x;
        vt.pos[1] = p.y;
        vt.pos[2] = p.z;
        SV sv;
        if (spatial64(v, p, &sv)) {
          vt.radius = std::sqrt(1.0 / (1.0 / (sv.radius * sv.radius)));
          val_to_rgb(sv.a / (sv.a + sv.b), 0.4, 1.0, vt.color);
        } else {
          vt.radius = 0.0;
          vt.color[0] = vt.color[1] = vt.color[2] = 0.0;
        }
        // GetSpatialSDFGradient (spatial_query.h:68-97)
        const float h = 0.5f * g.vs;
        bool ok = true;
        double sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
        for (int k = 0; k < 3 && ok; ++k) {
          F3 pm = p, pq = p;
          if (k == 0) { pm.x = p.x - h; pq.x = p.x + h; }
          if (k == 1) { pm.y = p.y - h; pq.y = p.y + h; }
          if (k == 2) { pm.z = p.z - h; pq.z = p.z + h; }
          SV q;
          ok = spatial64(v, pm, &q);
          sn[k] = q.sdf;
          if (!ok) break;
          ok = spatial64(v, pq, &q);
          sp[k] = q.sdf;
        }
        const double gr[3] = {(sp[0] - sn[0]) / g.vs, (sp[1] - sn[1]) / g.vs, (sp[2] - sn[2]) / g.vs};
        const double l = std::sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
        for (int k = 0; k < 3; ++k) vt.normal[k] = (l > 0.0 && ok) ? gr[k] / l : 0.0;
        mine.emplace_back(VKey{{o.x, o.y, o.z, ax}}, vt);
      }
    }
  }
  for (const auto& list : found)
    for (const auto& kv : list) {
      Vtx& vt = v->verts[kv.first];  // keeps the reference count of an existing vertex
      std::memcpy(vt.pos, kv.second.pos, sizeof(vt.pos));
      std::memcpy(vt.normal, kv.second.normal, sizeof(vt.normal));
      std::memcpy(vt.color, kv.second.color, sizeof(vt.color));
      vt.radius = kv.second.radius;
    }
  // pass 2 (TriangleExtractionKernel, :180-275)
  for (int ptr : v->cand) {
    for (int i = 0; i < 512; ++i) {
      const Block& B = v->blocks[ptr];
      const I3 vp = voxel_of(B, i);
      const int cube = B.cube[i] & 255;
      if (cube == 0 || cube == 255) continue;
      VKey keys[12];
      for (int e = 0; e < 12; ++e) {
        int ax = 0;
        for (int q = 0; q < 3; ++q)
          if (corner(ea(e), q) != corner(eb(e), q)) ax = q;
        keys[e] = VKey{{vp.x + std::min(corner(ea(e), 0), corner(eb(e), 0)),
                        vp.y + std::min(corner(ea(e), 1), corner(eb(e), 1)),
                        vp.z + std::min(corner(ea(e), 2), corner(eb(e), 2)), ax}};
      }
      for (int t = 0; t < ntri(cube); ++t) {
        const TKey tk{{vp.x, vp.y, vp.z, t}};
        auto it = v->tris.find(tk);
        if (it != v->tris.end())
          for (const VKey& k : it->second) v->verts[k].ref -= 1;
        std::array<VKey, 3> tv{keys[tri_edge(cube, 3 * t)], keys[tri_edge(cube, 3 * t + 1)],
                               keys[tri_edge(cube, 3 * t + 2)]};
        v->tris[tk] = tv;
        for (const VKey& k : tv) v->verts[k].ref += 1;
      }
    }
  }
  // RecycleTrianglesKernel (:278-301)
  for (int ptr : v->cand) {
    const Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      const I3 vp = voxel_of(B, i);
      for (int t = ntri(B.cube[i] & 255); t < 5; ++t) {
        auto it = v->tris.find(TKey{{vp.x, vp.y, vp.z, t}});
        if (it == v->tris.end()) continue;
        for (const VKey& k : it->second) v->verts[k].ref -= 1;
        v->tris.erase(it);
      }
    }
  }
  // RecycleVerticesKernel (:303-321); the surfel counts of
  // TriangleExtractionKernel (:187-205) are taken over the allocated vertex
  // slots before they are released.
  recycle_vertices(v, v->cand);
  return 0;
}

// ---------------------------------------------------------------- recycling
// CollectLowSurfelBlocksKernel (mapping/recycle.cu:81-115), in two passes
// (ageing, then flagging against the aged neighbours).
int orc_collect_low_surfel(orc_volume* v) {
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    B.life = (B.inner < 10 && B.boundary == 0) ? B.life - 1 : 3;
  }
  v->cand_flag.assign(v->cand.size(), 0);
  const I3 off[6] = {{0, 0, 1}, {0, 0, -1}, {0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}};
  for (size_t i = 0; i < v->cand.size(); ++i) {
    const Block& B = v->blocks[v->cand[i]];
    if (B.life > 0) continue;
    uint8_t f = 1;
    for (const I3& o : off) {
      const int q = find_block(v, I3{B.pos.x + o.x, B.pos.y + o.y, B.pos.z + o.z});
      if (q >= 0 && v->blocks[q].life != 0) f = 0;
    }
    v->cand_flag[i] = f;
  }
  return 0;
}

// StarveOccupiedBlocksKernel (recycle.cu:15-24).
int orc_starve(orc_volume* v) {
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) B.w[i] = std::max(0.0, B.w[i] - 1.0);
  }
  return 0;
}

// CollectGarbageBlockArrayKernel (recycle.cu:27-77); EPSILON is the double
// 1e-6 compared with the FP32 weight.
int orc_collect_garbage(orc_volume* v) {
  const float t = v->g.truncate(5.0f);
  v->cand_flag.assign(v->cand.size(), 0);
  for (size_t k = 0; k < v->cand.size(); ++k) {
    const Block& B = v->blocks[v->cand[k]];
    float mn = INFINITY, mx = 0.0f;
    for (int i = 0; i < 512; ++i) {
      const float w = (float)B.w[i];
      mn = std::min(mn, (double)w < 1e-6 ? INFINITY : std::fabs((float)B.sdf[i]));
      mx = std::max(mx, w);
    }
    v->cand_flag[k] = (mn >= t || (double)mx < 1e-6) ? 1 : 0;
  }
  return 0;
}

// RecycleGarbageBlockArray (recycle.cu:117-179, 231-260): the flagged
// blocks' triangles are released (their cubes get configuration 0), their
// own unreferenced vertices freed, and a flagged block with no vertex left
// is removed from the table (FreeEntry) and cleared.
uint32_t orc_recycle_garbage(orc_volume* v) {
  std::vector<int> gc;
  for (size_t k = 0; k < v->cand.size() && k < v->cand_flag.size(); ++k)
    if (v->cand_flag[k]) gc.push_back(v->cand[k]);
  for (int ptr : gc) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      const I3 vp = voxel_of(B, i);
      for (int t = 0; t < 5; ++t) {
        auto it = v->tris.find(TKey{{vp.x, vp.y, vp.z, t}});
        if (it == v->tris.end()) continue;
        for (const VKey& k : it->second) v->verts[k].ref -= 1;
        v->tris.erase(it);
      }
      B.cube[i] = (uint16_t)((B.cube[i] & 255) << 8);
    }
  }
  recycle_vertices(v, gc);
  uint32_t freed = 0;
  for (int ptr : gc) {
    Block& B = v->blocks[ptr];
    bool any = false;  // a vertex slot still referenced by a neighbour's triangle
    for (int i = 0; i < 512 && !any; ++i) {
      const I3 vp = voxel_of(B, i);
      for (int ax = 0; ax < 3 && !any; ++ax) any = v->verts.count(VKey{{vp.x, vp.y, vp.z, ax}}) > 0;
    }
    if (any) continue;
    if (!free_entry(v, B.pos)) continue;
    v->where.erase(B.pos);
    B.clear();
    B.freed = true;
    v->free_ptrs.push_back(ptr);
    ++freed;
  }
  return freed;
}

// CastKernel (visualization/ray_caster.cu:16-150) with BisectionIntersection
// (geometry/isosurface_intersection.h:21-55), FP32 throughout.
int orc_raycast(orc_volume* v, const mh_raycaster_params* rp, const float* cTw, float* depth_out,
                float* vertex_out, float* normal_out) {
  float wTc[16];
  inv4<float>(cTw, wTc);
  const Cam c{rp->fx, rp->fy, rp->cx, rp->cy, rp->width, rp->height, rp->min_raycast_depth,
              rp->max_raycast_depth};
#pragma omp parallel for schedule(dynamic, 2) num_threads(g_threads)
  for (uint32_t y = 0; y < rp->height; ++y) {
    for (uint32_t x = 0; x < rp->width; ++x) {
      const uint32_t pix = y * rp->width + x;
      float od = -INFINITY;
      F3 ov{-INFINITY, -INFINITY, -INFINITY}, on{-INFINITY, -INFINITY, -INFINITY};
      const F3 dc = normalizef(c.reproj(x, y, 1.0f));
      const float t_min = rp->min_raycast_depth / dc.z;
      const float t_max = rp->max_raycast_depth / dc.z;
      const F3 o = xf(wTc, F3{0.0f, 0.0f, 0.0f});
      const F3 dir = normalizef(xfd(wTc, dc));
      float ps = 0.0f, pt = 0.0f;
      bool pok = false;
      for (float t = t_min; t < t_max; t += rp->raycast_step) {
        float s;
        if (!spatial32(v, F3{o.x + t * dir.x, o.y + t * dir.y, o.z + t * dir.z}, &s)) continue;
        if (pok && ps > 0.0f && s < 0.0f) {
          float l = pt, r = t, m = (l + r) * 0.5f, ls = ps, rs = s;
          bool found = true;
          for (int k = 0; k < 3; ++k) {
            m = l + (ls / (ls - rs)) * (r - l);
            float ms;
            if (!spatial32(v, F3{o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z}, &ms)) {
              found = false;
              break;
            }
            if (ls * ms > 0.0f) { l = m; ls = ms; }
            else { r = m; rs = ms; }
          }
          if (found && std::fabs(ps - s) < rp->sample_sdf_threshold && std::fabs(s) < rp->sdf_threshold) {
            od = m * dc.z;
            ov = c.reproj(x, y, od);
            if (rp->enable_gradients) {
              const F3 iso{o.x + m * dir.x, o.y + m * dir.y, o.z + m * dir.z};
              const float h = 0.5f * v->g.vs;
              bool ok = true;
              float sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
              for (int k = 0; k < 3 && ok; ++k) {
                F3 pm = iso, pq = iso;
                if (k == 0) { pm.x = iso.x - h; pq.x = iso.x + h; }
                if (k == 1) { pm.y = iso.y - h; pq.y = iso.y + h; }
                if (k == 2) { pm.z = iso.z - h; pq.z = iso.z + h; }
                ok = spatial32(v, pm, &sn[k]);
                if (!ok) break;
                ok = spatial32(v, pq, &sp[k]);
              }
              const float vs = v->g.vs;
              const F3 gr{(sp[0] - sn[0]) / vs, (sp[1] - sn[1]) / vs, (sp[2] - sn[2]) / vs};
              const float l2 = std::sqrt(dotf(gr, gr));
              F3 nw = (l2 > 0.0f && ok) ? F3{gr.x / l2, gr.y / l2, gr.z / l2} : F3{0.0f, 0.0f, 0.0f};
              nw = F3{-nw.x, -nw.y, -nw.z};
              on = xfd(cTw, nw);
            }
            break;
          }
        }
        ps = s;
        pt = t;
        pok = true;
      }
      if (depth_out) depth_out[pix] = od;
      if (vertex_out) { vertex_out[3 * pix] = ov.x; vertex_out[3 * pix + 1] = ov.y; vertex_out[3 * pix + 2] = ov.z; }
      if (normal_out) { normal_out[3 * pix] = on.x; normal_out[3 * pix + 1] = on.y; normal_out[3 * pix + 2] = on.z; }
    }
  }
  return 0;
}

// PointToSurfaceKernel (localizing/point_to_psdf.cu:19-82) on the current
// frame's depth: odd pixels (both coordinates odd) with a depth that is
// finite, non-zero and below sdf_upper_bound; the world point wTc * p in
// FP32; GetSpatialValue and GetSpatialSDFGradient (spatial_query.h:23-97) in
// FP32; J = (g, p x g) for the tangent (translation, rotation).  mode 0 puts
// g.x in J's third slot, as the reference does (:69-74); mode 1 puts g.z.
// Per pixel J J^T, d J and d^2 are FP32 products (the reference's mat6x6
// arithmetic), summed in FP64.  out: A[36] row-major, b[6], err, count.
int orc_point_to_psdf(const orc_volume* v, const float* wTc, int mode, double* out) {
  const Cam& c = v->cam;
  const float vs = v->g.vs;
  const float h = 0.5f * vs;
  for (int k = 0; k < 44; ++k) out[k] = 0.0;
  for (uint32_t y = 1; y < c.h; y += 2) {
    for (uint32_t x = 1; x < c.w; x += 2) {
      const float depth = v->depth[(size_t)y * c.w + x];
      if (!depth_ok(depth, v->g.ub)) continue;
      const F3 pw = xf(wTc, c.reproj(x, y, depth));
      float d;
      if (!spatial32(v, pw, &d)) continue;
      bool ok = true;
      float sn[3] = {0, 0, 0}, sp[3] = {0, 0, 0};
      for (int k = 0; k < 3 && ok; ++k) {
        F3 pm = pw, pq = pw;
        if (k == 0) { pm.x = pw.x - h; pq.x = pw.x + h; }
        if (k == 1) { pm.y = pw.y - h; pq.y = pw.y + h; }
        if (k == 2) { pm.z = pw.z - h; pq.z = pw.z + h; }
        ok = spatial32(v, pm, &sn[k]);
        if (ok) ok = spatial32(v, pq, &sp[k]);
      }
      if (!ok) continue;
      const F3 g{(sp[0] - sn[0]) / vs, (sp[1] - sn[1]) / vs, (sp[2] - sn[2]) / vs};
      const float J[6] = {g.x, g.y, mode ? g.z : g.x,
                          -g.y * pw.z + g.z * pw.y,
                          -g.z * pw.x + g.x * pw.z,
                          -g.x * pw.y + g.y * pw.x};
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) out[6 * i + j] += (double)(J[i] * J[j]);
      for (int i = 0; i < 6; ++i) out[36 + i] += (double)(d * J[i]);
      out[42] += (double)(d * d);
      out[43] += 1.0;
    }
  }
  return 0;
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------- raycast adjoint
// The backward pass of the raycaster (no counterpart in the reference): the
// depth of a pixel as a function of a left pose perturbation xi = (omega,
// upsilon), wTc <- exp(xi^) wTc, and of the voxel SDFs, differentiated
// through the bracketing samples and the three regula-falsi steps of the
// march (visualization/ray_caster.cu:56-92, isosurface_intersection.h:24-55)
// and the trilinear interpolant (spatial_query.h:23-66).
//   T = float : the march exactly as orc_raycast (FP32 positions, values and
//               decisions); derivatives in FP64.  The GPU's parity target.
//   T = double: the whole march in FP64.  The finite-difference target.
template <typename T>
struct V3 {
  T x, y, z;
};
template <typename T>
V3<T> nrm(V3<T> v) {
  const T s = T(1) / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  return V3<T>{v.x * s, v.y * s, v.z * s};
}
template <typename T>
int w2v1T(T w, T vs) {
  const T p = w / vs;
  const T q = p + (p < T(0) ? T(-1) : T(1)) * T(0.5);
  if (!(q == q)) return 0;
  if (q >= T(2147483647.0)) return 2147483647;
  if (q <= T(-2147483648.0)) return -2147483647 - 1;
  return (int)q;
}

template <typename T>
struct TriG {
  T val;
  double g[3];
  I3 lo;  // lowest corner voxel (cell id for path checks)
  int bi[8], li[8];
  double w[8];
};

template <typename T>
bool tri_eval(const orc_volume* v, V3<T> pos, TriG<T>* o) {
  const T vs = (T)v->g.vs;
  const T hv = T(0.5) * vs;
  const T pc[3] = {pos.x - hv, pos.y - hv, pos.z - hv};
  const T pp[3] = {pos.x, pos.y, pos.z};
  int cv[3][2];
  T r[3];
  for (int k = 0; k < 3; ++k) {
    cv[k][0] = w2v1T<T>(pc[k] + T(0), vs);
    cv[k][1] = w2v1T<T>(pc[k] + vs, vs);
    const T q = pp[k] / vs;
    r[k] = q - std::floor(q);
  }
  T val = T(0);
  double g[3] = {0, 0, 0};
  for (int i = 0; i < 8; ++i) {
    const int m[3] = {(i >> 2) & 1, (i >> 1) & 1, i & 1};
    int bi, li;
    if (!voxel_at(v, I3{cv[0][m[0]], cv[1][m[1]], cv[2][m[2]]}, &bi, &li)) return false;
    const T c = (T)v->blocks[bi].sdf[li];
    const T wx = m[0] ? r[0] : T(1) - r[0], wy = m[1] ? r[1] : T(1) - r[1], wz = m[2] ? r[2] : T(1) - r[2];
    const T w = wx * wy * wz;
    val += w * c;
    const double cd = (double)c;
    g[0] += (m[0] ? cd : -cd) * (double)wy * (double)wz;
    g[1] += (m[1] ? cd : -cd) * (double)wx * (double)wz;
    g[2] += (m[2] ? cd : -cd) * (double)wx * (double)wy;
    o->bi[i] = bi;
    o->li[i] = li;
    o->w[i] = (double)w;
  }
  o->val = val;
  for (int k = 0; k < 3; ++k) o->g[k] = g[k] / (double)vs;
  o->lo = I3{cv[0][0], cv[1][0], cv[2][0]};
  return true;
}

constexpr int kTanN = 11;  // omega xyz, upsilon xyz, eps_0..eps_4

struct PixAdj {
  double depth;
  double dd[kTanN];
  int eval_b[5][8], eval_l[5][8];
  double eval_w[5][8];
  int64_t path;
};

template <typename T>
void sample_tan(const V3<T>& x, const double* g, int slot, double* t) {
  const double X[3] = {(double)x.x, (double)x.y, (double)x.z};
  t[0] = X[1] * g[2] - X[2] * g[1];
  t[1] = X[2] * g[0] - X[0] * g[2];
  t[2] = X[0] * g[1] - X[1] * g[0];
  t[3] = g[0];
  t[4] = g[1];
  t[5] = g[2];
  for (int k = 6; k < kTanN; ++k) t[k] = k == slot ? 1.0 : 0.0;
}

uint64_t mix64(uint64_t h, uint64_t x) {
  h ^= x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h * 0xBF58476D1CE4E5B9ull;
}

template <typename T>
bool adj_pixel(const orc_volume* v, const mh_raycaster_params* rp, const T* wTc, uint32_t x,
               uint32_t y, PixAdj* out) {
  const T fx = rp->fx, fy = rp->fy, cx = rp->cx, cy = rp->cy;
  V3<T> dc{((T)x - cx) / fx, ((T)y - cy) / fy, T(1)};
  dc = V3<T>{T(1) * dc.x, T(1) * dc.y, T(1)};
  dc = nrm(dc);
  const T t_min = (T)rp->min_raycast_depth / dc.z;
  const T t_max = (T)rp->max_raycast_depth / dc.z;
  const T step = (T)rp->raycast_step;
  const V3<T> o{wTc[0] * T(0) + wTc[1] * T(0) + wTc[2] * T(0) + wTc[3] * T(1),
                wTc[4] * T(0) + wTc[5] * T(0) + wTc[6] * T(0) + wTc[7] * T(1),
                wTc[8] * T(0) + wTc[9] * T(0) + wTc[10] * T(0) + wTc[11] * T(1)};
  const V3<T> dir = nrm(V3<T>{wTc[0] * dc.x + wTc[1] * dc.y + wTc[2] * dc.z + wTc[3] * T(0),
                              wTc[4] * dc.x + wTc[5] * dc.y + wTc[6] * dc.z + wTc[7] * T(0),
                              wTc[8] * dc.x + wTc[9] * dc.y + wTc[10] * dc.z + wTc[11] * T(0)});
  auto at = [&](T t) { return V3<T>{o.x + t * dir.x, o.y + t * dir.y, o.z + t * dir.z}; };
  T ps = T(0), pt = T(0);
  bool pok = false;
  int64_t k = 0;
  for (T t = t_min; t < t_max; t += step, ++k) {
    TriG<T> s;
    if (!tri_eval(v, at(t), &s)) continue;
    if (pok && ps > T(0) && s.val < T(0)) {
      T l = pt, r = t, m = (l + r) * T(0.5), ls = ps, rs = s.val;
      bool found = true;
      for (int j = 0; j < 3; ++j) {
        m = l + (ls / (ls - rs)) * (r - l);
        TriG<T> q;
        if (!tri_eval(v, at(m), &q)) {
          found = false;
          break;
        }
        if (ls * q.val > T(0)) { l = m; ls = q.val; }
        else { r = m; rs = q.val; }
      }
      if (found && std::fabs(ps - s.val) < (T)rp->sample_sdf_threshold &&
          std::fabs(s.val) < (T)rp->sdf_threshold) {
        // accepted: replay the refinement with tangents
        TriG<T> e[5];
        V3<T> X[5] = {at(pt), at(t)};
        tri_eval(v, X[0], &e[0]);
        tri_eval(v, X[1], &e[1]);
        double dl[kTanN] = {0}, dr[kTanN] = {0}, dls[kTanN], drs[kTanN], dm[kTanN] = {0};
        sample_tan(X[0], e[0].g, 6, dls);
        sample_tan(X[1], e[1].g, 7, drs);
        T L = pt, Rr = t, LS = e[0].val, RS = e[1].val, M = T(0);
        uint64_t path = mix64(0, (uint64_t)k);
        for (int j = 0; j < 3; ++j) {
          const double D = (double)LS - (double)RS, q = (double)LS / D;
          M = L + (LS / (LS - RS)) * (Rr - L);
          for (int c = 0; c < kTanN; ++c) {
            const double dq = ((double)LS * drs[c] - (double)RS * dls[c]) / (D * D);
            dm[c] = dl[c] + dq * ((double)Rr - (double)L) + q * (dr[c] - dl[c]);
          }
          X[2 + j] = at(M);
          tri_eval(v, X[2 + j], &e[2 + j]);
          double dms[kTanN];
          sample_tan(X[2 + j], e[2 + j].g, 8 + j, dms);
          const double gd = e[2 + j].g[0] * (double)dir.x + e[2 + j].g[1] * (double)dir.y +
                            e[2 + j].g[2] * (double)dir.z;
          for (int c = 0; c < kTanN; ++c) dms[c] += gd * dm[c];
          const bool left = LS * e[2 + j].val > T(0);
          path = mix64(path, left ? 1u : 2u);
          if (left) {
            L = M;
            LS = e[2 + j].val;
            std::memcpy(dl, dm, sizeof(dl));
            std::memcpy(dls, dms, sizeof(dls));
          } else {
            Rr = M;
            RS = e[2 + j].val;
            std::memcpy(dr, dm, sizeof(dr));
            std::memcpy(drs, dms, sizeof(drs));
          }
        }
        for (int q = 0; q < 5; ++q) {
          path = mix64(path, (uint64_t)(uint32_t)e[q].lo.x);
          path = mix64(path, (uint64_t)(uint32_t)e[q].lo.y);
          path = mix64(path, (uint64_t)(uint32_t)e[q].lo.z);
          for (int c = 0; c < 8; ++c) {
            out->eval_b[q][c] = e[q].bi[c];
            out->eval_l[q][c] = e[q].li[c];
            out->eval_w[q][c] = e[q].w[c];
          }
        }
        out->depth = (double)(M * dc.z);
        for (int c = 0; c < kTanN; ++c) out->dd[c] = dm[c] * (double)dc.z;
        out->path = (int64_t)(path >> 1);
        return true;
      }
    }
    ps = s.val;
    pt = t;
    pok = true;
  }
  return false;
}

template <typename T>
void adj_image(const orc_volume* v, const mh_raycaster_params* rp, const T* wTc, const float* obs,
               double max_residual, double* depth_out, double* jac_out, int64_t* path_out,
               std::vector<double>* gsdf, double* summary) {
  double acc[30] = {0};
  const uint32_t W = rp->width, H = rp->height;
  for (uint32_t y = 0; y < H; ++y)
    for (uint32_t x = 0; x < W; ++x) {
      const uint32_t pix = y * W + x;
      PixAdj a;
      const bool hit = adj_pixel<T>(v, rp, wTc, x, y, &a);
      if (depth_out) depth_out[pix] = hit ? a.depth : -INFINITY;
      if (path_out) path_out[pix] = hit ? a.path : -1;
      if (jac_out)
        for (int c = 0; c < kTanN; ++c) jac_out[(size_t)pix * kTanN + c] = hit ? a.dd[c] : 0.0;
      if (!hit || !obs) continue;
      const float od = obs[pix];
      if (!(od > 0.0f && od < INFINITY)) continue;
      const double r = a.depth - (double)od;
      if (!(std::fabs((float)(a.depth - (double)od)) <= (float)max_residual)) {
        acc[29] += 1.0;  // valid but gated out
        continue;
      }
      acc[0] += 0.5 * r * r;
      for (int c = 0; c < 6; ++c) acc[1 + c] += r * a.dd[c];
      int t = 7;
      for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) acc[t++] += a.dd[i] * a.dd[j];
      acc[28] += 1.0;
      if (gsdf)
        for (int q = 0; q < 5; ++q)
          for (int c = 0; c < 8; ++c)
            (*gsdf)[(size_t)a.eval_b[q][c] * 512 + a.eval_l[q][c]] += r * a.dd[6 + q] * a.eval_w[q][c];
    }
  if (summary) std::memcpy(summary, acc, sizeof(acc));
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- primal-dual
// optimize/primal_dual.cu:10-223, geometry/primal_dual_query.h:15-234.
static bool below_eps(double w) { return (double)(float)w < 1e-6; }

static bool pd_voxel(const orc_volume* v, I3 vox, int* bi, int* li) { return voxel_at(v, vox, bi, li); }

int orc_pd_init(orc_volume* v) {
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      if (below_eps(B.w[i])) continue;
      const I3 vp = voxel_of(B, i);
      double g[3] = {0, 0, 0};
      bool ok = true;
      for (int k = 0; k < 3 && ok; ++k) {
        int bi, li;
        const I3 q{vp.x + (k == 0), vp.y + (k == 1), vp.z + (k == 2)};
        if (!pd_voxel(v, q, &bi, &li) || below_eps(v->blocks[bi].w[li])) { ok = false; break; }
        g[k] = v->blocks[bi].sdf[li] - B.sdf[i];
      }
      if (!ok) g[0] = g[1] = g[2] = 0.0;
      B.p[i][0] = B.p[i][1] = B.p[i][2] = 0.0;
      B.sdf_bar[i] = 0.0;
      B.wpd[i] = std::exp(std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) / (double)v->g.vs);
      B.sdf0[i] = B.sdf[i];
      B.mask[i] = 1;
    }
  }
  return 0;
}

int orc_pd_iterate(orc_volume* v, float lambda, float sigma, float tau, double* energy) {
  const double alpha = 0.02;
  double ed = 0.0, et = 0.0;
  auto valid_nb = [&](I3 q, int* bi, int* li) {
    return pd_voxel(v, q, bi, li) && !below_eps(v->blocks[*bi].w[*li]) && v->blocks[*bi].mask[*li];
  };
  // pass 1
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      if (below_eps(B.w[i])) continue;
      const I3 vp = voxel_of(B, i);
      double dd = std::fabs(B.sdf[i] - B.sdf0[i]);
      dd *= dd;
      if ((double)(float)B.w[i] > 1e-6) ed += dd;
      double gs[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
      bool ok = true;
      for (int k = 0; k < 3 && ok; ++k) {
        int bi, li;
        if (!valid_nb(I3{vp.x + (k == 0), vp.y + (k == 1), vp.z + (k == 2)}, &bi, &li)) { ok = false; break; }
        gs[k] = v->blocks[bi].sdf[li] - B.sdf[i];
        gb[k] = v->blocks[bi].sdf_bar[li] - B.sdf_bar[i];
      }
      if (!ok) for (int k = 0; k < 3; ++k) gs[k] = gb[k] = 0.0;
      const double gl = std::sqrt(gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]);
      et += (std::fabs(gl) < alpha) ? 0.5 * gl * gl / alpha : (std::fabs(gl) - 0.5 * alpha);
      double p[3];
      for (int k = 0; k < 3; ++k) p[k] = (B.p[i][k] + (double)sigma * gb[k]) / (1.0 + (double)sigma * alpha);
      const double pl = std::max(1.0, std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
      for (int k = 0; k < 3; ++k) B.p[i][k] = p[k] / pl;
    }
  }
  if (energy) {
    energy[0] = ed;
    energy[1] = et;
    energy[2] = ed * (double)lambda / 2.0 + et;
  }
  // pass 2
  // in-place is Jacobi here: the stencil reads only p, mask and weights
  for (int ptr : v->cand) {
    Block& B = v->blocks[ptr];
    for (int i = 0; i < 512; ++i) {
      if (below_eps(B.w[i])) continue;
      const I3 vp = voxel_of(B, i);
      double div = 0.0;
      if (B.mask[i]) {
        double pn[3];
        bool ok = true;
        for (int k = 0; k < 3 && ok; ++k) {
          int bi, li;
          if (!valid_nb(I3{vp.x - (k == 0), vp.y - (k == 1), vp.z - (k == 2)}, &bi, &li)) { ok = false; break; }
          pn[k] = v->blocks[bi].p[li][k];
        }
        if (ok) div = (pn[0] - B.p[i][0]) + (pn[1] - B.p[i][1]) + (pn[2] - B.p[i][2]);
      }
      const double prev = B.sdf[i];
      const double lam = (double)lambda * B.wpd[i];
      double s = prev - (double)tau * div;
      s = (s + lam * tau * B.sdf0[i]) / (1.0 + lam * tau);
      if (B.wpd[i] > 2.0) s = B.sdf0[i];
      B.sdf[i] = s;
      B.sdf_bar[i] = 2.0 * s - prev;
    }
  }
  return 0;
}

// ---------------------------------------------------------------- dumps
static std::vector<int> sorted_ptrs(const orc_volume* v) {
  std::vector<int> idx;
  for (size_t i = 0; i < v->blocks.size(); ++i)
    if (!v->blocks[i].freed) idx.push_back((int)i);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return v->blocks[a].pos < v->blocks[b].pos; });
  return idx;
}

uint32_t orc_block_count(const orc_volume* v) { return (uint32_t)v->where.size(); }

int orc_dump_blocks(const orc_volume* v, int32_t* pos, double* sdf, double* w, uint32_t* col,
                    double* a, double* b, uint16_t* cube) {
  const std::vector<int> idx = sorted_ptrs(v);
  for (size_t n = 0; n < idx.size(); ++n) {
    const Block& B = v->blocks[idx[n]];
    if (pos) { pos[3 * n] = B.pos.x; pos[3 * n + 1] = B.pos.y; pos[3 * n + 2] = B.pos.z; }
    for (int i = 0; i < 512; ++i) {
      const size_t o = n * 512 + i;
      if (sdf) sdf[o] = B.sdf[i];
      if (w) w[o] = B.w[i];
      if (col) col[o] = (uint32_t)B.col[i][0] | ((uint32_t)B.col[i][1] << 8) | ((uint32_t)B.col[i][2] << 16);
      if (a) a[o] = B.a[i];
      if (b) b[o] = B.b[i];
      if (cube) cube[o] = B.cube[i];
    }
  }
  return 0;
}

int orc_dump_pd(const orc_volume* v, double* pdv, uint8_t* mask) {
  const std::vector<int> idx = sorted_ptrs(v);
  for (size_t n = 0; n < idx.size(); ++n) {
    const Block& B = v->blocks[idx[n]];
    for (int i = 0; i < 512; ++i) {
      double* d = pdv + n * 6 * 512;
      d[i] = B.sdf0[i];
      d[512 + i] = B.sdf_bar[i];
      d[1024 + i] = B.wpd[i];
      d[1536 + i] = B.p[i][0];
      d[2048 + i] = B.p[i][1];
      d[2560 + i] = B.p[i][2];
      if (mask) mask[n * 512 + i] = B.mask[i];
    }
  }
  return 0;
}

// Load blocks (FP32 fields as given) in the listed order.
int orc_load_blocks(orc_volume* v, uint32_t n, const int32_t* pos, const float* sdf, const float* w,
                    const uint32_t* col, const float* a, const float* b) {
  for (uint32_t i = 0; i < n; ++i) {
    const I3 p{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    if (!bucket_insert(v, p)) list_insert(v, p);
    const int bi = find_block(v, p);
    if (bi < 0) return -1;
    Block& B = v->blocks[bi];
    for (int k = 0; k < 512; ++k) {
      const size_t o = (size_t)i * 512 + k;
      B.sdf[k] = sdf[o];
      B.w[k] = w[o];
      if (col) { B.col[k][0] = col[o] & 255; B.col[k][1] = (col[o] >> 8) & 255; B.col[k][2] = (col[o] >> 16) & 255; }
      if (a) B.a[k] = a[o];
      if (b) B.b[k] = b[o];
    }
  }
  return 0;
}

// Hash records: entry index, home bucket, in-bucket flag, pos.
uint32_t orc_dump_hash(const orc_volume* v, int32_t* rec) {
  uint32_t n = 0;
  for (size_t i = 0; i < v->entries.size(); ++i) {
    const Entry& e = v->entries[i];
    if (e.ptr == kFree) continue;
    if (rec) {
      const uint32_t home = bucket_of(e.pos, v->hp.bucket_count);
      int32_t* r = rec + 6 * (size_t)n;
      r[0] = (int32_t)i;
      r[1] = (int32_t)home;
      r[2] = (i / v->hp.bucket_size) == home ? 1 : 0;
      r[3] = e.pos.x;
      r[4] = e.pos.y;
      r[5] = e.pos.z;
    }
    ++n;
  }
  return n;
}

uint32_t orc_dump_candidates(const orc_volume* v, int32_t* pos) {
  std::vector<I3> ps;
  for (int p : v->cand) ps.push_back(v->blocks[p].pos);
  std::sort(ps.begin(), ps.end());
  if (pos)
    for (size_t i = 0; i < ps.size(); ++i) { pos[3 * i] = ps[i].x; pos[3 * i + 1] = ps[i].y; pos[3 * i + 2] = ps[i].z; }
  return (uint32_t)ps.size();
}

// Flags of the last low-surfel / garbage pass as (x, y, z, flag) rows in
// block-position order.
uint32_t orc_dump_cand_flags(const orc_volume* v, int32_t* rec) {
  std::vector<std::array<int32_t, 4>> r;
  for (size_t k = 0; k < v->cand.size() && k < v->cand_flag.size(); ++k) {
    const I3 p = v->blocks[v->cand[k]].pos;
    r.push_back({p.x, p.y, p.z, (int32_t)v->cand_flag[k]});
  }
  std::sort(r.begin(), r.end());
  if (rec) std::memcpy(rec, r.data(), r.size() * sizeof(r[0]));
  return (uint32_t)r.size();
}

uint32_t orc_tri_count(const orc_volume* v) { return (uint32_t)v->tris.size(); }
uint32_t orc_vtx_count(const orc_volume* v) { return (uint32_t)v->verts.size(); }

int orc_dump_mesh(const orc_volume* v, int32_t* tri, int32_t* tri_cube, int32_t* vkey, double* vval,
                  int32_t* vref) {
  size_t n = 0;
  for (const auto& kv : v->tris) {
    if (tri)
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 4; ++j) tri[12 * n + 4 * k + j] = kv.second[k].k[j];
    if (tri_cube) std::memcpy(tri_cube + 4 * n, kv.first.k, sizeof(kv.first.k));
    ++n;
  }
  n = 0;
  for (const auto& kv : v->verts) {
    if (vkey) std::memcpy(vkey + 4 * n, kv.first.k, sizeof(kv.first.k));
    if (vval) {
      double* d = vval + 10 * n;
      d[0] = kv.second.pos[0]; d[1] = kv.second.pos[1]; d[2] = kv.second.pos[2];
      d[3] = kv.second.normal[0]; d[4] = kv.second.normal[1]; d[5] = kv.second.normal[2];
      d[6] = kv.second.color[0]; d[7] = kv.second.color[1]; d[8] = kv.second.color[2];
      d[9] = kv.second.radius;
    }
    if (vref) vref[n] = kv.second.ref;
    ++n;
  }
  return 0;
}

void orc_inlier(const orc_volume* v, double* out) { std::memcpy(out, v->inlier.data(), v->inlier.size() * 8); }
void orc_depth(const orc_volume* v, float* out) { std::memcpy(out, v->depth.data(), v->depth.size() * 4); }
// Unnormalised GetSpatialSDFGradient at an FP32 point (FP64 values);
// returns 1 when every sample was valid.
int orc_spatial_gradient(const orc_volume* v, const float* pos, double* grad) {
  const float h = 0.5f * v->g.vs;
  const F3 p{pos[0], pos[1], pos[2]};
  double sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
  bool ok = true;
  for (int k = 0; k < 3 && ok; ++k) {
    F3 pm = p, pq = p;
    if (k == 0) { pm.x = p.x - h; pq.x = p.x + h; }
    if (k == 1) { pm.y = p.y - h; pq.y = p.y + h; }
    if (k == 2) { pm.z = p.z - h; pq.z = p.z + h; }
    SV q;
    ok = spatial64(v, pm, &q);
    sn[k] = q.sdf;
    if (!ok) break;
    ok = spatial64(v, pq, &q);
    sp[k] = q.sdf;
  }
  for (int k = 0; k < 3; ++k) grad[k] = (sp[k] - sn[k]) / v->g.vs;
  return ok ? 1 : 0;
}

uint32_t orc_dropped(const orc_volume* v) { return v->dropped; }
uint64_t orc_voxel_updates(const orc_volume* v) { return v->voxel_updates; }
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }
uint32_t orc_list_inserts(const orc_volume* v) { return v->list_inserts; }

// MC table access for the table tests.
// Raycast adjoint (see adj_image).  mode 0: FP32 march from cTw32 (wTc =
// FP32 inverse, as orc_raycast); mode 1: FP64 march from wTc64.  Outputs per
// pixel (optional): depth, 11 tangents (d depth / d(omega, upsilon, eps0..4)),
// path id (-1: no hit); grad_sdf: dL/dsdf per voxel in orc_dump_blocks
// order; summary: loss, dL/dxi[6], J^T J upper triangle[21], pixels, and
// the valid pixels left out by |residual| > max_residual.
int orc_raycast_adjoint(const orc_volume* v, const mh_raycaster_params* rp, int mode,
                        const float* cTw32, const double* wTc64, const float* obs,
                        double max_residual, double* depth_out, double* jac_out,
                        int64_t* path_out, double* grad_sdf, double* summary) {
  std::vector<double> g;
  if (grad_sdf) g.assign(v->blocks.size() * 512, 0.0);
  if (mode == 0) {
    float wTc[16];
    inv4<float>(cTw32, wTc);
    adj_image<float>(v, rp, wTc, obs, max_residual, depth_out, jac_out, path_out,
                     grad_sdf ? &g : nullptr, summary);
  } else {
    adj_image<double>(v, rp, wTc64, obs, max_residual, depth_out, jac_out, path_out,
                      grad_sdf ? &g : nullptr, summary);
  }
  if (grad_sdf) {
    const std::vector<int> order = sorted_ptrs(v);
    for (size_t i = 0; i < order.size(); ++i)
      std::memcpy(grad_sdf + i * 512, g.data() + (size_t)order[i] * 512, 512 * sizeof(double));
  }
  return 0;
}

// Adds delta to one voxel's SDF (block i of orc_dump_blocks order, voxel li).
int orc_perturb_sdf(orc_volume* v, uint32_t block, uint32_t li, double delta) {
  const std::vector<int> order = sorted_ptrs(v);
  if (block >= order.size() || li >= 512) return -1;
  v->blocks[order[block]].sdf[li] += delta;
  return 0;
}

// Test hooks for closed-form pins: ValToRGB itself, and a hand-placed
// vertex record in a voxel's x/y/z slot (the predict step reads a voxel's
// own three slots, update_bayesian.cu:76-95).
void orc_val_to_rgb(double v, double vmin, double vmax, double* out) { val_to_rgb(v, vmin, vmax, out); }
int orc_set_vertex(orc_volume* v, const int32_t* key, const float* pos, const double* normal,
                   double radius) {
  if (key[3] < 0 || key[3] > 2) return -1;
  Vtx vt;
  for (int k = 0; k < 3; ++k) {
    vt.pos[k] = pos[k];
    vt.normal[k] = normal[k];
    vt.color[k] = 0.0;
  }
  vt.radius = radius;
  v->verts[VKey{{key[0], key[1], key[2], key[3]}}] = vt;
  return 0;
}

// Replaces the payload (position, normal, colour, radius) of n vertices,
// keyed like orc_dump_mesh, keeping their reference counts: lockstep tests
// hand the library's vertices to the predict step.
int orc_load_vertices(orc_volume* v, uint32_t n, const int32_t* keys, const double* vals) {
  for (uint32_t i = 0; i < n; ++i) {
    const int32_t* k = keys + 4 * (size_t)i;
    if (k[3] < 0 || k[3] > 2) return -1;
    Vtx& vt = v->verts[VKey{{k[0], k[1], k[2], k[3]}}];
    const double* d = vals + 10 * (size_t)i;
    for (int j = 0; j < 3; ++j) {
      vt.pos[j] = (float)d[j];
      vt.normal[j] = d[3 + j];
      vt.color[j] = d[6 + j];
    }
    vt.radius = d[9];
  }
  return 0;
}

int orc_mc_ntri(int cube) { return ntri(cube); }
int orc_mc_tri_edge(int cube, int k) { return tri_edge(cube, k); }
uint32_t orc_mc_edge_mask(int cube) { return edge_mask(cube); }
int orc_mc_corner(int i, int ax) { return corner(i, ax); }
int orc_mc_edge_a(int e) { return ea(e); }
int orc_mc_edge_b(int e) { return eb(e); }

}  // extern "C"
