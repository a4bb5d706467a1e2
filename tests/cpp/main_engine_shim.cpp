// main_engine_shim.cpp -- MeshHashing's MainEngine stages over the C ABI,
// compiled as a reference-side maintainer would compile them: plain C++ with
// g++, against include/meshhashing.h only, linked to libmeshhashing_amd.so.
// The class keeps the reference's method names and schedule
// (engine/main_engine.cc:100-192 Mapping / Meshing / Recycle, :231-239 and
// visualizing_engine.cc:262-274 RenderRayCaster, :299-311 FinalLog).
//
// Test driver: reads the parameter PODs from <dir>/config.bin (the header's
// structs, byte for byte), runs the seeded synthetic sequence through the
// engine, and writes the fused blocks, the status, the last frame's raycast
// and the final PLY mesh under <dir>.  Only the driver touches HIP (device
// buffers for the raycast images); the engine class does not.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "meshhashing.h"

namespace {

void mh_check(int rc) {  // the reference aborts on CUDA errors
  if (rc != MH_OK && rc != MH_ERR_EMPTY) {
    std::fprintf(stderr, "meshhashing: %d %s\n", rc, mh_last_error());
    std::abort();
  }
}

class MainEngine {
 public:
  MainEngine(const mh_hash_params& hp, const mh_volume_params& vp, const mh_mesh_params& mp,
             const mh_sensor_params& sp, const mh_raycaster_params& rp, const mh_mc_params& mc,
             bool map_psdf)
      : rp_(rp), mc_(mc), map_psdf_(map_psdf) {
    mh_check(mh_volume_create(&hp, &vp, &mp, &sp, map_psdf ? 1 : 0, &vol_));
  }
  ~MainEngine() { mh_volume_destroy(vol_); }
  MainEngine(const MainEngine&) = delete;
  MainEngine& operator=(const MainEngine&) = delete;

  // main_engine.cc:100-154
  void Mapping(const uint16_t* depth_u16_host, const float wTc[16]) {
    mh_check(mh_frame_upload_depth_u16(vol_, depth_u16_host, wTc));
    mh_check(mh_alloc_blocks(vol_));
    mh_check(mh_collect_visible(vol_, nullptr));
    if (map_psdf_) {
      mh_check(mh_predict_inlier(vol_));
      mh_check(mh_integrate_psdf(vol_));
    } else {
      mh_check(mh_integrate_tsdf(vol_));
    }
    ++integrated_frame_count_;
  }
  // main_engine.cc:156-174
  void Meshing() {
    mh_check(mh_marching_cubes(vol_, &mc_));
    mh_check(mh_collect_low_surfel_blocks(vol_));
    if (integrated_frame_count_ % 10 == 0) mh_check(mh_recycle_garbage_blocks(vol_, nullptr));
  }
  // main_engine.cc:176-192
  void Recycle() {
    if (!map_psdf_ && integrated_frame_count_ % 15 == 14) {
      mh_check(mh_starve_blocks(vol_));
      mh_check(mh_collect_garbage_blocks(vol_));
      mh_check(mh_recycle_garbage_blocks(vol_, nullptr));
    }
  }
  // visualizing_engine.cc:262-274
  void RenderRayCaster(const float cTw[16], float* depth_dev, float* vertex_dev, float* normal_dev) {
    mh_check(mh_raycast(vol_, &rp_, cTw, depth_dev, vertex_dev, normal_dev));
  }
  // main_engine.cc:299-311
  void FinalLog(const std::string& ply_path) {
    mh_check(mh_collect_all(vol_, nullptr));
    mh_check(mh_marching_cubes(vol_, &mc_));
    mh_check(mh_compress_mesh(vol_, nullptr, nullptr));
    mh_check(mh_save_ply(vol_, ply_path.c_str()));
  }
  mh_volume* volume() { return vol_; }

 private:
  mh_volume* vol_ = nullptr;
  mh_raycaster_params rp_;
  mh_mc_params mc_;
  bool map_psdf_;
  int integrated_frame_count_ = 0;
};

template <typename T>
void read_pod(FILE* f, T* out) {
  if (std::fread(out, sizeof(T), 1, f) != 1) {
    std::fprintf(stderr, "config.bin too short\n");
    std::exit(2);
  }
}

template <typename T>
void write_file(const std::string& path, const std::vector<T>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <dir with config.bin>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  FILE* f = std::fopen((dir + "/config.bin").c_str(), "rb");
  if (!f) {
    std::fprintf(stderr, "no config.bin in %s\n", dir.c_str());
    return 2;
  }
  mh_hash_params hp;
  mh_volume_params vp;
  mh_mesh_params mp;
  mh_sensor_params sp;
  mh_raycaster_params rp;
  mh_mc_params mc;
  mh_synth_spec ss;
  uint32_t frames = 0, step = 1;
  read_pod(f, &hp);
  read_pod(f, &vp);
  read_pod(f, &mp);
  read_pod(f, &sp);
  read_pod(f, &rp);
  read_pod(f, &mc);
  read_pod(f, &ss);
  read_pod(f, &frames);
  read_pod(f, &step);
  std::fclose(f);

  MainEngine engine(hp, vp, mp, sp, rp, mc, /*map_psdf=*/false);
  std::vector<uint16_t> depth((size_t)ss.width * ss.height);
  float wTc[16];
  for (uint32_t k = 0; k < frames; ++k) {
    const uint32_t fr = k * step;
    mh_check(mh_synth_pose(&ss, fr, wTc));
    mh_check(mh_synth_depth(&ss, fr, depth.data()));
    engine.Mapping(depth.data(), wTc);
    engine.Meshing();
    engine.Recycle();
    std::printf("frame %u done\n", fr);
    std::fflush(stdout);
  }
  // raycast at the last frame's pose (the volume's own FP32 inverse of it,
  // as Sensor::set_transform keeps cTw beside wTc)
  float cTw[16];
  mh_check(mh_frame_get_pose(engine.volume(), wTc, cTw));
  const size_t npix = (size_t)rp.width * rp.height;
  float *d_dev = nullptr, *n_dev = nullptr;
  if (hipMalloc((void**)&d_dev, npix * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&n_dev, 3 * npix * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 3;
  }
  engine.RenderRayCaster(cTw, d_dev, nullptr, n_dev);
  mh_check(mh_synchronize(engine.volume()));
  std::vector<float> dimg(npix), nimg(3 * npix);
  if (hipMemcpy(dimg.data(), d_dev, npix * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(nimg.data(), n_dev, 3 * npix * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return 3;
  (void)hipFree(d_dev);
  (void)hipFree(n_dev);
  write_file(dir + "/ray_depth.bin", dimg);
  write_file(dir + "/ray_normal.bin", nimg);
  write_file(dir + "/cTw.bin", std::vector<float>(cTw, cTw + 16));

  // fused blocks, sorted by position
  uint32_t nb = 0;
  mh_check(mh_dump_blocks(engine.volume(), 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &nb));
  std::vector<int32_t> pos(3 * (size_t)nb);
  std::vector<float> sdf((size_t)nb * 512), w((size_t)nb * 512);
  std::vector<uint32_t> col((size_t)nb * 512);
  mh_check(mh_dump_blocks(engine.volume(), nb, pos.data(), sdf.data(), w.data(), col.data(), nullptr,
                          nullptr, &nb));
  write_file(dir + "/pos.bin", pos);
  write_file(dir + "/sdf.bin", sdf);
  write_file(dir + "/weight.bin", w);
  write_file(dir + "/color.bin", col);
  mh_status st;
  mh_check(mh_get_status(engine.volume(), &st));
  write_file(dir + "/status.bin", std::vector<mh_status>{st});
  engine.FinalLog(dir + "/mesh.ply");
  std::printf("blocks %u triangles %u\n", nb, st.triangle_count);
  return 0;
}
