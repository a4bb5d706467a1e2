// synth.cpp -- seeded synthetic RGB-D depth sequences (host code).
//
// Scenes are analytic: an axis-aligned 4 x 3 x 4 m room seen from inside,
// centred on the origin, with spheres and axis-aligned boxes in it.  World
// axes: x right, y down, z forward, the camera convention of the pinhole
// model in geometry/geometry_helper.h:141-168.  Depth is the exact FP64 ray
// hit, perturbed by sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 Gaussian noise and
// rounded to millimetres.  All randomness comes from SplitMix64 keyed on
// (seed, frame, pixel), so a frame is the same whatever the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/meshhashing.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(splitmix(seed ^ 0xA0761D6478BD642Full)) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in (0, 1]
  double uniform() { return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 unit(V3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }
double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct Scene {
  V3 room_min{-2.0, -1.5, -2.0}, room_max{2.0, 1.5, 2.0};
  std::vector<V3> sph_c;
  std::vector<double> sph_r;
  std::vector<V3> box_min, box_max;
};

struct Pose {
  double R[9];  // camera axes as columns
  V3 t;
};

// Distance from p to the horizontal unit circle around the origin.
double dist_to_circle(V3 p) {
  const double rho = std::sqrt(p.x * p.x + p.z * p.z);
  return std::sqrt((rho - 1.0) * (rho - 1.0) + p.y * p.y);
}

Scene make_scene(const mh_synth_spec& s) {
  Scene sc;
  Rng rng(s.seed * 7919u + 17u);
  if (s.kind == 0) {
    for (int i = 0; i < 3; ++i) {
      for (int tries = 0; tries < 10000; ++tries) {
        const double r = rng.uniform(0.3, 0.6);
        const V3 c{rng.uniform(sc.room_min.x + r + 0.05, sc.room_max.x - r - 0.05),
                   rng.uniform(sc.room_min.y + r + 0.05, sc.room_max.y - r - 0.05),
                   rng.uniform(sc.room_min.z + r + 0.05, sc.room_max.z - r - 0.05)};
        if (dist_to_circle(c) < r + 0.25) continue;
        bool clash = false;
        for (size_t j = 0; j < sc.sph_c.size(); ++j)
          if (std::sqrt(dot(c - sc.sph_c[j], c - sc.sph_c[j])) < r + sc.sph_r[j] + 0.05) clash = true;
        if (clash) continue;
        sc.sph_c.push_back(c);
        sc.sph_r.push_back(r);
        break;
      }
    }
  } else {
    // 4 spheres + 4 boxes in a ring between 1.1 m and 1.8 m from the centre.
    for (int i = 0; i < 8; ++i) {
      for (int tries = 0; tries < 10000; ++tries) {
        const double size = rng.uniform(0.15, 0.35);
        const double ang = rng.uniform(0.0, 2.0 * kPi);
        const double rad = rng.uniform(1.1 + size, 1.85 - size);
        const V3 c{rad * std::cos(ang), rng.uniform(-1.0, 1.0), rad * std::sin(ang)};
        bool clash = false;
        for (size_t j = 0; j < sc.sph_c.size(); ++j)
          if (std::sqrt(dot(c - sc.sph_c[j], c - sc.sph_c[j])) < size + sc.sph_r[j] + 0.1) clash = true;
        for (size_t j = 0; j < sc.box_min.size(); ++j) {
          const V3 bc = (sc.box_min[j] + sc.box_max[j]) * 0.5;
          const double bh = (sc.box_max[j].x - sc.box_min[j].x) * 0.5;
          if (std::sqrt(dot(c - bc, c - bc)) < size + bh * 1.8 + 0.1) clash = true;
        }
        if (clash) continue;
        if (i % 2 == 0) {
          sc.sph_c.push_back(c);
          sc.sph_r.push_back(size);
        } else {
          sc.box_min.push_back(c - V3{size, size, size});
          sc.box_max.push_back(c + V3{size, size, size});
        }
        break;
      }
    }
  }
  return sc;
}

Pose look_pose(V3 pos, V3 fwd) {
  const V3 z = unit(fwd);
  const V3 down{0.0, 1.0, 0.0};
  const V3 y = unit(down - z * dot(down, z));
  const V3 x = cross(y, z);
  Pose p;
  p.R[0] = x.x; p.R[1] = y.x; p.R[2] = z.x;
  p.R[3] = x.y; p.R[4] = y.y; p.R[5] = z.y;
  p.R[6] = x.z; p.R[7] = y.z; p.R[8] = z.z;
  p.t = pos;
  return p;
}

double catmull(double p0, double p1, double p2, double p3, double u) {
  const double u2 = u * u, u3 = u2 * u;
  return 0.5 * ((2.0 * p1) + (-p0 + p2) * u + (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) * u2 +
                (-p0 + 3.0 * p1 - 3.0 * p2 + p3) * u3);
}

// Spline trajectory: knots 1 s apart (30 frames at 30 Hz), positions in a
// 0.5 m disc around the centre, yaw changing <= 25 deg and pitch within
// +-8 deg per knot; positions and angles are then scaled down until the
// sampled speed is <= 0.5 m/s and the angular rate <= 30 deg/s.
struct Spline {
  std::vector<double> kx, ky, kz, kyaw, kpitch;
  double pos_scale = 1.0, ang_scale = 1.0;
};

Spline make_spline(const mh_synth_spec& s) {
  Spline sp;
  Rng rng(s.seed * 104729u + 3u);
  const int knots = (int)(s.frames / 30u) + 4;
  double yaw = rng.uniform(0.0, 2.0 * kPi);
  for (int i = 0; i < knots; ++i) {
    const double ang = rng.uniform(0.0, 2.0 * kPi), rad = 0.5 * std::sqrt(rng.uniform());
    sp.kx.push_back(rad * std::cos(ang));
    sp.kz.push_back(rad * std::sin(ang));
    sp.ky.push_back(rng.uniform(-0.15, 0.15));
    yaw += rng.uniform(-25.0, 25.0) * kPi / 180.0;
    sp.kyaw.push_back(yaw);
    sp.kpitch.push_back(rng.uniform(-8.0, 8.0) * kPi / 180.0);
  }
  return sp;
}

Pose spline_pose(const Spline& sp, double tsec) {
  const int i = (int)std::floor(tsec) + 1;
  const double u = tsec - std::floor(tsec);
  auto ev = [&](const std::vector<double>& k) { return catmull(k[i - 1], k[i], k[i + 1], k[i + 2], u); };
  const V3 pos = V3{ev(sp.kx), ev(sp.ky), ev(sp.kz)} * sp.pos_scale;
  const double yaw0 = sp.kyaw[1];
  const double yaw = yaw0 + (ev(sp.kyaw) - yaw0) * sp.ang_scale;
  const double pitch = ev(sp.kpitch) * sp.ang_scale;
  const V3 fwd{std::cos(pitch) * std::sin(yaw), std::sin(pitch), std::cos(pitch) * std::cos(yaw)};
  return look_pose(pos, fwd);
}

double rot_angle(const Pose& a, const Pose& b) {
  double tr = 0.0;  // trace(Ra^T Rb)
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) tr += a.R[3 * r + c] * b.R[3 * r + c];
  return std::acos(std::max(-1.0, std::min(1.0, (tr - 1.0) * 0.5)));
}

Spline calibrated_spline(const mh_synth_spec& s) {
  Spline sp = make_spline(s);
  const double dt = 1.0 / 30.0;
  for (int iter = 0; iter < 20; ++iter) {
    double vmax = 0.0, wmax = 0.0;
    Pose prev = spline_pose(sp, 0.0);
    for (uint32_t f = 1; f < std::max(2u, s.frames); ++f) {
      const Pose cur = spline_pose(sp, f * dt);
      const V3 d = cur.t - prev.t;
      vmax = std::max(vmax, std::sqrt(dot(d, d)) / dt);
      wmax = std::max(wmax, rot_angle(prev, cur) / dt);
      prev = cur;
    }
    bool ok = true;
    if (vmax > 0.5) {
      sp.pos_scale *= 0.49 / vmax;
      ok = false;
    }
    if (wmax > 30.0 * kPi / 180.0) {
      sp.ang_scale *= (29.0 * kPi / 180.0) / wmax;
      ok = false;
    }
    if (ok) break;
  }
  return sp;
}

Pose frame_pose(const mh_synth_spec& s, uint32_t frame) {
  if (s.kind == 0) {
    const uint32_t n = std::max(1u, s.frames);
    const double th = 2.0 * kPi * (double)frame / (double)n;
    const V3 pos{std::cos(th), 0.0, std::sin(th)};
    return look_pose(pos, V3{0.0, 0.0, 0.0} - pos);
  }
  const Spline sp = calibrated_spline(s);
  return spline_pose(sp, frame / 30.0);
}

// Nearest positive ray hit (parameter s along p = o + s d).
double trace(const Scene& sc, V3 o, V3 d) {
  double best = INFINITY;
  for (int i = 0; i < 3; ++i) {
    const double di = comp(d, i), oi = comp(o, i);
    if (di > 0.0) best = std::min(best, (comp(sc.room_max, i) - oi) / di);
    else if (di < 0.0) best = std::min(best, (comp(sc.room_min, i) - oi) / di);
  }
  for (size_t k = 0; k < sc.sph_c.size(); ++k) {
    const V3 oc = o - sc.sph_c[k];
    const double a = dot(d, d), b = 2.0 * dot(d, oc), c = dot(oc, oc) - sc.sph_r[k] * sc.sph_r[k];
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) continue;
    const double sq = std::sqrt(disc);
    double t = (-b - sq) / (2.0 * a);
    if (t <= 0.0) t = (-b + sq) / (2.0 * a);
    if (t > 0.0) best = std::min(best, t);
  }
  for (size_t k = 0; k < sc.box_min.size(); ++k) {
    double t0 = -INFINITY, t1 = INFINITY;
    bool miss = false;
    for (int i = 0; i < 3; ++i) {
      const double di = comp(d, i), oi = comp(o, i);
      const double lo = comp(sc.box_min[k], i), hi = comp(sc.box_max[k], i);
      if (di == 0.0) {
        if (oi < lo || oi > hi) miss = true;
        continue;
      }
      double a = (lo - oi) / di, b = (hi - oi) / di;
      if (a > b) std::swap(a, b);
      t0 = std::max(t0, a);
      t1 = std::min(t1, b);
    }
    if (miss || t1 < t0) continue;
    if (t0 > 0.0) best = std::min(best, t0);
    else if (t1 > 0.0) best = std::min(best, t1);
  }
  return best;
}

void pose_to_wtc(const Pose& p, float* wTc) {
  const double m[16] = {p.R[0], p.R[1], p.R[2], p.t.x, p.R[3], p.R[4], p.R[5], p.t.y,
                        p.R[6], p.R[7], p.R[8], p.t.z, 0.0,    0.0,    0.0,    1.0};
  for (int i = 0; i < 16; ++i) wTc[i] = (float)m[i];
}

Pose wtc_to_pose(const float* wTc) {
  Pose p;
  p.R[0] = wTc[0]; p.R[1] = wTc[1]; p.R[2] = wTc[2];
  p.R[3] = wTc[4]; p.R[4] = wTc[5]; p.R[5] = wTc[6];
  p.R[6] = wTc[8]; p.R[7] = wTc[9]; p.R[8] = wTc[10];
  p.t = V3{wTc[3], wTc[7], wTc[11]};
  return p;
}

template <typename F>
void parallel_rows(uint32_t h, F&& fn) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (uint32_t y = t; y < h; y += nt) fn(y);
    });
  for (auto& x : th) x.join();
}

bool spec_ok(const mh_synth_spec* s) {
  return s && s->width > 0 && s->height > 0 && s->width <= 16384 && s->height <= 16384 &&
         (s->kind == 0 || s->kind == 1) && s->fx > 0 && s->fy > 0;
}

}  // namespace

extern "C" {

int mh_synth_pose(const mh_synth_spec* s, uint32_t frame, float wTc[16]) {
  if (!spec_ok(s) || !wTc) return MH_ERR_ARG;
  pose_to_wtc(frame_pose(*s, frame), wTc);
  return MH_OK;
}

int mh_synth_depth_exact(const mh_synth_spec* s, uint32_t frame, const float wTc[16],
                         double* depth_host) {
  (void)frame;
  if (!spec_ok(s) || !wTc || !depth_host) return MH_ERR_ARG;
  const Scene sc = make_scene(*s);
  const Pose p = wtc_to_pose(wTc);
  parallel_rows(s->height, [&](uint32_t y) {
    for (uint32_t x = 0; x < s->width; ++x) {
      const V3 dc{((double)x - s->cx) / s->fx, ((double)y - s->cy) / s->fy, 1.0};
      const V3 dw{p.R[0] * dc.x + p.R[1] * dc.y + p.R[2] * dc.z,
                  p.R[3] * dc.x + p.R[4] * dc.y + p.R[5] * dc.z,
                  p.R[6] * dc.x + p.R[7] * dc.y + p.R[8] * dc.z};
      depth_host[(size_t)y * s->width + x] = trace(sc, p.t, dw);
    }
  });
  return MH_OK;
}

int mh_synth_depth(const mh_synth_spec* s, uint32_t frame, uint16_t* depth_host) {
  if (!spec_ok(s) || !depth_host) return MH_ERR_ARG;
  float wTc[16];
  pose_to_wtc(frame_pose(*s, frame), wTc);
  std::vector<double> exact((size_t)s->width * s->height);
  int rc = mh_synth_depth_exact(s, frame, wTc, exact.data());
  if (rc != MH_OK) return rc;
  parallel_rows(s->height, [&](uint32_t y) {
    for (uint32_t x = 0; x < s->width; ++x) {
      const size_t i = (size_t)y * s->width + x;
      double z = exact[i];
      const uint64_t key = splitmix(s->seed ^ splitmix(((uint64_t)frame << 32) ^ (uint64_t)i));
      Rng r(key);
      const double u_hole = r.uniform(), u_out = r.uniform(), u1 = r.uniform(), u2 = r.uniform();
      uint16_t mm = 0;
      if (!(u_hole <= (double)s->hole_fraction) && std::isfinite(z)) {
        if (u_out <= (double)s->outlier_fraction) {
          z = 0.3 + 4.7 * u1;
        } else if (s->noise) {
          const double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
          z += sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
        }
        const double v = std::floor(z * 1000.0 + 0.5);
        mm = (uint16_t)std::max(0.0, std::min(65535.0, v));
      }
      depth_host[i] = mm;
    }
  });
  return MH_OK;
}

// Colour image of frame f: a smooth procedural texture of the world hit
// point, 8-bit BGRA (alpha 255), every channel in [28, 228]; 0 where the ray
// leaves the room (no surface).
int mh_synth_color(const mh_synth_spec* s, uint32_t frame, uint8_t* bgra_host) {
  if (!spec_ok(s) || !bgra_host) return MH_ERR_ARG;
  const Scene sc = make_scene(*s);
  const Pose p = frame_pose(*s, frame);
  parallel_rows(s->height, [&](uint32_t y) {
    for (uint32_t x = 0; x < s->width; ++x) {
      const V3 dc{((double)x - s->cx) / s->fx, ((double)y - s->cy) / s->fy, 1.0};
      const V3 dw{p.R[0] * dc.x + p.R[1] * dc.y + p.R[2] * dc.z,
                  p.R[3] * dc.x + p.R[4] * dc.y + p.R[5] * dc.z,
                  p.R[6] * dc.x + p.R[7] * dc.y + p.R[8] * dc.z};
      const double t = trace(sc, p.t, dw);
      uint8_t* o = bgra_host + 4 * ((size_t)y * s->width + x);
      if (!std::isfinite(t)) {
        o[0] = o[1] = o[2] = o[3] = 0;
        continue;
      }
      const V3 X{p.t.x + t * dw.x, p.t.y + t * dw.y, p.t.z + t * dw.z};
      auto ch = [](double v) { return (uint8_t)std::floor(128.0 + 100.0 * std::sin(v) + 0.5); };
      o[0] = ch(4.0 * X.z + 2.0);  // B
      o[1] = ch(4.0 * X.y + 1.0);  // G
      o[2] = ch(4.0 * X.x);        // R
      o[3] = 255;
    }
  });
  return MH_OK;
}

}  // extern "C"
