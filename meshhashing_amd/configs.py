"""Workload configurations (the reference's YAML parameter sets, restated).

Each config carries the five parameter PODs of core/params.h:16-74, the
marching-cubes thresholds, and the synthetic sequence that feeds it.

  c1  plain TSDF + MC, 320x240, 2 cm voxels, trunc 8 cm, 2^17 x 10 hash
  c2  full-rate plain pipeline, 640x480, 5 mm voxels, trunc 2 cm, 2^20 x 10
  c3  PSDF + primal-dual: c1 with 10 % outliers, MC keeps rho >= 0.4,
      50 PD iterations (lambda 0.5, tau = sigma = 0.25)
  c5  raycast volume: c1's scene fused at 5 mm, raycast 640x480

Depth valid range 0.1-5 m and sdf_upper_bound 5 m for all configs; the
PSDF outlier term's 1/(5.0 - 0.1) (mapping/update_bayesian.cu:167-168)
assumes exactly that range.  Plain-TSDF configs mesh with weight_min = 1
(at least one observation): the reference's (1/voxel_size)^2 threshold
(meshing/marching_cubes.cu:126-128) is tuned for PSDF precisions and would
reject almost every plain-TSDF cube.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

from . import _capi as A


@dataclass
class Config:
    name: str
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    voxel_size: float
    truncation: float
    bucket_count: int
    value_capacity: int
    max_vertices: int
    max_triangles: int
    frames: int
    synth_kind: int
    seed: int = 1
    psdf: bool = False
    hole_fraction: float = 0.0
    outlier_fraction: float = 0.0
    bucket_size: int = 10
    linked_list_size: int = 7
    min_depth: float = 0.1
    max_depth: float = 5.0
    sdf_upper_bound: float = 5.0
    weight_sample: int = 10
    raycast_step: float = 0.0  # 0: half the truncation distance
    mc_weight_min: float = 1.0
    mc_rho_min: float = 0.1
    pd_iterations: int = 0
    pd_lambda: float = 0.5
    pd_sigma: float = 0.25
    pd_tau: float = 0.25
    extra: dict = field(default_factory=dict)

    # -- PODs ------------------------------------------------------------
    def hash_params(self) -> A.HashParams:
        return A.HashParams(self.bucket_count, self.bucket_size,
                            self.bucket_count * self.bucket_size, self.linked_list_size,
                            self.value_capacity)

    def volume_params(self) -> A.VolumeParams:
        return A.VolumeParams(self.voxel_size, 0.0, self.truncation, self.sdf_upper_bound,
                              self.weight_sample, 255)

    def mesh_params(self) -> A.MeshParams:
        return A.MeshParams(self.max_vertices, self.max_triangles)

    def sensor_params(self) -> A.SensorParams:
        return A.SensorParams(self.fx, self.fy, self.cx, self.cy, self.width, self.height,
                              self.min_depth, self.max_depth, 0.001)

    def raycaster_params(self) -> A.RayCasterParams:
        step = self.raycast_step or 0.5 * self.truncation
        return A.RayCasterParams(self.fx, self.fy, self.cx, self.cy, self.width, self.height,
                                 self.min_depth, self.max_depth, step, 0.808, 0.800, 1)

    def mc_params(self) -> A.McParams:
        return A.McParams(0.20, self.mc_rho_min, self.mc_weight_min, 0.008, 1)

    def synth_spec(self) -> A.SynthSpec:
        return A.SynthSpec(self.seed, self.synth_kind, self.width, self.height, self.fx, self.fy,
                           self.cx, self.cy, self.frames, self.hole_fraction,
                           self.outlier_fraction, 1)


C1 = Config("c1", 320, 240, 262.5, 262.5, 160.0, 120.0, 0.02, 0.08, 1 << 17, 1 << 15,
            1 << 20, 1 << 21, 20, 0)
C2 = Config("c2", 640, 480, 525.0, 525.0, 319.5, 239.5, 0.005, 0.02, 1 << 20, 1 << 19,
            8 << 20, 16 << 20, 300, 1, hole_fraction=0.02)
C3 = replace(C1, name="c3", psdf=True, outlier_fraction=0.10, mc_weight_min=-1.0,
             mc_rho_min=0.4, pd_iterations=50)
C5 = replace(C1, name="c5", width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5,
             voxel_size=0.005, truncation=0.02, bucket_count=1 << 20, value_capacity=1 << 19,
             max_vertices=1 << 20, max_triangles=1 << 21)

# Coarse copies of C5's scene (same frames, 4 cm and 2 cm voxels, truncation
# four voxels) for coarse-to-fine pose refinement, and the Gauss-Newton steps
# per level (4 cm, 2 cm, 5 mm): ten in all.
C5_PYRAMID = (replace(C5, name="c5_4cm", voxel_size=0.04, truncation=0.16, bucket_count=1 << 16,
                      value_capacity=1 << 13),
              replace(C5, name="c5_2cm", voxel_size=0.02, truncation=0.08, bucket_count=1 << 17,
                      value_capacity=1 << 15))
C5_PYRAMID_STEPS = (3, 3, 4)

CONFIGS = {c.name: c for c in (C1, C2, C3, C5)}

# Batched form: independent rooms of 320x240 frames, voxel sizes drawn from
# {1, 2, 3, 4} cm (seeded), truncation 4 voxels, each with its own seed.
C4_VOXEL_SIZES = (0.01, 0.02, 0.03, 0.04)
C4_CAPACITY = {0.01: 1 << 15, 0.02: 1 << 14, 0.03: 1 << 13, 0.04: 1 << 13}
C4_VERTICES = {0.01: 1 << 21, 0.02: 1 << 20, 0.03: 1 << 19, 0.04: 1 << 18}


def c4_scenes(n: int = 64, frames: int = 4, seed: int = 4) -> list[Config]:
    import numpy as np

    pick = np.random.default_rng(seed).integers(0, len(C4_VOXEL_SIZES), n)
    out = []
    for i in range(n):
        vs = C4_VOXEL_SIZES[int(pick[i])]
        out.append(replace(C1, name=f"c4_{i}", voxel_size=vs, truncation=4 * vs,
                           bucket_count=1 << 15, value_capacity=C4_CAPACITY[vs],
                           max_vertices=C4_VERTICES[vs], max_triangles=2 * C4_VERTICES[vs],
                           frames=frames,
                           seed=1000 + i))
    return out


def small(cfg: Config, **kw) -> Config:
    """A reduced copy for quick tests (same geometry rules, fewer resources)."""
    return replace(cfg, **kw)
