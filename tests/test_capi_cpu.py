"""C ABI checks that need no GPU: the library loads, exports every symbol
include/meshhashing.h declares, and its host-side pieces (parameter
defaults, the synthetic sequence generator) behave."""
import ctypes as C
import re
from dataclasses import replace
from pathlib import Path

import numpy as np

from meshhashing_amd import _capi as A
from meshhashing_amd.configs import C1, C2
from meshhashing_amd.engine import synth_depth, synth_depth_exact, synth_pose

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "meshhashing.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(A.SIGNATURES), set(names) ^ set(A.SIGNATURES)


def test_struct_layouts():
    assert C.sizeof(A.HashParams) == 20
    assert C.sizeof(A.VolumeParams) == 24
    assert C.sizeof(A.SensorParams) == 36
    assert C.sizeof(A.RayCasterParams) == 48
    assert C.sizeof(A.McParams) == 32
    assert C.sizeof(A.Status) == 88
    assert C.sizeof(A.RaycastGrad) == 360


def test_mc_defaults_are_the_reference_constants():
    p = A.McParams()
    A.load().mh_mc_params_default(C.byref(p))
    assert abs(p.sdf_threshold - 0.20) < 1e-7 and abs(p.rho_min - 0.1) < 1e-7
    assert p.snap_distance == 0.008 and p.weight_min < 0 and p.enable_sdf_gradient == 1


def test_synth_deterministic_and_close_to_exact():
    cfg = replace(C1, frames=20)
    d0 = synth_depth(cfg, 3)
    d1 = synth_depth(cfg, 3)
    assert np.array_equal(d0, d1)
    pose = synth_pose(cfg, 3)
    ex = synth_depth_exact(cfg, pose)
    valid = d0 > 0
    assert valid.mean() > 0.95
    err = np.abs(d0[valid] * 1e-3 - ex[valid])
    sig = 0.0012 + 0.0019 * (ex[valid] - 0.4) ** 2
    assert np.mean(err < 4 * sig + 0.0006) > 0.999
    R = pose[:3, :3].astype(np.float64)
    assert np.abs(R.T @ R - np.eye(3)).max() < 1e-6


def test_synth_spline_speed_limits():
    cfg = replace(C2, frames=300)
    poses = [synth_pose(cfg, f).astype(np.float64) for f in range(0, 300)]
    v = [np.linalg.norm(b[:3, 3] - a[:3, 3]) * 30 for a, b in zip(poses, poses[1:])]
    w = []
    for a, b in zip(poses, poses[1:]):
        r = a[:3, :3].T @ b[:3, :3]
        w.append(np.degrees(np.arccos(np.clip((np.trace(r) - 1) / 2, -1, 1))) * 30)
    assert max(v) <= 0.5 + 1e-3 and max(w) <= 30.0 + 0.5
    holes = synth_depth(cfg, 10) == 0
    assert 0.01 < holes.mean() < 0.05


def test_header_constants_match_binding():
    text = (ROOT / "include" / "meshhashing.h").read_text()
    for name in ("MH_NUM_TILINGS", "MH_NUM_TUNE_GRIDS", "MH_NUM_RAY_SHAPES", "MH_NUM_MC_SHAPES"):
        m = re.search(rf"#define {name} (\d+)", text)
        assert m and int(m.group(1)) == getattr(A, name), name
    assert C.sizeof(A.TuneResult) == 40
    assert C.sizeof(A.TuneShapeResult) == 24
    assert C.sizeof(A.TuneMcResult) == 32
