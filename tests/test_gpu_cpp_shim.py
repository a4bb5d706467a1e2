"""A compiled C++ caller of the C ABI: tests/cpp/main_engine_shim.cpp is
MeshHashing's MainEngine (Mapping / Meshing / Recycle / RenderRayCaster /
FinalLog, engine/main_engine.cc:100-192,231-239,299-311) written against
include/meshhashing.h only and built with g++.  On the 320x240 room
sequence it must match the oracle's own fusion of the first five frames,
and after fifteen frames (recycling included) end bit-identical to the same
schedule driven from Python (MainEngine in meshhashing_amd/engine.py, which
the GC tests hold against the oracle): fused blocks with SDF / weight /
colour bits, status counters, the raycast of the last frame and the final
PLY file.

The binary is built with the library (`make -C tests/cpp`); the test skips
if it is missing.
"""
import subprocess
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from meshhashing_amd.configs import C1
from meshhashing_amd.engine import MainEngine, synth_depth, synth_pose

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
SHIM = ROOT / "tests" / "cpp" / "_build" / "main_engine_shim"


def run_shim(cfg, frames, out):
    blob = b"".join(bytes(x) for x in (cfg.hash_params(), cfg.volume_params(), cfg.mesh_params(),
                                       cfg.sensor_params(), cfg.raycaster_params(), cfg.mc_params(),
                                       cfg.synth_spec()))
    (out / "config.bin").write_bytes(blob + np.array([frames, 1], np.uint32).tobytes())
    print(f"[shim] compiled MainEngine, {frames} frames", flush=True)
    r = subprocess.run([str(SHIM), str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0, r.stderr
    return {"pos": np.fromfile(out / "pos.bin", np.int32).reshape(-1, 3),
            "sdf": np.fromfile(out / "sdf.bin", np.float32).reshape(-1, 512),
            "weight": np.fromfile(out / "weight.bin", np.float32).reshape(-1, 512),
            "color": np.fromfile(out / "color.bin", np.uint32).reshape(-1, 512)}


def test_compiled_main_engine_matches_the_oracle(hip_device, oracle_mod, tmp_path):
    """Five frames (before any recycling point): the block set bit-exact and
    SDF / weight within 1e-5 of the oracle's own fusion, colour exact."""
    if not SHIM.exists():
        pytest.skip("tests/cpp/_build/main_engine_shim not built")
    cfg = replace(C1, frames=20)
    g = run_shim(cfg, 5, tmp_path)
    orc = oracle_mod.Oracle(cfg)
    for f in range(5):
        orc.frame(synth_depth(cfg, f), synth_pose(cfg, f), mesh=False)
    o = orc.dump_blocks()
    assert np.array_equal(g["pos"], o["pos"])
    obs = o["weight"] > 0
    assert np.array_equal(g["weight"] > 0, obs)
    es = np.abs(g["sdf"][obs] - o["sdf"][obs]) / np.maximum(np.abs(o["sdf"][obs]), cfg.truncation)
    ew = np.abs(g["weight"][obs] - o["weight"][obs]) / np.maximum(o["weight"][obs], 1.0)
    print(f"[shim] {len(g['pos'])} blocks, sdf {es.max():.2e} weight {ew.max():.2e}")
    assert es.max() < 1e-5 and ew.max() < 1e-5
    assert np.array_equal(g["color"], o["color"])


def test_compiled_main_engine_matches_python_driver(hip_device, tmp_path):
    """Fifteen frames with the recycling schedule (starve + garbage at frame
    count 14, low-surfel recycling at 10): bit-identical to the Python
    driver, whose schedule the GC tests hold against the oracle."""
    import torch
    if not SHIM.exists():
        pytest.skip("tests/cpp/_build/main_engine_shim not built")
    cfg = replace(C1, frames=20)
    frames = list(range(15))
    g = run_shim(cfg, len(frames), tmp_path)
    eng = MainEngine(cfg)
    for f in frames:
        eng.frame(synth_depth(cfg, f), synth_pose(cfg, f))
    vol = eng.vol
    n = cfg.width * cfg.height
    d = torch.empty(n, dtype=torch.float32, device=hip_device)
    nn = torch.empty(3 * n, dtype=torch.float32, device=hip_device)
    cTw = vol.pose()[1]
    vol.raycast(cTw, d.data_ptr(), 0, nn.data_ptr())
    vol.synchronize()
    assert np.array_equal(np.fromfile(tmp_path / "cTw.bin", np.float32).reshape(4, 4), cTw)
    assert np.array_equal(np.fromfile(tmp_path / "ray_depth.bin", np.float32).view(np.uint32),
                          d.cpu().numpy().view(np.uint32))
    assert np.array_equal(np.fromfile(tmp_path / "ray_normal.bin", np.float32).view(np.uint32),
                          nn.cpu().numpy().view(np.uint32))
    p = vol.dump_blocks()
    assert np.array_equal(g["pos"], p["pos"])
    for k in ("sdf", "weight", "color"):
        assert np.array_equal(g[k].view(np.uint32), p[k].view(np.uint32)), k
    from meshhashing_amd import _capi as A
    st = A.Status.from_buffer_copy((tmp_path / "status.bin").read_bytes())
    mine = vol.status()
    for k, _ in A.Status._fields_:
        if k != "mc_vertex_writes":  # the claimer of a fresh slot is a race
            assert getattr(st, k) == mine[k], k
    assert st.flags == 0
    # FinalLog: collect all, mesh, compress, PLY
    vol.collect_all()
    vol.marching_cubes()
    vol.compress_mesh()
    vol.save_ply(str(tmp_path / "python.ply"))
    a, b = ply_triangles(tmp_path / "mesh.ply"), ply_triangles(tmp_path / "python.ply")
    assert len(a) > 10_000 and a == b


def ply_triangles(path):
    """The PLY's triangles as a sorted list of vertex-record triples (vertex
    numbering follows the vertex pool, whose order is a race)."""
    lines = path.read_text().splitlines()
    nv = int(next(l for l in lines if l.startswith("element vertex")).split()[-1])
    nf = int(next(l for l in lines if l.startswith("element face")).split()[-1])
    start = lines.index("end_header") + 1
    verts = lines[start:start + nv]
    faces = lines[start + nv:start + nv + nf]
    tris = []
    for f in faces:
        k = [int(x) for x in f.split()]
        assert k[0] == 3
        tris.append(tuple(sorted(verts[i] for i in k[1:])))
    return sorted(tris)
