"""Mesh compaction (mh_compress_mesh) and PLY output (mh_save_ply).

The whole-volume mesh of MainEngine::FinalLog (collect all blocks, mesh,
compress) against the oracle's marching cubes on the same SDF: the
compacted triangles, as ordered triples of vertex positions, equal the
oracle's triangles bit for bit (as a multiset); every vertex is used; the
PLY file holds the same mesh (positions to the stream's 6 digits).
"""
import numpy as np
import pytest

from meshhashing_amd.configs import C1
from meshhashing_amd.engine import Volume, synth_depth, synth_pose

pytestmark = pytest.mark.gpu


def tri_positions(tri, pos):
    return np.concatenate([pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]], axis=1)


def sort_rows(a):
    return a[np.lexsort(a.T[::-1])]


def test_compressed_mesh_matches_oracle(hip_device, oracle_mod, tmp_path):
    cfg = C1
    vol = Volume(cfg)
    for f in range(0, 8, 2):
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
        vol.marching_cubes()
    vol.collect_all()
    vol.marching_cubes()
    cm = vol.compress_mesh()
    g = vol.dump_blocks()
    orc = oracle_mod.Oracle(cfg)
    orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
    orc.collect_all()
    orc.marching_cubes()
    om = orc.dump_mesh()
    keypos = {tuple(k): v[:3] for k, v in zip(om["vkey"], om["vval"])}
    otri = np.array([[c for k in range(3) for c in keypos[tuple(t[4 * k:4 * k + 4])]]
                     for t in om["tri"]], np.float64)
    gtri = tri_positions(cm["tri"], cm["pos"]).astype(np.float64)
    print(f"[compact] {len(cm['pos'])} vertices, {len(cm['tri'])} triangles "
          f"(oracle {len(om['vkey'])} / {len(om['tri'])})")
    assert len(cm["tri"]) == len(om["tri"]) > 1000
    assert len(cm["pos"]) == len(om["vkey"])
    assert np.array_equal(sort_rows(gtri), sort_rows(otri))
    assert np.array_equal(np.unique(cm["tri"]), np.arange(len(cm["pos"])))  # every vertex used
    # PLY
    path = tmp_path / "mesh.ply"
    vol.save_ply(str(path))
    lines = path.read_text().splitlines()
    end = lines.index("end_header")
    assert lines[0] == "ply" and lines[1] == "format ascii 1.0"
    assert f"element vertex {len(cm['pos'])}" in lines[:end]
    assert f"element face {len(cm['tri'])}" in lines[:end]
    vtx = np.array([l.split() for l in lines[end + 1:end + 1 + len(cm["pos"])]], np.float64)
    assert np.allclose(vtx[:, :3], cm["pos"], rtol=1e-5, atol=1e-6)
    assert np.allclose(vtx[:, 3:6], cm["normal"], rtol=1e-5, atol=1e-6)
    # int(255 c) with NaN -> 0 (plain-TSDF vertices colour ValToRGB(0/0))
    want = np.nan_to_num(np.float32(255.0) * cm["color"], nan=0.0).astype(np.int64)
    assert np.array_equal(vtx[:, 6:9].astype(np.int64), want)
    faces = np.array([l.split() for l in lines[end + 1 + len(cm["pos"]):]], np.int64)
    assert np.array_equal(faces[:, 0], np.full(len(faces), 3))
    assert np.array_equal(faces[:, 1:], cm["tri"])
