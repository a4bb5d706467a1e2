"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on
the same seeded inputs.

Bars (tolerances are written in each test):
  * block set, hash layout (canonical), worklist, MC cube indices, triangle
    topology and vertex positions: bit-exact given identical SDF;
  * SDF, weight, PSDF (mu, 1/sigma^2, a, b): 1e-5 relative, relative to
    max(|oracle|, scale) with scale = truncation for SDF, 1 otherwise;
  * raycast depth: within 1e-4 m given identical SDF (observed: bit-exact).
"""
from dataclasses import replace

import numpy as np
import pytest

from meshhashing_amd.configs import C1, C3
from meshhashing_amd.engine import Volume, synth_color, synth_depth, synth_pose

pytestmark = pytest.mark.gpu


def canon_blocks(d):
    order = np.lexsort((d["pos"][:, 2], d["pos"][:, 1], d["pos"][:, 0]))
    return {k: v[order] for k, v in d.items()}


def rel_err(gpu, ref, scale):
    return np.abs(gpu.astype(np.float64) - ref) / np.maximum(np.abs(ref), scale)


def run_frames(cfg, frames, vol, orc, mesh=False, color=False):
    for f in frames:
        depth = synth_depth(cfg, f)
        pose = synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        if color:
            bgra = synth_color(cfg, f)
            vol.upload_color(bgra)
            orc.set_frame(depth, pose)
            orc.set_color(bgra)
            orc.alloc()
            orc.collect_visible()
            orc.integrate_tsdf()
        else:
            orc.frame(depth, pose, mesh=False)
        vol.alloc()
        vol.collect_visible()
        if cfg.psdf:
            vol.predict_inlier()
            vol.integrate_psdf()
        else:
            vol.integrate_tsdf()
        if mesh:
            vol.marching_cubes()
            orc.marching_cubes()


def test_alloc_blockset_and_hash_layout(hip_device, oracle_mod):
    cfg = C1
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    for f in range(0, cfg.frames, 4):
        print(f"[alloc] frame {f}", flush=True)
        depth, pose = synth_depth(cfg, f), synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        vol.alloc()
        n = vol.collect_visible(sync=True)
        orc.set_frame(depth, pose)
        orc.alloc()
        m = orc.collect_visible()
        g = vol.dump_blocks()["pos"]
        o = orc.dump_blocks()["pos"]
        assert np.array_equal(g, o), f"frame {f}: block sets differ"
        assert n == m
        assert np.array_equal(vol.dump_candidates(), orc.dump_candidates())
    hg, ho = vol.dump_hash(), orc.dump_hash()
    # canonical layout: per home bucket, the set of keys and the in-bucket count
    def canon(rec):
        out = {}
        for r in rec:
            out.setdefault(int(r[1]), []).append((tuple(r[3:]), int(r[2])))
        return {k: (sorted(p for p, _ in v), sum(f for _, f in v)) for k, v in out.items()}
    assert canon(hg) == canon(ho)
    st = vol.status()
    assert st["flags"] == 0, st
    assert st["block_count"] == len(o)


def test_alloc_with_linked_lists(hip_device, oracle_mod):
    """A 200-bucket table (~6 keys per bucket) forces the overflow lists;
    a 40-slot list window keeps every key placeable, so the result is still
    lossless and must equal the oracle's."""
    cfg = replace(C1, bucket_count=200, value_capacity=1 << 14, linked_list_size=40)
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    depth, pose = synth_depth(cfg, 0), synth_pose(cfg, 0)
    vol.upload_depth(depth, pose)
    vol.alloc()
    orc.set_frame(depth, pose)
    orc.alloc()
    st = vol.status()
    g = vol.dump_blocks()["pos"]
    o = orc.dump_blocks()["pos"]
    assert orc.dropped() == 0 and st["dropped_inserts"] == 0 and st["flags"] == 0, st
    assert np.array_equal(g, o)
    def canon(rec):
        out = {}
        for r in rec:
            out.setdefault(int(r[1]), []).append((tuple(r[3:]), int(r[2])))
        return {k: (sorted(p for p, _ in v), sum(f for _, f in v)) for k, v in out.items()}
    assert canon(vol.dump_hash()) == canon(orc.dump_hash())
    rec = vol.dump_hash()
    assert len({tuple(r[3:]) for r in rec}) == len(rec)  # no duplicates
    assert (rec[:, 2] == 0).any(), "no list entries were exercised"
    for b in np.unique(rec[rec[:, 2] == 0][:, 1]):
        assert (rec[:, 0] // 10 == b).sum() == 10  # a list entry only behind a full bucket
    assert st["list_inserts"] > 0


def test_parallel_list_insert_over_frames(hip_device, oracle_mod):
    """The overflow lists under load: a 500-bucket table (3 578 blocks in 5 000
    slots) puts ~200 keys in lists over five frames, each key queued by many
    rays.  Every
    frame the block set equals the oracle's (sequential inserts), no key is
    stored twice, every stored key is found again by a lookup (each list is
    intact) and fusion through the lists matches the oracle."""
    cfg = replace(C1, bucket_count=500, value_capacity=1 << 14, linked_list_size=400)
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    total = 0
    for f in range(0, 10, 2):
        depth, pose = synth_depth(cfg, f), synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
        orc.frame(depth, pose, mesh=False)
        st = vol.status()
        assert st["dropped_inserts"] == 0 and st["flags"] == 0, st
        rec = vol.dump_hash()
        assert len({tuple(r[3:]) for r in rec}) == len(rec), f"frame {f}: duplicate keys"
        g, o = vol.dump_blocks(), orc.dump_blocks()
        assert np.array_equal(g["pos"], o["pos"]), f"frame {f}: block sets differ"
        total = st["list_inserts"]
    g = vol.dump_blocks()
    o = orc.dump_blocks()
    obs = o["weight"] > 0
    assert np.array_equal(g["weight"] > 0, obs)
    es = np.abs(g["sdf"][obs] - o["sdf"][obs]) / np.maximum(np.abs(o["sdf"][obs]), cfg.truncation)
    assert es.max() < 1e-5
    # every stored key is reachable through its bucket and list: the raycast
    # (hash lookups of every block it meets) equals the oracle's over the
    # same blocks
    import torch
    orc2 = oracle_mod.Oracle(cfg)
    orc2.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
    cTw = np.linalg.inv(synth_pose(cfg, 5).astype(np.float64)).astype(np.float32)
    d = torch.empty(cfg.width * cfg.height, dtype=torch.float32, device=hip_device)
    vol.raycast(cTw, d.data_ptr(), 0, 0)
    vol.synchronize()
    dg = d.cpu().numpy().reshape(cfg.height, cfg.width)
    do = orc2.raycast(cTw)[0]
    assert np.array_equal(np.isfinite(dg), np.isfinite(do)) and np.isfinite(dg).mean() > 0.5
    assert np.abs(dg[np.isfinite(dg)] - do[np.isfinite(do)]).max() <= 1e-4
    print(f"[lists] {total} list inserts, {len(g['pos'])} blocks")
    assert total > 150 and int((vol.dump_hash()[:, 2] == 0).sum()) == total
    vol.close()


@pytest.mark.parametrize("color", [False, True])
def test_tsdf_fusion_end_to_end(hip_device, oracle_mod, color):
    cfg = C1
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    run_frames(cfg, range(0, cfg.frames, 2), vol, orc, color=color)
    g = vol.dump_blocks()
    o = orc.dump_blocks()
    assert np.array_equal(g["pos"], o["pos"])
    obs = o["weight"] > 0
    assert np.array_equal(g["weight"] > 0, obs)
    e_sdf = rel_err(g["sdf"][obs], o["sdf"][obs], cfg.truncation)
    e_w = rel_err(g["weight"][obs], o["weight"][obs], 1.0)
    print(f"[tsdf] voxels {obs.sum()} sdf max rel {e_sdf.max():.3e} weight max rel {e_w.max():.3e}")
    assert e_sdf.max() < 1e-5
    assert e_w.max() < 1e-5
    assert np.array_equal(g["color"], o["color"])
    if color:  # the colour image was used: more than (0, 255, 0) blends
        assert len(np.unique(g["color"][obs])) > 1000


def test_frame_conversion_without_allocation(hip_device, oracle_mod):
    """The frame's depth / colour conversion normally runs inside the
    allocation kernel; a frame fused without an allocation pass (worklist
    of the blocks already allocated) is converted by the standalone path and
    must give the same result as the oracle."""
    cfg = C1
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    run_frames(cfg, [0], vol, orc, color=True)
    for f in (2, 4):
        depth, pose, bgra = synth_depth(cfg, f), synth_pose(cfg, f), synth_color(cfg, f)
        vol.upload_depth(depth, pose)
        vol.upload_color(bgra)
        vol.collect_visible()
        vol.integrate_tsdf()
        orc.set_frame(depth, pose)
        orc.set_color(bgra)
        orc.collect_visible()
        orc.integrate_tsdf()
    g, o = vol.dump_blocks(), orc.dump_blocks()
    assert np.array_equal(g["pos"], o["pos"])
    obs = o["weight"] > 0
    assert np.array_equal(g["weight"] > 0, obs)
    assert rel_err(g["sdf"][obs], o["sdf"][obs], cfg.truncation).max() < 1e-5
    assert np.array_equal(g["color"], o["color"])
    depth_dev = vol.depth_ptr()  # converted depth of the last frame
    assert depth_dev


def test_mc_lockstep_given_identical_sdf(hip_device, oracle_mod):
    """Per frame, the oracle takes the GPU's fused SDF, then both mesh the
    same worklist incrementally; topology and vertex positions must agree
    bit for bit, normals to 1e-4."""
    cfg = C1
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    for f in range(0, 12, 2):
        print(f"[mc] frame {f}", flush=True)
        depth, pose = synth_depth(cfg, f), synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
        orc.set_frame(depth, pose)
        orc.alloc()
        orc.collect_visible()
        g = vol.dump_blocks()
        orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
        vol.marching_cubes()
        orc.marching_cubes()
        cg = vol.dump_cube_indices()
        co = orc.dump_blocks()["cube"]
        assert np.array_equal(cg, co), f"frame {f}: cube indices differ"
        mg, mo = vol.dump_mesh(), orc.dump_mesh()
        assert len(mg["tri"]) > 100
        assert np.array_equal(mg["tri_cube"], mo["tri_cube"])
        assert np.array_equal(mg["tri"], mo["tri"])
        assert np.array_equal(mg["vkey"], mo["vkey"])
        assert np.array_equal(mg["vref"], mo["vref"])
        assert np.array_equal(mg["vval"][:, :3], mo["vval"][:, :3].astype(np.float32))
        # Normals to 1e-4, except where the SDF gradient itself is at the FP32
        # rounding floor (|grad| < 1e-3 against ~1 for a true SDF): there the
        # direction is undefined in FP32 and FP64 alike.
        dnv = np.abs(mg["vval"][:, 3:6] - mo["vval"][:, 3:6]).max(axis=1)
        bad = np.nonzero(dnv >= 1e-4)[0]
        for i in bad:
            gr, ok = orc.spatial_gradient(mg["vval"][i, :3])
            print("  ill-conditioned normal", mg["vkey"][i], "|grad|", np.linalg.norm(gr), ok)
            assert np.linalg.norm(gr) < 1e-3, (mg["vkey"][i], gr, mg["vval"][i], mo["vval"][i])
        assert len(bad) <= 1e-4 * len(dnv) + 2
        dn = np.delete(dnv, bad).max()
        print(f"[mc] frame {f}: {len(mg['tri'])} triangles, {len(mg['vkey'])} vertices, "
              f"normal err {dn:.2e}, ill-conditioned {len(bad)}")
        assert dn < 1e-4
    st = vol.status()
    assert st["flags"] == 0, st


def test_raycast_given_identical_sdf(hip_device, oracle_mod):
    import torch
    cfg = C1
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    run_frames(cfg, range(0, 8, 2), vol, orc)
    g = vol.dump_blocks()
    orc2 = oracle_mod.Oracle(cfg)
    orc2.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
    pose = synth_pose(cfg, 3)
    cTw = np.linalg.inv(pose.astype(np.float64)).astype(np.float32)
    n = cfg.width * cfg.height
    d = torch.empty(n, dtype=torch.float32, device=hip_device)
    v = torch.empty(3 * n, dtype=torch.float32, device=hip_device)
    nn = torch.empty(3 * n, dtype=torch.float32, device=hip_device)
    vol.raycast(cTw, d.data_ptr(), v.data_ptr(), nn.data_ptr())
    vol.synchronize()
    dg = d.cpu().numpy().reshape(cfg.height, cfg.width)
    ng = nn.cpu().numpy().reshape(cfg.height, cfg.width, 3)
    do, vo, no = orc2.raycast(cTw)
    hit_g, hit_o = np.isfinite(dg), np.isfinite(do)
    assert np.array_equal(hit_g, hit_o)
    assert hit_g.mean() > 0.5
    err = np.abs(dg[hit_g] - do[hit_g])
    print(f"[raycast] hits {hit_g.sum()} max |d| err {err.max():.3e}, bit-exact {np.mean(err == 0):.4f}")
    assert err.max() <= 1e-4
    assert np.abs(ng[hit_g] - no[hit_g]).max() < 1e-4


def test_psdf_and_primal_dual(hip_device, oracle_mod):
    cfg = replace(C3, frames=20)
    vol = Volume(cfg)
    orc = oracle_mod.Oracle(cfg)
    errs = {}
    for f in range(0, 8, 2):
        print(f"[psdf] frame {f}", flush=True)
        depth, pose = synth_depth(cfg, f), synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        vol.alloc()
        vol.collect_visible()
        orc.set_frame(depth, pose)
        orc.alloc()
        orc.collect_visible()
        # identical state in: the GPU's fields and (from the second frame) the
        # GPU's vertices, whose normals can differ from the oracle's where
        # the SDF gradient is at the FP32 floor (checked below)
        g = vol.dump_blocks()
        orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"], g["a"], g["b"])
        if f > 0:
            orc.load_vertices(mg["vkey"], mg["vval"])
        vol.predict_inlier()
        orc.predict_inlier()
        vol.synchronize()
        gi = d2h_f32(vol.inlier_ptr(), cfg.width * cfg.height).reshape(cfg.height, cfg.width)
        oi = orc.inlier()
        errs["inlier"] = max(errs.get("inlier", 0), float(np.abs(gi - oi).max()))
        vol.integrate_psdf()
        orc.integrate_psdf()
        g = vol.dump_blocks()
        o = orc.dump_blocks()
        obs = o["weight"] > 0
        assert np.array_equal(g["weight"] > 0, obs)
        for k, s in (("sdf", cfg.truncation), ("weight", 1.0), ("a", 1.0), ("b", 1.0)):
            e = rel_err(g[k][obs], o[k][obs], s)
            errs[k] = max(errs.get(k, 0), float(e.max()))
        # meshing with the reference's default rho_min (0.1) and a low
        # precision threshold, so the frames carry a mesh (with C3's 0.4 and
        # (1/vs)^2 almost no cube passes in these frames), whose vertices
        # the next frame's inlier prediction then uses
        mcp = replace(cfg, mc_weight_min=100.0, mc_rho_min=0.1).mc_params()
        vol.marching_cubes(mcp)
        orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"], g["a"], g["b"])
        orc.marching_cubes(mcp)
        mg, mo = vol.dump_mesh(), orc.dump_mesh()
        assert np.array_equal(mg["tri"], mo["tri"])
        assert np.array_equal(mg["vkey"], mo["vkey"])
        assert np.array_equal(mg["vval"][:, :3], mo["vval"][:, :3].astype(np.float32))
        dn = np.abs(mg["vval"][:, 3:6] - mo["vval"][:, 3:6]).max(axis=1)
        bad = np.nonzero(dn >= 1e-4)[0]
        for i in bad:
            # The normal is the SDF gradient normalised: an FP32 gradient
            # carries an absolute error of ~1e-6 (sample rounding over the
            # stencil's half-voxel spacing), so where |grad| is small the
            # direction moves by ~1e-6 / |grad|.  PSDF fields have many such
            # points (unobserved voxels keep sdf = 0 inside allocated blocks).
            gr, _ = orc.spatial_gradient(mg["vval"][i, :3])
            assert dn[i] * np.linalg.norm(gr) < 2e-6, (mg["vkey"][i], gr, mg["vval"][i], mo["vval"][i])
        errs["normal"] = max(errs.get("normal", 0), float(np.delete(dn, bad).max()) if len(dn) > len(bad) else 0.0)
        errs["ill_conditioned"] = max(errs.get("ill_conditioned", 0), len(bad) / max(1, len(dn)))
        # vertex payload (marching_cubes.cu:59-69): radius = sqrt(1 / inv_sigma2)
        # of the interpolated voxel, colour = ValToRGB(a / (a + b), 0.4, 1)
        rg, ro = mg["vval"][:, 9].astype(np.float64), mo["vval"][:, 9]
        fin = np.isfinite(ro)
        assert np.array_equal(np.isfinite(rg), fin)
        # a/(a+b) is NaN where an interpolation corner was never observed
        # (a = b = 0): ValToRGB then gives (1, NaN, 0) in both
        cg, co = mg["vval"][:, 6:9].astype(np.float64), mo["vval"][:, 6:9]
        cfin = np.isfinite(co)
        assert np.array_equal(np.isfinite(cg), cfin)
        errs["vertices"] = errs.get("vertices", 0) + len(ro)
        if fin.any():
            errs["radius"] = max(errs.get("radius", 0), float(rel_err(rg[fin], ro[fin], 1e-3).max()))
            errs["colour"] = max(errs.get("colour", 0), float(np.abs(cg[cfin] - co[cfin]).max()))
        gi2 = d2h_f32(vol.inlier_ptr(), cfg.width * cfg.height)
        errs["inlier_above_reset"] = max(errs.get("inlier_above_reset", 0),
                                         int((gi2 > np.float32(0.1)).sum()))
    print("[psdf] errors (max over frames)", errs)
    assert errs["inlier"] < 1e-5
    for k in ("sdf", "weight", "a", "b", "radius"):
        assert errs[k] < 1e-5, (k, errs)
    assert errs["colour"] < 1e-5, errs
    assert errs["vertices"] > 10000, errs
    assert errs["inlier_above_reset"] > 100, errs  # the prediction saw vertices
    assert errs["normal"] < 1e-4 and errs["ill_conditioned"] < 0.05, errs
    # primal-dual over all blocks from identical state
    g = vol.dump_blocks()
    orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"], g["a"], g["b"])
    vol.collect_all()
    orc.collect_all()
    vol.pd_init()
    orc.pd_init()
    for it in range(cfg.pd_iterations):
        eg = vol.pd_iterate(cfg.pd_lambda, cfg.pd_sigma, cfg.pd_tau, energy=(it % 10 == 0))
        eo = orc.pd_iterate(cfg.pd_lambda, cfg.pd_sigma, cfg.pd_tau)
        if eg is not None:
            assert abs(eg[2] - eo[2]) <= 1e-4 * max(abs(eo[2]), 1e-6), (it, eg, eo)
    g = vol.dump_blocks()
    o = orc.dump_blocks()
    obs = o["weight"] > 1e-6
    e = rel_err(g["sdf"][obs], o["sdf"][obs], cfg.truncation)
    print(f"[pd] sdf max rel {e.max():.3e}")
    assert e.max() < 1e-5


def d2h_f32(dev_ptr, n):
    """Copy n floats from device memory with hipMemcpy (no torch involved)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.zeros(n, np.float32)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(dev_ptr),
                       ctypes.c_size_t(4 * n), ctypes.c_int(2))
    assert rc == 0, rc
    return out.astype(np.float64)
