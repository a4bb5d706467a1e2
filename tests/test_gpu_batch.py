"""Batched scenes (mh_batch_*) and integration tilings (mh_tune_integrate,
mh_set_integrate_tiling) through the C ABI.

Bar: bit-identical per-scene results.  A batch runs every stage as one
launch over all scenes, eagerly or as a replayed HIP graph; each scene must
end exactly where running it alone through the single-volume entry points
ends (canonical block dump with SDF / weight / colour bits, canonical hash
layout, worklist, cube indices, mesh, status counters, raycast images of
every frame).  Integration tilings and the autotuner must not change a bit.
"""
from dataclasses import replace

import numpy as np
import pytest

from meshhashing_amd import _capi as A
from meshhashing_amd.configs import C1, c4_scenes
from meshhashing_amd.engine import Batch, Volume, mc_tune_calls, synth_depth, synth_pose

pytestmark = pytest.mark.gpu

# counters that depend only on the data (vertex writes also count the
# claimer of a fresh slot, whose identity is a race)
STATUS_KEYS = ("flags", "dropped_inserts", "dda_cap_hits", "list_inserts", "block_count",
               "candidate_count", "vertex_count", "triangle_count", "voxel_updates",
               "blocks_integrated", "mc_triangle_writes", "mc_blocks", "ray_samples")


def canon_hash(rec):
    out = {}
    for r in rec:
        out.setdefault(int(r[1]), []).append((tuple(r[3:]), int(r[2])))
    return {k: (sorted(p for p, _ in v), sum(f for _, f in v)) for k, v in out.items()}


def scene_state(vol, mesh=True):
    b = vol.dump_blocks()
    st = vol.status()
    out = {"pos": b["pos"], "sdf": b["sdf"].view(np.uint32), "weight": b["weight"].view(np.uint32),
           "color": b["color"], "cand": vol.dump_candidates(), "hash": canon_hash(vol.dump_hash()),
           "status": {k: st[k] for k in STATUS_KEYS}}
    if mesh:
        m = vol.dump_mesh()
        out["cube"] = vol.dump_cube_indices()
        out["tri"] = m["tri"]
        out["vkey"] = m["vkey"]
        out["vval"] = m["vval"].view(np.uint32)
        out["vref"] = m["vref"]
    return out


def assert_same(a, b, what):
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k]), f"{what}: {k} differs"
        else:
            assert a[k] == b[k], f"{what}: {k} differs: {a[k]} vs {b[k]}"


def sequential(cfg, frames, hip_device, raycast=True):
    """One scene alone through the single-volume entry points."""
    import torch
    vol = Volume(cfg)
    n = cfg.width * cfg.height
    images = []
    for f in frames:
        depth, pose = synth_depth(cfg, f), synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
        vol.marching_cubes()
        if raycast:
            d = torch.empty(n, dtype=torch.float32, device=hip_device)
            nn = torch.empty(3 * n, dtype=torch.float32, device=hip_device)
            vol.raycast(vol.pose()[1], d.data_ptr(), 0, nn.data_ptr())
            vol.synchronize()
            images.append((d.cpu().numpy().view(np.uint32), nn.cpu().numpy().view(np.uint32)))
    return vol, images


def batched(cfgs, frames, hip_device, graph=False, tune_first=False, mc_tune=False):
    import torch
    b = Batch(cfgs)
    if mc_tune:
        b.tune_marching_cubes(1)
    W, H, S = cfgs[0].width, cfgs[0].height, len(cfgs)
    d = torch.empty(S * W * H, dtype=torch.float32, device=hip_device)
    nn = torch.empty(S * W * H * 3, dtype=torch.float32, device=hip_device)
    b.set_raycast(d.data_ptr(), 0, nn.data_ptr())
    images = []
    tune = None
    for i, f in enumerate(frames):
        depth = np.stack([synth_depth(c, f) for c in cfgs])
        pose = np.stack([synth_pose(c, f) for c in cfgs])
        b.frame(depth, pose)
        d.fill_(-1.0)  # eager torch work between graph replays
        r = b.run(mesh=True, raycast=True, graph=graph and not (tune_first and i == 0),
                  tune=tune_first and i == 0)
        tune = tune or r
        b.synchronize()
        dn, nn_ = d.cpu().numpy().view(np.uint32), nn.cpu().numpy().view(np.uint32)
        images.append([(dn[s * W * H:(s + 1) * W * H], nn_[3 * s * W * H:3 * (s + 1) * W * H])
                       for s in range(S)])
    return b, images, tune


def test_batch_matches_sequential(hip_device):
    cfgs = c4_scenes(n=8, frames=3)
    sizes = {c.voxel_size for c in cfgs}
    assert len(sizes) >= 3, sizes
    frames = range(3)
    b, bimg, _ = batched(cfgs, frames, hip_device)
    for s, cfg in enumerate(cfgs):
        print(f"[batch] scene {s} voxel {cfg.voxel_size}", flush=True)
        vol, simg = sequential(cfg, frames, hip_device)
        assert_same(scene_state(b.volumes[s]), scene_state(vol), f"scene {s}")
        for f in frames:
            assert np.array_equal(bimg[f][s][0], simg[f][0]), f"scene {s} frame {f}: depth"
            assert np.array_equal(bimg[f][s][1], simg[f][1]), f"scene {s} frame {f}: normal"
        st = b.volumes[s].status()
        assert st["flags"] == 0 and st["block_count"] > 50, st
        vol.close()
    b.close()


def test_batch_graph_replays_match_eager(hip_device):
    """The captured frame graph, replayed on every frame (new depth, poses and
    MC stamps through the device table) with eager work in between, equals
    the eager batch after every frame."""
    cfgs = c4_scenes(n=6, frames=5, seed=9)
    frames = range(5)
    be, eimg, _ = batched(cfgs, frames, hip_device, graph=False)
    bg, gimg, _ = batched(cfgs, frames, hip_device, graph=True)
    for f in frames:
        for s in range(len(cfgs)):
            assert np.array_equal(eimg[f][s][0], gimg[f][s][0]), f"frame {f} scene {s}: depth"
            assert np.array_equal(eimg[f][s][1], gimg[f][s][1]), f"frame {f} scene {s}: normal"
    for s in range(len(cfgs)):
        assert_same(scene_state(bg.volumes[s]), scene_state(be.volumes[s]), f"scene {s}")
    be.close()
    bg.close()


def test_batch_tuning_changes_nothing(hip_device):
    cfgs = c4_scenes(n=6, frames=3, seed=5)
    frames = range(3)
    b0, img0, _ = batched(cfgs, frames, hip_device)
    b1, img1, tune = batched(cfgs, frames, hip_device, graph=True, tune_first=True)
    print(f"[tune] batch: tiling {tune.tiling} ({tune.workgroup_size} lanes, "
          f"{tune.blocks_per_workgroup} blocks/wg, {tune.workgroups_per_cu} wg/CU) "
          f"{tune.best_ms:.4f} ms vs default {tune.default_ms:.4f} ms")
    assert tune.candidates == A.MH_NUM_TILINGS * A.MH_NUM_TUNE_GRIDS and 0 <= tune.tiling < A.MH_NUM_TILINGS
    assert tune.best_ms <= tune.default_ms
    assert 0 <= tune.ray_shape < A.MH_NUM_RAY_SHAPES  # the raycast tile shape, tuned too
    for f in frames:
        for s in range(len(cfgs)):
            assert np.array_equal(img0[f][s][0], img1[f][s][0])
    for s in range(len(cfgs)):
        assert_same(scene_state(b1.volumes[s]), scene_state(b0.volumes[s]), f"scene {s}")
    b0.close()
    b1.close()


def run_tsdf(cfg, frames, tiling=None, tune_at=None):
    vol = Volume(cfg)
    if tiling is not None:
        vol.set_integrate_tiling(*tiling)
    res = None
    for f in frames:
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        if f == tune_at:
            res = vol.tune_integrate(reps=2)
        vol.integrate_tsdf()
    return vol, res


def test_integrate_tilings_are_output_invariant(hip_device):
    cfg = replace(C1, frames=6)
    frames = range(0, 6, 2)
    ref, _ = run_tsdf(cfg, frames)
    want = scene_state(ref, mesh=False)
    for t in range(A.MH_NUM_TILINGS):
        for g in (2, 32):
            print(f"[tiling] {t} x {g}", flush=True)
            vol, _ = run_tsdf(cfg, frames, tiling=(t, g))
            assert_same(scene_state(vol, mesh=False), want, f"tiling {t} grid {g}")
            vol.close()


def test_tune_integrate_changes_nothing(hip_device):
    cfg = replace(C1, frames=6)
    frames = range(0, 6, 2)
    ref, _ = run_tsdf(cfg, frames)
    vol, res = run_tsdf(cfg, frames, tune_at=2)
    print(f"[tune] volume: tiling {res.tiling} {res.best_ms:.4f} ms vs default {res.default_ms:.4f} ms")
    assert res.candidates == A.MH_NUM_TILINGS * A.MH_NUM_TUNE_GRIDS
    assert_same(scene_state(vol, mesh=False), scene_state(ref, mesh=False), "tuned")


def test_c4_all_scenes_batched_equal_sequential(hip_device):
    """The whole batched workload: 64 scenes x 4 frames of 320x240, voxel
    sizes 1-4 cm, replayed as one HIP graph per frame (integration tuned on
    the first frame), against each scene run alone: every scene's state and
    every frame's raycast images bit-identical."""
    cfgs = c4_scenes(64, 4)
    frames = range(4)
    b, bimg, tune = batched(cfgs, frames, hip_device, graph=True, tune_first=True)
    counts = []
    for s, cfg in enumerate(cfgs):
        vol, simg = sequential(cfg, frames, hip_device)
        assert_same(scene_state(b.volumes[s]), scene_state(vol), f"scene {s}")
        for f in frames:
            assert np.array_equal(bimg[f][s][0], simg[f][0]), f"scene {s} frame {f}: depth"
            assert np.array_equal(bimg[f][s][1], simg[f][1]), f"scene {s} frame {f}: normal"
        st = vol.status()
        assert st["flags"] == 0, st
        counts.append(st["block_count"])
        vol.close()
    print(f"[c4] 64 scenes equal; blocks per scene {min(counts)}-{max(counts)}, tiling {tune.tiling}")
    assert len(set(c.voxel_size for c in cfgs)) == 4
    b.close()


def test_raycast_shapes_are_output_invariant(hip_device):
    """Every raycast tile shape (8x8, 16x4, 4x16, 32x2 pixels per wave) gives
    bit-identical depth, vertex and normal images; mh_tune_raycast picks one
    and leaves the work counters as they were."""
    import torch
    cfg = replace(C1, frames=8)
    vol = Volume(cfg)
    for f in range(0, 8, 2):
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
    cTw = vol.pose()[1]
    n = cfg.width * cfg.height

    def cast():
        d = torch.empty(n, dtype=torch.float32, device=hip_device)
        v = torch.empty(3 * n, dtype=torch.float32, device=hip_device)
        nn = torch.empty(3 * n, dtype=torch.float32, device=hip_device)
        vol.raycast(cTw, d.data_ptr(), v.data_ptr(), nn.data_ptr())
        vol.synchronize()
        return [x.cpu().numpy().view(np.uint32) for x in (d, v, nn)]

    want = cast()
    assert np.isfinite(want[0].view(np.float32)).mean() > 0.5
    for shape in range(A.MH_NUM_RAY_SHAPES):
        vol.set_raycast_shape(shape)
        got = cast()
        for a, b in zip(got, want):
            assert np.array_equal(a, b), f"shape {shape}"
    before = vol.status()
    r = vol.tune_raycast(cTw, reps=2)
    print(f"[ray shape] {r.tile_width}x{r.tile_height}: {r.best_ms:.4f} ms vs 8x8 {r.default_ms:.4f} ms")
    assert 0 <= r.shape < A.MH_NUM_RAY_SHAPES and r.candidates == A.MH_NUM_RAY_SHAPES
    assert r.tile_width * r.tile_height == 64 and r.best_ms <= r.default_ms
    assert vol.status() == before
    for a, b in zip(cast(), want):
        assert np.array_equal(a, b)
    with pytest.raises(A.MHError):
        vol.set_raycast_shape(A.MH_NUM_RAY_SHAPES)
    vol.close()


def meshed_run(cfg, frames, shape=None, tune_reps=None):
    vol = Volume(cfg)
    if shape is not None:
        vol.set_marching_cubes_shape(shape)
    if tune_reps is not None:
        vol.tune_marching_cubes(tune_reps)
    for f in frames:
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
        vol.marching_cubes()
    return vol


def test_marching_cubes_shapes_are_output_invariant(hip_device):
    """Every marching-cubes launch shape (surfel workgroups per CU, static
    share of the worklist) leaves the same canonical mesh, cube indices,
    vertex payloads and counters."""
    cfg = replace(C1, frames=8)
    frames = range(0, 8, 2)
    want = scene_state(meshed_run(cfg, frames))
    for k in range(A.MH_NUM_MC_SHAPES):
        print(f"[mc shape] {k}", flush=True)
        vol = meshed_run(cfg, frames, shape=k)
        assert_same(scene_state(vol), want, f"mc shape {k}")
        vol.close()
    with pytest.raises(A.MHError):
        meshed_run(cfg, frames, shape=A.MH_NUM_MC_SHAPES)


def test_tune_marching_cubes_online(hip_device):
    """Online tuning: the first mc_tune_calls(1) meshing calls alternate shape
    0 with each other shape, the next one keeps the chosen shape; the mesh is
    the untuned one."""
    cfg = replace(C1, frames=20)
    frames = range(mc_tune_calls(1) + 1)
    ref = scene_state(meshed_run(cfg, frames))
    vol = meshed_run(cfg, frames, tune_reps=1)
    r = vol.marching_cubes_tuning()
    print(f"[mc tune] shape {r.shape} ({r.workgroups_per_cu} wg/CU, {r.static_quarters}/4 static): "
          f"{1e3 * r.best_ms:.3f} us/block vs default {1e3 * r.default_ms:.3f} us/block")
    assert r.done == 1 and 0 <= r.shape < A.MH_NUM_MC_SHAPES and r.candidates == A.MH_NUM_MC_SHAPES
    assert r.best_ms <= r.default_ms
    assert_same(scene_state(vol), ref, "tuned")
    vol.close()


def test_batch_tunes_marching_cubes_over_graph_replays(hip_device):
    """A batch's online marching-cubes tuning alternates the shapes over its
    graph replays (a new graph each time the shape changes) and then keeps
    one: every replay's images and the final scenes equal an untuned run."""
    cfgs = c4_scenes(n=4, frames=20, seed=3)
    frames = range(mc_tune_calls(1) + 2)
    b0, img0, _ = batched(cfgs, frames, hip_device, graph=True)
    b1, img1, _ = batched(cfgs, frames, hip_device, graph=True, mc_tune=True)
    r = b1.marching_cubes_tuning()
    print(f"[batch mc tune] shape {r.shape} ({r.workgroups_per_cu} wg/CU, {r.static_quarters}/4 static): "
          f"{r.best_ms:.4f} ms vs default {r.default_ms:.4f} ms per replay")
    assert r.done == 1 and 0 <= r.shape < A.MH_NUM_MC_SHAPES and r.candidates == A.MH_NUM_MC_SHAPES
    assert r.best_ms <= r.default_ms and r.blocks_per_ms == 0.0
    for f in frames:
        for s in range(len(cfgs)):
            assert np.array_equal(img0[f][s][0], img1[f][s][0]), f"frame {f} scene {s}: depth"
            assert np.array_equal(img0[f][s][1], img1[f][s][1]), f"frame {f} scene {s}: normal"
    for s in range(len(cfgs)):
        assert_same(scene_state(b1.volumes[s]), scene_state(b0.volumes[s]), f"scene {s}")
    b0.close()
    b1.close()


def test_tune_psdf_integration_changes_nothing(hip_device):
    """mh_tune_integrate on a PSDF volume tunes the PSDF fusion's workgroup
    size and grid with mu / (1/sigma^2) / a / b and the counters restored
    after every trial: the fused state equals the untuned run's bit for bit,
    and so does every fixed PSDF tiling."""
    from meshhashing_amd.configs import C3
    cfg = replace(C3, frames=8)
    frames = range(0, 8, 2)

    def run(tune_at=None, tiling=None):
        vol = Volume(cfg)
        if tiling is not None:
            vol.set_integrate_tiling(*tiling)
        res = None
        for f in frames:
            vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
            vol.alloc()
            vol.collect_visible()
            vol.predict_inlier()
            if f == tune_at:
                res = vol.tune_integrate(reps=2)
            vol.integrate_psdf()
        d = vol.dump_blocks()
        st = vol.status()
        vol.close()
        return {k: d[k].view(np.uint32) for k in ("sdf", "weight", "a", "b")}, st["voxel_updates"], res

    want, upd, _ = run()
    got, upd2, res = run(tune_at=2)
    print(f"[tune psdf] {res.workgroup_size} lanes x {res.workgroups_per_cu} wg/CU: {res.best_ms:.4f} ms "
          f"vs default {res.default_ms:.4f} ms")
    assert res.candidates == 15 and res.workgroup_size in (128, 256, 512) and res.best_ms <= res.default_ms
    assert upd2 == upd
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    for t in range(3):
        got, _, _ = run(tiling=(t, 4))
        for k in want:
            assert np.array_equal(got[k], want[k]), (t, k)
