"""Per-frame reconstruction benchmark on one MI355X.

Workload: the full-rate plain pipeline (640x480 uint16 depth, 5 mm voxels,
2 cm truncation, 2^20 x 10 hash, 8 analytic primitives, spline trajectory,
2 % holes).  One step = one frame of
    depth upload (already in HBM) -> AllocBlockArray -> CollectBlocksInFrustum
    -> UpdateBlocksSimple -> MarchingCubes -> RayCaster::Cast (640x480).
All input frames are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--no-batched] [--no-pose]

Prints one JSON line.  `value` is voxel-updates/s of the whole chain and
`ms_per_step` ms/frame, with each frame's chain replayed as one HIP graph.
The same frames run eagerly, stage by stage, give per-stage times (HIP
events on the volume's stream) and the roofline.  `batched` holds the
batched form: 64 independent 320x240 scenes with voxel sizes 1-4 cm, per-
scene throughput of the fused batch against one-scene-at-a-time runs.
`pose_refinement` holds the differentiable raycast: Gauss-Newton pose
refinement of 32 perturbed 640x480 views and the adjoint's error against
the CPU FP64 adjoint, beside the reference's own point-to-PSDF localiser
(20 steps, its Jacobian and lambda, and the corrected Jacobian) from the
same starting poses.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s


def algorithmic_bytes(stage, c, frames, w, h):
    """Algorithmic bytes of one stage over the timed frames:
    integrate 24 B per voxel-update + 24 B per worklist block + 6 B per
    pixel (raw and metric depth); marching cubes 729 * 8 B per block (the
    9^3 shell of sdf and weight) + 1 B per cube index + 28 B per vertex
    written + 12 B per triangle; raycast 16 B per pixel + 4 KiB per distinct
    block sampled."""
    if stage == "integrate":
        return 24 * c["voxel_updates"] + 24 * c["blocks_integrated"] + 6 * w * h * frames
    if stage == "mesh":
        return (729 * 8 + 512) * c["mc_blocks"] + 28 * c["mc_vertex_writes"] + 12 * c["mc_triangle_writes"]
    if stage == "raycast":
        if c.get("ray_blocks") is None:
            return None
        return 16 * w * h * frames + 512 * 8 * c["ray_blocks"]
    return None


KERNEL_OF_STAGE = {"integrate": "integrate_tsdf_kernel", "mesh": "mc_surfel_kernel",
                   "raycast": "raycast_kernel"}


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (each scaled by the counter's reading on
    a copy of known size); None if the summary does not cover it."""
    path = ROOT / "profiles" / "c2_640x480_5mm_traffic.json"
    if not path.exists():
        return None
    prof = json.loads(path.read_text())
    k = prof["kernels"].get(kernel)
    if not k or k.get("hbm_bytes") is None:
        return None
    return {"bytes_per_launch": k["hbm_bytes"], "kernel": kernel,
            "fetch_factor": prof["fetch_factor"], "write_factor": prof["write_factor"],
            "l2_hit_rate": k.get("l2_hit_rate"), "source": "profiles/c2_640x480_5mm_traffic.json"}


def host_cores():
    """Host cores this process may use, capped at the 16 a one-GPU share of
    the host gives it."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(cfg, runs=5):
    """The FP64 oracle on the same 640x480 sequence, full chain per frame
    (alloc, collect, TSDF, marching cubes, raycast).  Its block- and
    pixel-parallel stages (integrate, marching cubes pass 1, raycast) run on
    all host cores (OpenMP); allocation and the triangle / vertex
    bookkeeping stay serial.  Frame 0 warms up; frames 1..runs are timed one
    by one and the median frame is reported; then one more frame on one
    thread."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    from meshhashing_amd.engine import synth_depth, synth_pose

    cores = host_cores()
    orc.set_threads(cores)
    o = orc.Oracle(cfg)
    rp = cfg.raycaster_params()

    def frame(f):
        d, p = synth_depth(cfg, f), synth_pose(cfg, f)
        u0 = o.voxel_updates()
        t0 = time.perf_counter()
        o.frame(d, p, mesh=True)
        o.raycast(np.linalg.inv(p.astype(np.float64)).astype(np.float32), rp)
        return time.perf_counter() - t0, o.voxel_updates() - u0

    frame(0)
    timed = [frame(f) for f in range(1, runs + 1)]
    rates = sorted(u / t for t, u in timed)
    ms = sorted(1e3 * t for t, _ in timed)
    orc.set_threads(1)
    t1, u1 = frame(runs + 1)
    cpu = platform.processor() or platform.machine()
    return {"value": rates[len(rates) // 2], "unit": "voxel-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"frames 1-{runs} of the same 640x480 sequence after a warm-up frame, full "
                      f"chain (alloc, collect, TSDF, marching cubes, raycast), FP64 oracle, "
                      f"integrate / MC pass 1 / raycast on {cores} threads, median of {runs} frames; "
                      f"{os.cpu_count()} host cores visible, {cpu}",
            "ms_per_frame": ms[len(ms) // 2],
            "one_thread": {"value": u1 / t1, "ms_per_frame": 1e3 * t1, "cores": 1,
                           "sample": f"frame {runs + 1} of the same sequence, one run"}}


def wall_clock(fn):
    """Host milliseconds of fn, between two device-wide synchronisations."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0)


def c4_batched(dev, stream, n_scenes=64, n_frames=4):
    """Batched form: n_scenes independent rooms (voxel sizes 1-4 cm), one frame
    of all scenes per step, fused per stage and replayed as a HIP graph,
    against the same scene-frames run one scene after another through the
    single-volume entry points.  Both timed after a full warm-up pass and a
    reset of every volume."""
    import torch

    from meshhashing_amd.configs import c4_scenes
    from meshhashing_amd.engine import Batch, Volume, synth_depth, synth_pose

    cfgs = c4_scenes(n_scenes, n_frames)
    W, H = cfgs[0].width, cfgs[0].height
    depth = np.stack([np.stack([synth_depth(c, f) for c in cfgs]) for f in range(n_frames)])
    poses = np.stack([np.stack([synth_pose(c, f) for c in cfgs]) for f in range(n_frames)])
    depth_dev = torch.from_numpy(depth.view(np.int16)).to(dev)
    fbytes = n_scenes * W * H * 2
    rd = torch.empty(n_scenes * W * H, dtype=torch.float32, device=dev)
    rn = torch.empty(3 * n_scenes * W * H, dtype=torch.float32, device=dev)

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    # batched
    bt = Batch(cfgs)
    bt.set_stream(stream.cuda_stream)
    bt.set_raycast(rd.data_ptr(), 0, rn.data_ptr())
    tune = None

    def batch_pass(graph, tune_first=False):
        nonlocal tune
        for f in range(n_frames):
            bt.frame(depth_dev.data_ptr() + f * fbytes, poses[f])
            r = bt.run(mesh=True, raycast=True, graph=graph and not (tune_first and f == 0),
                       tune=tune_first and f == 0)
            tune = r or tune

    batch_pass(True, tune_first=True)  # tunes on frame 0, captures the graph on frame 1
    torch.cuda.synchronize()
    for v in bt.volumes:
        v.reset()
    c0 = [v.status() for v in bt.volumes]
    ms_b = timed(lambda: batch_pass(True))
    st_b = [v.status() for v in bt.volumes]
    for v in bt.volumes:
        v.reset()
    ms_b_wall = wall_clock(lambda: batch_pass(True))
    upd_b = sum(b["voxel_updates"] - a["voxel_updates"] for a, b in zip(c0, st_b))
    blocks = [s["block_count"] for s in st_b]
    bt.close()

    # sequential single-volume runs of the same scene-frames
    vols = [Volume(c) for c in cfgs]
    for v in vols:
        v.set_stream(stream.cuda_stream)
    sbytes = W * H * 2

    def seq_pass():
        for s, v in enumerate(vols):
            for f in range(n_frames):
                v.set_depth_device(depth_dev.data_ptr() + f * fbytes + s * sbytes, poses[f][s])
                v.alloc()
                v.collect_visible()
                v.integrate_tsdf()
                v.marching_cubes()
                v.raycast(v.pose()[1], rd.data_ptr(), 0, rn.data_ptr())

    seq_pass()
    torch.cuda.synchronize()
    for v in vols:
        v.reset()
    ms_s = timed(seq_pass)
    for v in vols:
        v.close()

    # the same scene-frames one scene after another, each scene replaying its
    # own one-scene frame graph with the batch's tuned integration tiling:
    # the batched form's gain over the plain path with the same launch
    # machinery
    # (each batch keeps its own stream: several batches capturing and
    # replaying graphs on one shared stream were refused by the runtime, so
    # both this leg and the batched one are timed on the host clock around a
    # device-wide synchronisation)
    ones = [Batch([c]) for c in cfgs]
    for s, b1 in enumerate(ones):
        b1.set_raycast(rd.data_ptr(), 0, rn.data_ptr())
        if tune is not None:
            b1.set_integrate_tiling(tune.tiling, tune.workgroups_per_cu)

    def seq_graph_pass():
        for s, b1 in enumerate(ones):
            for f in range(n_frames):
                b1.frame(depth_dev.data_ptr() + f * fbytes + s * sbytes, poses[f][s][None])
                b1.run(mesh=True, raycast=True, graph=True)
            b1.synchronize()  # one scene after another

    seq_graph_pass()
    torch.cuda.synchronize()
    for b1 in ones:
        b1.volumes[0].reset()
    ms_g = wall_clock(seq_graph_pass)
    for b1 in ones:
        b1.close()
    sf = n_scenes * n_frames
    vs = [c.voxel_size for c in cfgs]
    return {
        "workload": f"{n_scenes} scenes x {n_frames} frames of 320x240, voxel sizes "
                    f"{ {round(x * 100): vs.count(x) for x in sorted(set(vs))} } cm, "
                    "alloc->collect->TSDF->marching cubes->raycast per frame",
        "blocks_per_scene": [min(blocks), max(blocks)],
        "batched_ms_per_step": ms_b / n_frames,
        "scene_frames_per_s_batched": sf / (ms_b * 1e-3),
        "scene_frames_per_s_sequential": sf / (ms_g * 1e-3),
        "speedup_per_scene": ms_g / ms_b_wall,
        "timing": "speedup_per_scene: host clock around device-wide synchronisations for both "
                  "legs; batched_ms_per_step: HIP events",
        "scene_frames_per_s_sequential_eager": sf / (ms_s * 1e-3),
        "speedup_vs_eager_default_tiling": ms_s / ms_b,
        "voxel_updates_per_s_batched": upd_b / (ms_b * 1e-3),
        "mode": "per-frame chain replayed as one HIP graph, integration tiling autotuned; "
                "sequential: each scene alone, replaying its own one-scene frame graph with the "
                "same tuned tiling (and, as _eager, through the single-volume entry points with "
                "the default tiling)",
        "tuned_tiling": None if tune is None else {
            "workgroup_size": tune.workgroup_size, "blocks_per_workgroup": tune.blocks_per_workgroup,
            "workgroups_per_cu": tune.workgroups_per_cu},
    }


def se3_exp(xi):
    """exp of xi = (omega, upsilon) in se(3) as a 4x4."""
    w, u = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(w)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-9:
        a, b, c = 1.0, 0.5, 1.0 / 6.0
    else:
        a, b, c = np.sin(th) / th, (1 - np.cos(th)) / th ** 2, (th - np.sin(th)) / th ** 3
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + a * W + b * W @ W
    T[:3, 3] = (np.eye(3) + b * W + c * W @ W) @ u
    return T


def c5_pose_refinement(dev, stream, n_poses=32, iterations=10, grad_check=True):
    """Differentiable raycast: the room of the 320x240 sequence fused at 5 mm
    from 640x480 frames; n_poses poses around the room perturbed by a
    rotation of N(0, 2 deg) about a random axis and a translation of
    N(0, (2 cm)^2 / 3) per axis; iterations Gauss-Newton steps each on the
    observed 640x480 depth.  grad_check: the GPU adjoint against the
    oracle's (same FP32 march, FP64 tangents) at the first query pose."""
    import torch
    from dataclasses import replace

    from meshhashing_amd.configs import C5, C5_PYRAMID, C5_PYRAMID_STEPS
    from meshhashing_amd.engine import Volume, refine_pose_pyramid, synth_depth, synth_pose

    # the 5 mm volume and its coarse copies (4 cm, 2 cm) of the same frames
    vol = Volume(C5)
    coarse = [Volume(c) for c in C5_PYRAMID]
    for v_ in [vol] + coarse:
        v_.set_stream(stream.cuda_stream)
    for f in range(C5.frames):
        for v_ in [vol] + coarse:
            v_.upload_depth(synth_depth(v_.cfg, f), synth_pose(v_.cfg, f))
            v_.alloc()
            v_.collect_visible()
            v_.integrate_tsdf()
    levels = [(v_, v_.cfg.raycaster_params(), n)
              for v_, n in zip(coarse + [vol], C5_PYRAMID_STEPS)]
    q = replace(C5, frames=n_poses)
    rng = np.random.default_rng(5)
    err0, err1, err_single, err_floor, ms = [], [], [], [], []
    loc = {"reference": ([], []), "corrected": ([], [])}
    first = None
    for i in range(n_poses):
        gt = synth_pose(q, i).astype(np.float64)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        xi = np.concatenate([axis * np.radians(rng.normal(0.0, 2.0)), rng.normal(size=3) * 0.02 / np.sqrt(3)])
        start = se3_exp(xi) @ gt
        obs_np = (synth_depth(q, i) * np.float32(0.001)).astype(np.float32)
        obs = torch.from_numpy(obs_np).to(dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = refine_pose_pyramid(levels, start.astype(np.float32), obs.data_ptr())
        ms.append(1e3 * (time.perf_counter() - t0))
        single, _ = vol.refine_pose(start.astype(np.float32), obs.data_ptr(), iterations)
        # the 5 mm objective's own optimum: refinement started at the truth
        floor, _ = vol.refine_pose(gt.astype(np.float32), obs.data_ptr(), iterations)
        # the reference's own solver from the same start: point-to-PSDF on
        # the observed frame, 20 steps at lambda 1e5 with its Jacobian
        # (MainEngine::Localizing as app/slam.cc calls it), and the
        # corrected Jacobian at lambda 4e3
        vol.upload_depth(synth_depth(q, i), start.astype(np.float32))
        found = {}
        for key, corr, lam in (("reference", False, 1e5), ("corrected", True, 4e3)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            found[key] = vol.localize(start.astype(np.float32), iterations=20, lam=lam,
                                      corrected=corr)["wTc"]
            loc[key][1].append(1e3 * (time.perf_counter() - t0))
        for lst, T in ((err0, start), (err1, out), (err_single, single), (err_floor, floor),
                       (loc["reference"][0], found["reference"]),
                       (loc["corrected"][0], found["corrected"])):
            D = np.linalg.inv(gt) @ T.astype(np.float64)
            ang = np.degrees(np.arccos(np.clip((np.trace(D[:3, :3]) - 1) / 2, -1, 1)))
            lst.append((ang, 1e3 * np.linalg.norm(D[:3, 3])))
        if first is None:
            first = (np.linalg.inv(start).astype(np.float32), obs_np, obs)
    res = {
        "workload": f"{n_poses} poses, 640x480 depth, 5 mm volume of the 320x240 room "
                    f"({vol.status()['block_count']} blocks), {sum(C5_PYRAMID_STEPS)} Gauss-Newton steps "
                    f"each (coarse to fine)",
        "initial_error_deg_mm_median": [float(np.median([e[0] for e in err0])), float(np.median([e[1] for e in err0]))],
        "final_error_deg_mm_median": [float(np.median([e[0] for e in err1])), float(np.median([e[1] for e in err1]))],
        "final_error_deg_mm_max": [float(max(e[0] for e in err1)), float(max(e[1] for e in err1))],
        "poses_worse_than_start": int(sum(e1[0] > e0[0] or e1[1] > e0[1] for e0, e1 in zip(err0, err1))),
        "poses_worse_than_start_by_0.05deg_or_1mm": int(sum(e1[0] > e0[0] + 0.05 or e1[1] > e0[1] + 1.0
                                                            for e0, e1 in zip(err0, err1))),
        "ms_per_pose": float(np.median(ms)),
        "schedule": "Gauss-Newton steps per level (4 cm, 2 cm, 5 mm copies of the room): "
                    + ", ".join(map(str, C5_PYRAMID_STEPS)),
        "objective_floor_5mm": {
            "what": "10 steps on the 5 mm volume started at the true pose: where its own optimum lies",
            "final_error_deg_mm_median": [float(np.median([e[0] for e in err_floor])),
                                          float(np.median([e[1] for e in err_floor]))],
            "final_error_deg_mm_max": [float(max(e[0] for e in err_floor)),
                                       float(max(e[1] for e in err_floor))]},
        "single_level_5mm": {
            "final_error_deg_mm_median": [float(np.median([e[0] for e in err_single])),
                                          float(np.median([e[1] for e in err_single]))],
            "final_error_deg_mm_max": [float(max(e[0] for e in err_single)),
                                       float(max(e[1] for e in err_single))]},
    }
    for key, (errs, times) in loc.items():
        res[f"point_to_psdf_{key}"] = {
            "final_error_deg_mm_median": [float(np.median([e[0] for e in errs])),
                                          float(np.median([e[1] for e in errs]))],
            "ms_per_pose": float(np.median(times)),
        }
    if grad_check:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as orc

        cTw, obs_np, obs = first
        b = vol.dump_blocks()
        o = orc.Oracle(C5)
        o.load_blocks(b["pos"], b["sdf"], b["weight"], b["color"])
        g = vol.raycast_backward(cTw, obs.data_ptr())
        ref = o.raycast_adjoint(False, cTw, obs_np)
        res["grad_rel_err_vs_cpu_fp64_adjoint"] = float(np.abs(g["grad"] - ref["grad"]).max() /
                                                        np.abs(ref["grad"]).max())
        res["grad_check_pixels"] = [g["pixels"], ref["pixels"]]
    vol.close()
    for v_ in coarse:
        v_.close()
    return res


RAY_TILES = [(8, 8), (16, 4), (4, 16), (32, 2)]  # raycast tile shapes, by index


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=280)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-pose", action="store_true")
    ap.add_argument("--stage-events", type=int, default=1)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run 5 plain 256 MiB device copies (reference dispatches "
                         "for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not "
                                   "part of it (the sharding design is described in DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from meshhashing_amd import _capi as A
    from meshhashing_amd.configs import C2
    from meshhashing_amd.engine import Batch, Volume, mc_tune_calls, synth_depth, synth_pose

    cfg = C2
    total = args.warmup + args.steps
    W, H = cfg.width, cfg.height
    t_gen = time.perf_counter()
    frames = np.stack([synth_depth(cfg, f % cfg.frames) for f in range(total)])
    poses = [synth_pose(cfg, f % cfg.frames) for f in range(total)]
    t_gen = time.perf_counter() - t_gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.Stream(device=dev)  # default priority
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    depth_dev = torch.from_numpy(frames.view(np.int16)).to(dev)  # resident in HBM
    out_depth = torch.empty(W * H, dtype=torch.float32, device=dev)
    out_normal = torch.empty(3 * W * H, dtype=torch.float32, device=dev)
    frame_bytes = W * H * 2
    base_ptr = depth_dev.data_ptr()
    K = args.steps
    tune_frame = args.warmup // 2

    # ---- 1. eager stage-by-stage run (per-stage HIP events, autotuned integration)
    vol = Volume(cfg)
    vol.set_stream(stream.cuda_stream)
    stages = ["alloc", "collect", "integrate", "mesh", "raycast"]
    ev = {s: [] for s in stages}
    tune = None

    def run_frame(f, record):
        nonlocal tune
        e = [torch.cuda.Event(enable_timing=True) for _ in range(6)] if record else None
        if record:
            e[0].record(stream)
        vol.set_depth_device(base_ptr + f * frame_bytes, poses[f])
        vol.alloc()
        if record:
            e[1].record(stream)
        vol.collect_visible()
        if f == tune_frame:
            tune = vol.tune_integrate(reps=3)
        if record:
            e[2].record(stream)
        vol.integrate_tsdf()
        if record:
            e[3].record(stream)
        vol.marching_cubes()
        if record:
            e[4].record(stream)
        if f == tune_frame:
            vol.tune_raycast(vol.pose()[1], reps=3)
        vol.raycast(vol.pose()[1], out_depth.data_ptr(), 0, out_normal.data_ptr())
        if record:
            e[5].record(stream)
            for i, s in enumerate(stages):
                ev[s].append((e[i], e[i + 1]))

    # the marching-cubes launch shape is tuned online over warm-up frames
    # 2.. (each meshes as usual; the first two allocate most blocks), if
    # there are enough of them
    mc_tune_from = 2 if args.warmup > 2 + mc_tune_calls(1) else -1
    for f in range(args.warmup):
        if f == mc_tune_from:
            vol.tune_marching_cubes(reps=1)
        run_frame(f, False)
    torch.cuda.synchronize()
    mc_tune = vol.marching_cubes_tuning()
    # per-kernel HIP events and the raycast's sampled-block count for the
    # timed frames (the graph leg below runs without them)
    vol.set_diagnostics(kernel_times=True, ray_blocks=True)
    vol.kernel_times(reset=True)
    c0 = vol.status()
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    for f in range(args.warmup, total):
        run_frame(f, bool(args.stage_events))
    t_end.record(stream)
    torch.cuda.synchronize()
    ms_eager = t_start.elapsed_time(t_end)
    c1 = vol.status()
    ktimes = vol.kernel_times(reset=True)
    cnt = {k: c1[k] - c0[k] for k in ("voxel_updates", "blocks_integrated", "mc_vertex_writes",
                                      "mc_triangle_writes", "mc_blocks", "ray_samples",
                                      "ray_blocks")}
    stage_ms = {s: sum(a.elapsed_time(b) for a, b in ev[s]) for s in stages} if args.stage_events else {}
    st = vol.status()
    vol.close()

    # roofline per kernel: the stage's algorithmic bytes per launch over the
    # kernel's own average duration (HIP events on the volume's stream)
    kernels = {}
    for stage, name in KERNEL_OF_STAGE.items():
        if name not in ktimes:
            continue
        launches, tot = ktimes[name]
        b = algorithmic_bytes(stage, cnt, K, W, H)
        avg = tot / launches
        row = {"launches": launches, "avg_us": 1e3 * avg}
        if b is not None:
            bpl = b / launches
            row.update({"algorithmic_bytes_per_launch": bpl, "GB/s": bpl / (avg * 1e-3) / 1e9,
                        "frac": bpl / (avg * 1e-3) / HBM_PEAK})
        kernels[name] = row
    other = {n: {"launches": l, "avg_us": 1e3 * t / l} for n, (l, t) in ktimes.items()
             if n not in kernels}
    roof = None
    priced = [n for n in kernels if "frac" in kernels[n]]
    if priced:
        dom = max(priced, key=lambda n: kernels[n]["avg_us"])
        k = kernels[dom]
        tr = measured_traffic(dom)
        roof = {"bound": "hbm", "kernel": dom, "achieved": k["GB/s"], "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": k["frac"],
                "traffic": None if tr is None else tr["bytes_per_launch"],
                "traffic_source": tr, "algorithmic_bytes_per_launch": k["algorithmic_bytes_per_launch"],
                "avg_kernel_us": k["avg_us"], "kernels": kernels, "other_kernels": other}
    if stage_ms:
        rows = {}
        for s in ("integrate", "mesh", "raycast"):
            b = algorithmic_bytes(s, cnt, K, W, H)
            rows[s] = {"ms_per_frame": stage_ms[s] / K,
                       "GB/s": None if b is None else b / (stage_ms[s] * 1e-3) / 1e9,
                       "frac": None if b is None else b / (stage_ms[s] * 1e-3) / HBM_PEAK}
        if roof is not None:
            roof["stages"] = rows

    # ---- 2. the same frames as one HIP graph per frame (a one-scene batch)
    bt = Batch([cfg])
    bt.set_stream(stream.cuda_stream)
    bt.set_raycast(out_depth.data_ptr(), 0, out_normal.data_ptr())
    # The replayed chain tunes its own marching-cubes shape over warm-up
    # replays 1.. (a replay can prefer another shape than eager launches);
    # with too few warm-up frames it takes the eager leg's choice.
    graph_warm = [f for f in range(1, args.warmup) if f != tune_frame]
    gmc_from = 1 if len(graph_warm) > mc_tune_calls(1) else -1
    if gmc_from < 0 and mc_tune.done:
        bt.set_marching_cubes_shape(mc_tune.shape)
    gtune = None
    for f in range(args.warmup):
        if f == gmc_from:
            bt.tune_marching_cubes(1)
        bt.frame(base_ptr + f * frame_bytes, poses[f][None])
        r = bt.run(mesh=True, raycast=True, graph=f != tune_frame, tune=f == tune_frame)
        gtune = r or gtune
    torch.cuda.synchronize()
    gmc = bt.marching_cubes_tuning()
    gv = bt.volumes[0]
    g0 = gv.status()
    fev = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    t_start.record(stream)
    wall0 = time.perf_counter()
    for i, f in enumerate(range(args.warmup, total)):
        fev[i].record(stream)
        bt.frame(base_ptr + f * frame_bytes, poses[f][None])
        bt.run(mesh=True, raycast=True, graph=True)
    fev[K].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_graph = t_start.elapsed_time(t_end)
    frame_ms = sorted(fev[i].elapsed_time(fev[i + 1]) for i in range(K))
    g1 = gv.status()
    gupd = g1["voxel_updates"] - g0["voxel_updates"]
    bt.close()

    line = {
        "metric": "ms/frame & voxel-updates/s (640x480, 5 mm) plain vs batched; HBM GB/s; grad err",
        "value": gupd / (ms_graph * 1e-3),
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_graph / K,
        "ms_per_frame_median": frame_ms[K // 2],
        "ms_per_frame_p90": frame_ms[min(K - 1, (9 * K) // 10)],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c2 full-rate plain pipeline: 640x480 depth, 5 mm voxels, trunc 2 cm, "
                               "2^20x10 hash, alloc->collect->TSDF->marching cubes->raycast 640x480, "
                               "one HIP graph per frame",
                   "frames": total, "voxel_size": cfg.voxel_size, "truncation": cfg.truncation,
                   "bucket_count": cfg.bucket_count, "parallelism": "single GPU"},
        "roofline": roof,
        "eager": {"ms_per_step": ms_eager / K, "voxel_updates_per_s": cnt["voxel_updates"] / (ms_eager * 1e-3),
                  "stage_ms_per_frame": {s: v / K for s, v in stage_ms.items()}},
        "integrate_tiling": None if tune is None else {
            "workgroup_size": tune.workgroup_size, "blocks_per_workgroup": tune.blocks_per_workgroup,
            "workgroups_per_cu": tune.workgroups_per_cu, "tuned_ms": tune.best_ms,
            "default_ms": tune.default_ms},
        "raycast_tile": None if gtune is None or gtune.ray_shape < 0 else
        "{0}x{1}".format(*RAY_TILES[gtune.ray_shape]),
        "marching_cubes_shape": None if not mc_tune.done else {
            "workgroups_per_cu": mc_tune.workgroups_per_cu or "all resident",
            "static_quarters": mc_tune.static_quarters,
            "tuned_us_per_block": 1e3 * mc_tune.best_ms, "default_us_per_block": 1e3 * mc_tune.default_ms},
        "marching_cubes_shape_graph": None if not gmc.done else {
            "workgroups_per_cu": gmc.workgroups_per_cu or "all resident",
            "static_quarters": gmc.static_quarters,
            "tuned_ms_per_replay": gmc.best_ms, "default_ms_per_replay": gmc.default_ms},
        "graph_matches_eager_updates": gupd == cnt["voxel_updates"],
        "counters_per_frame": {k: v / K for k, v in cnt.items()},
        "blocks_allocated": st["block_count"],
        "triangles": st["triangle_count"],
        "status_flags": st["flag_names"],
        "host_wall_ms_per_frame": 1e3 * wall / K,
        "input_generation_s": t_gen,
    }
    if not args.no_batched:
        line["batched"] = c4_batched(dev, stream)
    if not args.no_pose:
        line["pose_refinement"] = c5_pose_refinement(dev, stream, grad_check=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
