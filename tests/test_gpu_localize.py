"""Point-to-PSDF localisation on the GPU (mh_point_to_psdf, mh_localize)
against the oracle.

The GPU evaluates the same FP32 per-pixel terms as the oracle and sums them
in FP64 in another fixed order, so A, b and err agree to 1e-9 relative and
the pixel count exactly.  mh_localize is replayed on the host from the
oracle's normal equations (numpy solve, FP64 exponential, FP32 pose between
steps, as the reference keeps it): poses agree to 1e-5.
"""
import numpy as np
import pytest

from meshhashing_amd.configs import C1
from meshhashing_amd.engine import Volume, synth_depth, synth_pose

pytestmark = pytest.mark.gpu


def se3_exp_tr(xi):
    u, w = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(w)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        A, B, Cc = 1.0, 0.5, 1.0 / 6.0
    else:
        A, B, Cc = np.sin(th) / th, (1 - np.cos(th)) / th ** 2, (th - np.sin(th)) / th ** 3
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + A * W + B * W @ W
    T[:3, 3] = (np.eye(3) + B * W + Cc * W @ W) @ u
    return T


def pose_error(A, B):
    D = np.linalg.inv(A.astype(np.float64)) @ B.astype(np.float64)
    c = np.clip((np.trace(D[:3, :3]) - 1) / 2, -1, 1)
    return np.degrees(np.arccos(c)), np.linalg.norm(D[:3, 3])


def fused_pair(cfg, oracle_mod):
    vol = Volume(cfg)
    for f in range(0, 8, 2):
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
    g = vol.dump_blocks()
    orc = oracle_mod.Oracle(cfg)
    orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
    return vol, orc


def perturbed(gt, rng, deg=2.0, cm=2.0):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    xi = np.concatenate([rng.normal(size=3) * cm / 100 / np.sqrt(3), axis * np.radians(deg)])
    return (se3_exp_tr(xi) @ gt).astype(np.float32)


def test_point_to_psdf_matches_oracle(hip_device, oracle_mod):
    cfg = C1
    vol, orc = fused_pair(cfg, oracle_mod)
    rng = np.random.default_rng(7)
    for f in (3, 5):
        depth, gt = synth_depth(cfg, f), synth_pose(cfg, f).astype(np.float64)
        vol.upload_depth(depth, gt)
        orc.set_frame(depth, gt)
        T = perturbed(gt, rng)
        for corrected in (False, True):
            g = vol.point_to_psdf(T, corrected=corrected)
            o = orc.point_to_psdf(T, corrected=corrected)
            print(f"[p2s] frame {f} corrected={corrected}: count {g['count']}, err {g['err']:.6e}",
                  flush=True)
            assert g["count"] == o["count"] > 10000
            assert np.abs(g["A"] - o["A"]).max() <= 1e-9 * np.abs(o["A"]).max()
            assert np.abs(g["b"] - o["b"]).max() <= 1e-9 * np.abs(o["b"]).max()
            assert abs(g["err"] - o["err"]) <= 1e-9 * o["err"]
            if not corrected:
                assert np.array_equal(g["A"][:, 2], g["A"][:, 0])


def test_localize_matches_host_replay_and_converges(hip_device, oracle_mod):
    cfg = C1
    vol, orc = fused_pair(cfg, oracle_mod)
    rng = np.random.default_rng(5)
    f = 3
    depth, gt = synth_depth(cfg, f), synth_pose(cfg, f).astype(np.float64)
    start = perturbed(gt, rng)
    for corrected, lam, steps in ((False, 1e5, 20), (True, 1e3, 20)):
        vol.upload_depth(depth, start)
        orc.set_frame(depth, start)
        out = vol.localize(start, iterations=steps, lam=lam, corrected=corrected)
        assert out["steps"] == steps
        T = start.copy()
        for it in range(steps):
            r = orc.point_to_psdf(T, corrected=corrected)
            assert abs(out["err"][it] - r["err"]) <= 1e-6 * r["err"], it
            assert out["count"][it] == r["count"], it
            x = np.linalg.solve(r["A"] + lam * np.eye(6), -r["b"])
            T = (se3_exp_tr(x) @ T.astype(np.float64)).astype(np.float32)
        assert np.abs(out["wTc"] - T).max() <= 1e-5
        r0, t0 = pose_error(gt, start)
        r1, t1 = pose_error(gt, out["wTc"])
        print(f"[localize] corrected={corrected} lambda={lam:g}: {r0:.2f} deg {100 * t0:.2f} cm -> "
              f"{r1:.3f} deg {100 * t1:.2f} cm, err/count {out['err'][0] / out['count'][0]:.2e} -> "
              f"{out['err'][-1] / out['count'][-1]:.2e}", flush=True)
        assert out["err"][-1] / out["count"][-1] < out["err"][0] / out["count"][0]
        if corrected:
            assert r1 < 0.3 and t1 < 0.01
        # the refined pose is the frame's pose from now on
        wTc, _ = vol.pose()
        assert np.array_equal(wTc, out["wTc"])
