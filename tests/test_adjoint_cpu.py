"""The raycast adjoint of the oracle against FP64 central finite differences.

The FP64 march (fp64=True) is differentiated analytically by the oracle;
here the same FP64 march is re-run at pose +-h (h = 1e-4) along each se(3)
direction (left perturbation wTc <- exp(xi^) wTc) and at voxel SDF +-h.
The march is piecewise smooth: a pixel whose accepted sample, regula-falsi
branches or evaluation cells differ between the three runs has a kink or a
jump inside +-h, so the comparison is made over the pixels whose path id is
identical in all three runs (the path id hashes exactly those choices).
Bar: 1e-4 relative (of the gradient's norm).  CPU only.
"""
from dataclasses import replace

import numpy as np

from meshhashing_amd.configs import C1
from meshhashing_amd.engine import synth_depth, synth_pose

H = 1e-4


def se3_exp(xi):
    """exp of (omega, upsilon) as a 4x4 (Rodrigues; V for the translation)."""
    w, u = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(w)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-9:
        A, B, Cc = 1.0, 0.5, 1.0 / 6.0
    else:
        A, B, Cc = np.sin(th) / th, (1 - np.cos(th)) / th ** 2, (th - np.sin(th)) / th ** 3
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + A * W + B * W @ W
    T[:3, 3] = (np.eye(3) + B * W + Cc * W @ W) @ u
    return T


def small_cfg():
    return replace(C1, width=96, height=72, fx=78.75, fy=78.75, cx=48.0, cy=36.0,
                   bucket_count=1 << 14, value_capacity=1 << 14)


def fused(oracle_mod, cfg, frames):
    o = oracle_mod.Oracle(cfg)
    for f in frames:
        o.frame(synth_depth(cfg, f), synth_pose(cfg, f), mesh=False)
    return o


def query(cfg, f):
    gt = synth_pose(cfg, f).astype(np.float64)
    xi = np.array([0.012, -0.008, 0.01, 0.015, -0.01, 0.008])
    return se3_exp(xi) @ gt, (synth_depth(cfg, f) * np.float32(0.001)).astype(np.float32)


def test_sample_tangent_identities():
    """se3_exp: a pure translation and a rotation about z; exp(-xi) exp(xi) = I."""
    T = se3_exp([0, 0, 0, 0.1, -0.2, 0.3])
    assert np.allclose(T[:3, 3], [0.1, -0.2, 0.3]) and np.allclose(T[:3, :3], np.eye(3))
    R = se3_exp([0, 0, np.pi / 2, 0, 0, 0])[:3, :3]
    assert np.allclose(R @ [1, 0, 0], [0, 1, 0])
    xi = np.array([0.3, -0.1, 0.2, 0.5, 0.4, -0.3])
    assert np.allclose(se3_exp(-xi) @ se3_exp(xi), np.eye(4))


def test_pose_adjoint_matches_central_differences(oracle_mod):
    """Plain central differences at h = 1e-4 carry a truncation error c h^2
    that is large on a few grazing pixels (the regula-falsi depth is a
    rational function with small denominators there; the error falls 100x
    per decade of h).  The comparison therefore uses the Richardson
    extrapolation of the h = 1e-4 and h / 2 central differences (error
    O(h^4)) and plain central differences at h = 1e-6, each over the pixels
    on which that difference has itself converged (|D(h) - D(h/2)| <=
    tol |D| + 1e-6, a property of the finite difference alone; tol = 1e-3
    for the O(h^4) extrapolation, 1e-5 for the plain O(h^2) difference)."""
    cfg = small_cfg()
    o = fused(oracle_mod, cfg, range(0, 8, 2))
    wTc, obs = query(cfg, 3)
    base = o.raycast_adjoint(True, wTc, obs, jac=True)
    hit = (base["path"] >= 0) & (obs > 0)
    assert hit.sum() > 2000, hit.sum()
    r = np.where(hit, base["depth"] - obs, 0.0)

    def central(j, h):
        e = np.zeros(6)
        e[j] = h
        p = o.raycast_adjoint(True, se3_exp(e) @ wTc, obs)
        m = o.raycast_adjoint(True, se3_exp(-e) @ wTc, obs)
        same = (p["path"] == base["path"]) & (m["path"] == base["path"])
        with np.errstate(invalid="ignore"):  # -inf - -inf off the surface
            return (p["depth"] - m["depth"]) / (2 * h), same

    errs, kept = {}, {}
    for name, h, tol in (("richardson(1e-4)", H, 1e-3), ("h=1e-6", 1e-6, 1e-5),
                         ("plain h=1e-4", H, None)):
        adj, fd = np.zeros(6), np.zeros(6)
        k = []
        for j in range(6):
            d1, s1 = central(j, h)
            d2, s2 = central(j, h / 2)
            use = hit & s1 & s2
            if tol is not None:
                use &= np.abs(d1 - d2) <= tol * np.abs(d2) + 1e-6
            d = (4 * d2 - d1) / 3 if name.startswith("richardson") else d1
            k.append(use.sum() / hit.sum())
            fd[j] = np.sum(r[use] * d[use])
            adj[j] = np.sum(r[use] * base["jac"][..., j][use])
        errs[name] = np.abs(adj - fd).max() / np.abs(adj).max()
        kept[name] = min(k)
    print(f"[adjoint] pose grad rel err {errs}, pixels kept {kept}")
    assert min(kept.values()) > 0.5
    assert errs["richardson(1e-4)"] < 1e-4 and errs["h=1e-6"] < 1e-4
    # the summary gradient is the same sum over every used pixel
    full = np.sum(r[hit][:, None] * base["jac"][hit][:, :6], axis=0)
    assert np.allclose(base["grad"], full, rtol=1e-12, atol=1e-12)
    J = base["jac"][hit][:, :6]
    assert np.allclose(base["hessian"], J.T @ J, rtol=1e-10)


def test_sdf_adjoint_matches_central_differences(oracle_mod):
    """dL/dsdf of the voxels with the largest gradients against the
    Richardson extrapolation of central differences of L at h = 1e-4 and
    h / 2 (voxels for which some pixel changes path within +-h skipped)."""
    cfg = small_cfg()
    o = fused(oracle_mod, cfg, range(0, 8, 2))
    wTc, obs = query(cfg, 3)
    base = o.raycast_adjoint(True, wTc, obs, grad_sdf=True)
    G = base["grad_sdf"]
    cand = np.argsort(-np.abs(G).ravel())[:40]
    checked = 0
    worst = 0.0

    def central(b, li, h):
        o.perturb_sdf(b, li, h)
        p = o.raycast_adjoint(True, wTc, obs)
        o.perturb_sdf(b, li, -2 * h)
        m = o.raycast_adjoint(True, wTc, obs)
        o.perturb_sdf(b, li, h)
        same = np.array_equal(p["path"], base["path"]) and np.array_equal(m["path"], base["path"])
        return (p["loss"] - m["loss"]) / (2 * h), same

    for flat in cand:
        b, li = divmod(int(flat), 512)
        d1, s1 = central(b, li, H)
        d2, s2 = central(b, li, H / 2)
        if not (s1 and s2):
            continue  # some pixel changes branch or cell within +-h
        fd = (4 * d2 - d1) / 3
        worst = max(worst, abs(fd - G[b, li]) / np.abs(G).max())
        checked += 1
    print(f"[adjoint] sdf grad: {checked} voxels checked, worst rel err {worst:.2e}")
    assert checked >= 10
    assert worst < 1e-4


def test_fp32_march_tangents_close_to_fp64(oracle_mod):
    """The FP32-decision adjoint (the GPU's parity target) differentiates the
    same function: on pixels with the same path its tangents agree with the
    FP64 march's to float precision."""
    cfg = small_cfg()
    o = fused(oracle_mod, cfg, range(0, 8, 2))
    wTc, obs = query(cfg, 3)
    a64 = o.raycast_adjoint(True, wTc, obs, jac=True)
    cTw32 = np.linalg.inv(wTc).astype(np.float32)
    a32 = o.raycast_adjoint(False, cTw32, obs, jac=True)
    both = (a64["path"] >= 0) & (a32["path"] >= 0)
    assert both.mean() > 0.5
    dd = np.abs(a64["depth"][both] - a32["depth"][both])
    assert np.median(dd) < 1e-5
    rel = np.abs(a64["jac"][both][:, :6] - a32["jac"][both][:, :6]).sum(1) / \
        (np.abs(a64["jac"][both][:, :6]).sum(1) + 1e-9)
    assert np.median(rel) < 1e-3
