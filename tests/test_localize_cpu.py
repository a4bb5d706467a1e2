"""The oracle's point-to-PSDF normal equations (PointToSurfaceKernel,
localizing/point_to_psdf.cu:19-82), pinned on a tilted plane where every
quantity has a closed form: trilinear interpolation of a linear field is
exact and so is its central difference, so d = n.p - c and g = n at every
world point p = wTc * reproject(pixel, depth).  Then

    A = sum J J^T,  b = sum d J,  err = sum d^2,  J = (g, p x g)

over the odd pixels with a valid depth (the reference's J puts g.x where
g.z belongs, :69-74; mode 0 keeps that).  Pixels whose world point lies
within 1e-4 voxel of a voxel or half-voxel plane are given no depth: there
GetSpatialValue's FP32 corner (rounded from pos - vs/2) and ratio
(frac(pos / vs)) can disagree by a voxel, as in the reference, and the
central difference is no longer the plane's slope.  b is also half the
derivative of err under a left pose perturbation (translation, rotation),
checked by central differences of the oracle's own err.
"""
from dataclasses import replace

import numpy as np

from meshhashing_amd.configs import C1


def se3_exp_tr(xi):
    """exp of the tangent (translation u, rotation w), as Sophus orders it."""
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


def plane_setup(oracle_mod, margin=1e-4):
    cfg = replace(C1, width=160, height=120, fx=131.25, fy=131.25, cx=80.0, cy=60.0,
                  bucket_count=1 << 14, value_capacity=1 << 14)
    n = np.array([0.2, -0.3, 1.0])
    n /= np.linalg.norm(n)
    c = float(n @ np.array([0.0, 0.0, 1.5]))
    ys, xs = np.mgrid[0:cfg.height, 0:cfg.width]
    ray = np.stack([(xs - cfg.cx) / cfg.fx, (ys - cfg.cy) / cfg.fy, np.ones_like(xs, float)], -1)
    z = c / (ray @ n)
    raw = np.round(z * 1000).astype(np.uint16)
    raw[::7, ::5] = 0                     # holes
    raw[40:44, :] = 6000                  # beyond the depth range
    depth = np.float32(0.001) * raw.astype(np.float32)
    valid = (raw > 0) & (depth <= np.float32(5.0))
    # the world points under the evaluation pose, and every block within two
    # voxels of one of them
    wTc = se3_exp_tr([0.01, -0.02, 0.015, 0.01, 0.02, -0.015])
    pc = ray * depth[..., None].astype(np.float64)
    pw = pc @ wTc[:3, :3].T + wTc[:3, 3]
    vs = cfg.voxel_size
    f = np.mod(pw / np.float32(vs), 0.5)
    tie = (np.minimum(f, 0.5 - f) < margin).any(-1)
    raw[tie] = 0
    valid &= ~tie
    vox = np.floor(pw[valid] / vs).astype(np.int64)
    blocks = set()
    for d in np.stack(np.meshgrid(*[np.arange(-2, 3)] * 3, indexing="ij"), -1).reshape(-1, 3):
        blocks.update(map(tuple, np.floor_divide(vox + d, 8).tolist()))
    pos = np.array(sorted(blocks), np.int32)
    li = np.arange(512)
    loc = np.stack([li % 8, (li // 8) % 8, li // 64], -1)
    X = (pos[:, None, :] * 8 + loc[None]).astype(np.float64) * np.float32(vs)
    sdf = (X @ n - c).astype(np.float32)
    orc = oracle_mod.Oracle(cfg)
    orc.load_blocks(pos, sdf, np.ones_like(sdf))
    orc.set_frame(raw, np.eye(4))
    return cfg, orc, n, c, wTc, pw, valid


def closed_form(n, c, pts, reference_bug):
    d = pts @ n - c
    J = np.concatenate([np.broadcast_to(n, pts.shape), np.cross(pts, n)], axis=1)
    if reference_bug:
        J[:, 2] = n[0]
    return J.T @ J, J.T @ d, float(d @ d), len(d)


def test_point_to_psdf_plane_closed_form(oracle_mod):
    cfg, orc, n, c, wTc, pw, valid = plane_setup(oracle_mod)
    odd = np.zeros_like(valid)
    odd[1::2, 1::2] = True
    pts = pw[valid & odd]
    for mode in (True, False):
        got = orc.point_to_psdf(wTc.astype(np.float32), corrected=mode)
        A, b, err, cnt = closed_form(n, c, pts, reference_bug=not mode)
        assert got["count"] == cnt > 3000
        assert np.abs(got["A"] - A).max() <= 1e-5 * np.abs(A).max()
        assert np.abs(got["b"] - b).max() <= 1e-4 * np.abs(b).max()
        assert abs(got["err"] - err) <= 1e-4 * err
        assert np.allclose(got["A"], got["A"].T, rtol=0, atol=0)
    # the reference's J: third column equals the first
    ref = orc.point_to_psdf(wTc.astype(np.float32), corrected=False)
    assert np.array_equal(ref["A"][:, 2], ref["A"][:, 0])


def test_point_to_psdf_b_is_half_the_error_gradient(oracle_mod):
    # the perturbed poses move points by < 0.004 voxel: keep 0.01 voxel away
    # from the planes where the FP32 corner choice flips
    cfg, orc, n, c, wTc, pw, valid = plane_setup(oracle_mod, margin=1e-2)
    g = orc.point_to_psdf(wTc.astype(np.float32), corrected=True)
    assert g["count"] > 2000
    h = 3e-5
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        ep = orc.point_to_psdf((se3_exp_tr(e) @ wTc).astype(np.float32), corrected=True)
        em = orc.point_to_psdf((se3_exp_tr(-e) @ wTc).astype(np.float32), corrected=True)
        assert ep["count"] == em["count"] == g["count"]
        fd[k] = (ep["err"] - em["err"]) / (2 * h)
    assert np.abs(0.5 * fd - g["b"]).max() <= 2e-3 * np.abs(g["b"]).max(), (0.5 * fd, g["b"])
