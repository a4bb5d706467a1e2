"""Pins of the FP64 oracle against what fixes its results independently:
the reference's own tables (golden fixture), closed forms, textbook
identities, invariants and brute force.  CPU only.
"""
import json
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from meshhashing_amd.configs import C1, C3

GOLDEN = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ hashing
def bucket_ref(x, y, z, n):
    """hash_table.h:259-269 evaluated with Python integers: 32-bit wrapping
    products (two's complement), xor, then an unsigned modulo."""
    m = 0xFFFFFFFF
    h = ((x * 73856093) & m) ^ ((y * 19349669) & m) ^ ((z * 83492791) & m)
    return h % n


@pytest.mark.parametrize("n", [1 << 17, 1 << 20, 100000, 500009])
def test_hash_bucket_kat(oracle_mod, n):
    rng = np.random.default_rng(1)
    pts = [(0, 0, 0), (1, 0, 0), (-1, -1, -1), (2 ** 20 - 1, -(2 ** 20 - 1), 7)]
    pts += [tuple(int(v) for v in rng.integers(-5000, 5000, 3)) for _ in range(500)]
    for p in pts:
        assert oracle_mod.hash_bucket(*p, n) == bucket_ref(*p, n)
    # a hand-computed value: (1,2,3) -> 73856093 ^ 38699338 ^ 250478373
    assert bucket_ref(1, 2, 3, 1 << 32) == (73856093 ^ 38699338 ^ 250478373)
    # signed negative coordinates go through the unsigned modulo, never < 0
    assert all(0 <= oracle_mod.hash_bucket(-x, -1, -1, n) < n for x in range(1, 50))


def test_pose_inverse(oracle_mod):
    rng = np.random.default_rng(2)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        T = np.eye(4)
        T[:3, :3] = R
        T[:3, 3] = rng.uniform(-3, 3, 3)
        inv = oracle_mod.invert4f(T.astype(np.float32))
        exact = np.linalg.inv(T.astype(np.float32).astype(np.float64))
        assert np.abs(inv - exact).max() < 2e-6


# ------------------------------------------------------------------ MC tables
def test_mc_tables_match_reference_fixture(oracle_mod):
    g = json.loads((GOLDEN / "mc_tables.json").read_text())
    L = oracle_mod.lib()
    for c in range(256):
        row = [L.orc_mc_tri_edge(c, k) for k in range(3 * L.orc_mc_ntri(c))]
        assert row == g["triangles"][c], c
        assert L.orc_mc_edge_mask(c) == g["cube_edges"][c], c
    for i in range(8):
        assert [L.orc_mc_corner(i, a) for a in range(3)] == g["corner_offsets"][i]
    for e in range(12):
        a, b = L.orc_mc_edge_a(e), L.orc_mc_edge_b(e)
        assert [a, b] == g["edge_endpoints"][e]
        ca = [L.orc_mc_corner(a, k) for k in range(3)]
        cb = [L.orc_mc_corner(b, k) for k in range(3)]
        axis = [k for k in range(3) if ca[k] != cb[k]]
        assert len(axis) == 1
        assert [min(u, v) for u, v in zip(ca, cb)] + axis == g["edge_owner_offset_axis"][e]


def test_mc_tables_brute_force(oracle_mod):
    """Every triangle of every configuration uses only edges whose endpoints
    straddle the surface, and every straddling edge is used."""
    L = oracle_mod.lib()
    for c in range(256):
        inside = [(c >> i) & 1 for i in range(8)]
        straddle = {e for e in range(12)
                    if inside[L.orc_mc_edge_a(e)] != inside[L.orc_mc_edge_b(e)]}
        used = {L.orc_mc_tri_edge(c, k) for k in range(3 * L.orc_mc_ntri(c))}
        assert used == straddle, c
        assert L.orc_mc_ntri(c) <= 5


def field_volume(oracle_mod, cfg, fn, lo, hi):
    """Oracle volume whose blocks cover voxels [lo, hi) on every axis, with
    sdf = fn(world position of each voxel) in FP32 and weight 50."""
    b0, b1 = lo // 8, -(-hi // 8)
    r = np.arange(b0, b1)
    pos = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    i = np.arange(512)
    off = np.stack([i % 8, (i // 8) % 8, i // 64], -1)
    vox = pos[:, None, :] * 8 + off[None]
    vs = float(np.float32(cfg.voxel_size))
    sdf = fn(vox * vs).astype(np.float32)
    o = oracle_mod.Oracle(cfg)
    o.load_blocks(pos, sdf, np.full(sdf.shape, 50.0, np.float32))
    o.collect_all()
    o.marching_cubes()
    return o, vs


def test_mc_plane_vertices_closed_form(oracle_mod):
    """A linear SDF makes the edge interpolation exact: every vertex lies on
    the plane, except where an endpoint's |sdf| < 0.008 snaps the vertex to
    that endpoint (marching_cubes.cu:20-21), checked in the canonical order
    (owner voxel first)."""
    cfg = plane_cfg()
    n = np.array([0.3, -0.5, 0.81])
    n /= np.linalg.norm(n)
    c0 = 0.2137
    o, vs = field_volume(oracle_mod, cfg, lambda p: p @ n - c0, 0, 32)
    m = o.dump_mesh()
    assert len(m["vkey"]) > 300
    snapped = 0
    for key, val in zip(m["vkey"], m["vval"]):
        a = key[:3].astype(np.float64)
        b = a.copy()
        b[key[3]] += 1
        pa, pb = a * vs, b * vs
        sa = float(np.float32(pa @ n - c0))
        sb = float(np.float32(pb @ n - c0))
        assert (sa < 0) != (sb < 0)
        if abs(sa) < 0.008:
            assert np.allclose(val[:3], pa, atol=1e-6)
            snapped += 1
        elif abs(sb) < 0.008:
            assert np.allclose(val[:3], pb, atol=1e-6)
            snapped += 1
        else:
            assert abs(val[:3] @ n - c0) < 2e-6
        # the normal is the normalised SDF gradient (its +-vs/2 stencil needs
        # every trilinear corner, so skip vertices next to the region's faces)
        if key[:3].min() >= 2 and key[:3].max() <= 29:
            assert np.allclose(val[3:6], n, atol=1e-5)
    assert 0 < snapped < len(m["vkey"])


def test_mc_sphere_is_closed_surface(oracle_mod):
    """A fully observed sphere gives a closed 2-manifold: every mesh edge is
    shared by exactly two triangles and V - E + F = 2."""
    cfg = plane_cfg()
    ctr = np.array([0.331, 0.272, 0.419])
    o, vs = field_volume(oracle_mod, cfg, lambda p: np.linalg.norm(p - ctr, axis=-1) - 0.15, 0, 48)
    m = o.dump_mesh()
    tri = m["tri"].reshape(-1, 3, 4)
    assert (tri[:, :, 0] != np.iinfo(np.int32).min).all()
    assert len(tri) > 500
    keys = {tuple(v) for t in tri for v in t}
    assert keys == {tuple(k) for k in m["vkey"]}
    edges = {}
    for t in tri:
        vs3 = [tuple(v) for v in t]
        assert len(set(vs3)) == 3
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = tuple(sorted((vs3[a], vs3[b])))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) == {2}
    assert len(keys) - len(edges) + len(tri) == 2
    # reference counts are the number of triangles using each vertex
    use = {}
    for t in tri:
        for v in t:
            use[tuple(v)] = use.get(tuple(v), 0) + 1
    assert all(use[tuple(k)] == r for k, r in zip(m["vkey"], m["vref"]))


# ------------------------------------------------------------------ TSDF
def plane_cfg(**kw):
    base = replace(C1, width=64, height=48, fx=60.0, fy=60.0, cx=32.0, cy=24.0,
                   bucket_count=1 << 12, value_capacity=1 << 14)
    return replace(base, **kw)


def test_tsdf_fronto_parallel_plane_closed_form(oracle_mod):
    """A plane at z = D seen head-on from the identity pose: every updated
    voxel holds clamp(D - z, +-trunc) and weight N * max(100 (1 - nd), 1)."""
    cfg = plane_cfg()
    D = 1.234
    depth = np.full((cfg.height, cfg.width), int(round(D * 1000)), np.uint16)
    o = oracle_mod.Oracle(cfg)
    for _ in range(3):
        o.frame(depth, np.eye(4), mesh=False)
    d = o.dump_blocks()
    Dm = np.float32(0.001) * np.float32(1234)
    dmin, dmax = float(np.float32(0.1)), float(np.float32(5.0))
    nd = (float(Dm) - dmin) / (dmax - dmin)
    w1 = max(100.0 * (1.0 - nd), 1.0)
    upd = d["weight"] > 0
    assert upd.sum() > 1000
    vz = (d["pos"][:, 2:3] * 8 + np.arange(512)[None, :] // 64) * np.float64(np.float32(0.02))
    tr = float(np.float32(0.08))  # the truncation is an FP32 parameter
    expect = np.clip(float(Dm) - vz, -tr, tr)
    assert np.abs(d["sdf"][upd] - expect[upd]).max() < 1e-9
    assert np.allclose(d["weight"][upd], 3 * w1, rtol=1e-12)
    # every voxel that received a value is within the truncation band in front
    # of the plane or behind it by less than the truncation distance
    assert (vz[upd] <= float(Dm) + 0.08 + 1e-6).all()


def test_tsdf_colour_closed_form(oracle_mod):
    """Colour fusion (preprocess.cu:44-58, update_simple.cu:68-73,
    voxel.h:107-115) on a plane: every updated voxel after two frames of a
    constant BGRA image holds c2 = uchar(0.5 c1 + 0.5 C + 0.5) with
    c1 = uchar(0.5 * 0 + 0.5 C + 0.5) and C = uchar(255 * (c / 255)) per
    channel, all in float32; a pixel with a zero channel contributes 0 and,
    with no colour image at all, (0, 255, 0)."""
    cfg = plane_cfg()
    depth = np.full((cfg.height, cfg.width), 1234, np.uint16)
    bgra = np.zeros((cfg.height, cfg.width, 4), np.uint8)
    bgra[...] = (37, 200, 254, 255)  # B, G, R, A
    bgra[:, : cfg.width // 2, 2] = 0  # left half invalid (R == 0)
    f32 = np.float32

    def fuse(c):
        return np.trunc(f32(255) * (f32(c) / f32(255))).astype(np.int64)

    def blend(old, new):
        return np.floor(f32(0.5) * f32(old) + f32(0.5) * f32(new) + f32(0.5)).astype(np.int64)

    C = np.array([fuse(254), fuse(200), fuse(37)])  # stored R, G, B
    o = oracle_mod.Oracle(cfg)
    for _ in range(2):
        o.set_frame(depth, np.eye(4))
        o.set_color(bgra)
        o.alloc()
        o.collect_visible()
        o.integrate_tsdf()
    d = o.dump_blocks()
    upd = d["weight"] > 0
    col = d["color"][upd]
    got = np.stack([col & 255, (col >> 8) & 255, (col >> 16) & 255], -1)
    valid_want = blend(blend(0, C), C)
    is_valid = (got == valid_want).all(-1)
    is_invalid = (got == 0).all(-1)
    assert (is_valid | is_invalid).all()
    assert is_valid.sum() > 500 and is_invalid.sum() > 500
    # no colour image: (0, 255, 0)
    o2 = oracle_mod.Oracle(cfg)
    o2.set_frame(depth, np.eye(4))
    o2.alloc()
    o2.collect_visible()
    o2.integrate_tsdf()
    c2 = o2.dump_blocks()["color"][o2.dump_blocks()["weight"] > 0]
    assert ((c2 & 255) == 0).all() and (((c2 >> 8) & 255) == blend(0, 255)).all()


def test_alloc_covers_dense_samples(oracle_mod):
    """Brute force: every block that a dense sampling of [d - tr, d + tr]
    along each pixel ray touches (and whose centre is in the frustum) is
    allocated by the DDA."""
    cfg = plane_cfg()
    rng = np.random.default_rng(5)
    depth = rng.integers(600, 3000, (cfg.height, cfg.width)).astype(np.uint16)
    th = 0.3
    pose = np.eye(4)
    pose[:3, :3] = [[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]]
    pose[:3, 3] = [0.3, -0.2, 0.1]
    o = oracle_mod.Oracle(cfg)
    o.set_frame(depth, pose)
    o.alloc()
    got = {tuple(p) for p in o.dump_blocks()["pos"]}
    vs, tr = 0.02, 0.08
    P = pose.astype(np.float32).astype(np.float64)
    Rinv = np.linalg.inv(P)
    missing = 0
    for y in range(0, cfg.height, 3):
        for x in range(0, cfg.width, 3):
            dm = 0.001 * float(depth[y, x])
            ray = np.array([(x - cfg.cx) / cfg.fx, (y - cfg.cy) / cfg.fy, 1.0])
            for s in np.linspace(dm - tr + 1e-4, dm + tr - 1e-4, 64):
                pw = P[:3, :3] @ (ray * s) + P[:3, 3]
                b = tuple(int(np.floor(np.floor(c / vs + 0.5) / 8)) for c in pw)
                # centre-in-frustum filter, as the DDA applies it
                cen = (np.array(b) * 8 + 3.5) * vs
                pc = Rinv[:3, :3] @ cen + Rinv[:3, 3]
                u = pc[0] * cfg.fx / pc[2] + cfg.cx
                v = pc[1] * cfg.fy / pc[2] + cfg.cy
                nx = (2 * u - 63) / 63 * 0.95
                ny = (47 - 2 * v) / 47 * 0.95
                nz = (pc[2] - 0.1) / 4.9 * 0.95
                inside = -1 < nx < 1 and -1 < ny < 1 and 0 < nz < 1
                margin = min(abs(abs(nx) - 1), abs(abs(ny) - 1), abs(nz), abs(nz - 1))
                if inside and margin > 1e-3 and b not in got:
                    missing += 1
    assert missing == 0


def test_hash_invariants(oracle_mod):
    """core/hash_table.cu:146-262 checks, on a table small enough to force
    linked lists: no duplicate positions, each entry findable from its home
    bucket, list entries only where the home bucket is full."""
    cfg = plane_cfg(bucket_count=64, value_capacity=4096)
    rng = np.random.default_rng(7)
    depth = rng.integers(600, 3000, (cfg.height, cfg.width)).astype(np.uint16)
    o = oracle_mod.Oracle(cfg)
    o.set_frame(depth, np.eye(4))
    o.alloc()
    rec = o.dump_hash()
    pos = {tuple(r[3:]) for r in rec}
    assert len(pos) == len(rec)
    blocks = {tuple(p) for p in o.dump_blocks()["pos"]}
    assert blocks == pos
    home = rec[:, 1]
    for b in np.unique(home[rec[:, 2] == 0]):
        # a key in a list means its bucket had no free slot when it came
        assert (rec[:, 0] // 10 == b).sum() == 10
    for r in rec:
        assert r[1] == oracle_mod.hash_bucket(int(r[3]), int(r[4]), int(r[5]), cfg.bucket_count)


def test_collect_matches_frustum_filter(oracle_mod):
    """CollectBlocksInFrustum (collect_block_array.cu:25-151) keeps exactly
    the allocated blocks whose centre lies in the 0.95-shrunk frustum of the
    current pose; CollectAllBlocks keeps every allocated block.  The filter is
    recomputed here in FP64; blocks within 1e-4 of a frustum face are
    skipped (FP32 rounding decides them)."""
    cfg = plane_cfg()
    rng = np.random.default_rng(11)
    depth = rng.integers(600, 3000, (cfg.height, cfg.width)).astype(np.uint16)
    o = oracle_mod.Oracle(cfg)
    o.set_frame(depth, np.eye(4))
    o.alloc()
    th = 0.25
    pose = np.eye(4)
    pose[:3, :3] = [[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]]
    pose[:3, 3] = [0.1, 0.05, 0.2]
    o.set_frame(depth, pose)  # a second pose: some blocks leave the frustum
    n = o.collect_visible()
    got = {tuple(p) for p in o.dump_candidates()}
    assert len(got) == n
    allb = {tuple(p) for p in o.dump_blocks()["pos"]}
    cTw = np.linalg.inv(pose.astype(np.float32).astype(np.float64))
    vs = float(np.float32(0.02))
    expect, ambiguous = set(), set()
    for b in allb:
        cen = (np.array(b) * 8 + 3.5) * vs
        pc = cTw[:3, :3] @ cen + cTw[:3, 3]
        u = pc[0] * cfg.fx / pc[2] + cfg.cx
        v = pc[1] * cfg.fy / pc[2] + cfg.cy
        nx = (2 * u - (cfg.width - 1)) / (cfg.width - 1) * 0.95
        ny = ((cfg.height - 1) - 2 * v) / (cfg.height - 1) * 0.95
        nz = (pc[2] - 0.1) / 4.9 * 0.95
        margin = min(abs(abs(nx) - 1), abs(abs(ny) - 1), abs(nz), abs(nz - 1))
        if margin < 1e-4:
            ambiguous.add(b)
        elif -1 <= nx <= 1 and -1 <= ny <= 1 and 0 <= nz <= 1:
            expect.add(b)
    assert 0 < len(expect) < len(allb)
    assert got - ambiguous == expect
    assert o.collect_all() == len(allb)
    assert {tuple(p) for p in o.dump_candidates()} == allb


# ------------------------------------------------------------------ PSDF
def psdf_single(oracle_mod, depth_mm, n, pose=None):
    cfg = replace(plane_cfg(), psdf=True)
    o = oracle_mod.Oracle(cfg)
    depth = np.full((cfg.height, cfg.width), depth_mm, np.uint16)
    for _ in range(n):
        o.set_frame(depth, np.eye(4) if pose is None else pose)
        o.alloc()
        o.collect_visible()
        o.integrate_psdf()
    return cfg, o.dump_blocks()


def test_psdf_first_observation(oracle_mod):
    cfg, d = psdf_single(oracle_mod, 1500, 1)
    obs = d["weight"] > 0
    D = float(np.float32(0.001) * np.float32(1500))
    tau = (D - 0.4) * 0.012 + 0.019
    assert np.allclose(d["weight"][obs], 1 / tau ** 2, rtol=1e-12)
    assert np.allclose(d["a"][obs], 1.0, rtol=1e-12)   # 10 * rho, rho = 0.1
    assert np.allclose(d["b"][obs], 9.0, rtol=1e-12)


def test_psdf_kalman_and_beta_limits(oracle_mod):
    """Identical repeated observations: the Gaussian part is the product of
    Gaussians (precisions add, the mean stays), and a, b follow the
    moment-matched Beta update of the mixture weights."""
    cfg, d1 = psdf_single(oracle_mod, 1500, 1)
    cfg, d2 = psdf_single(oracle_mod, 1500, 2)
    obs = d1["weight"] > 0
    D = float(np.float32(0.001) * np.float32(1500))
    tau2 = ((D - 0.4) * 0.012 + 0.019) ** 2
    x = d1["sdf"][obs]
    # second update, mu == x: C1 = rho N(0; 2 tau^2), C2 = (1-rho)/4.9
    rho = 0.1
    g = 1.0 / (np.sqrt(2 * np.pi) * np.sqrt(2 * tau2))
    C1 = rho * g
    C2 = (1 - rho) / 4.9
    C1, C2 = C1 / (C1 + C2), C2 / (C1 + C2)
    assert np.allclose(d2["sdf"][obs], x, rtol=1e-12, atol=1e-15)
    s2 = tau2 / 2
    var = C1 * (s2 + x * x) + C2 * (tau2 + x * x) - x * x
    assert np.allclose(d2["weight"][obs], 1 / var, rtol=1e-9)
    a, b = 1.0, 9.0
    f = C1 * (a + 1) / (a + b + 1) + C2 * a / (a + b + 1)
    e = C1 * (a + 1) * (a + 2) / ((a + b + 1) * (a + b + 2)) + C2 * a * (a + 1) / ((a + b + 1) * (a + b + 2))
    # moment matching of a Beta: mean f, second moment e
    a_new = (e - f) / (f - e / f)
    assert np.allclose(d2["a"][obs], a_new, rtol=1e-9)
    assert np.all(d2["a"][obs] > 0) and np.all(d2["b"][obs] > 0)
    # the Beta with mean f and second moment e: check its moments directly
    bb = d2["b"][obs]
    aa = d2["a"][obs]
    assert np.allclose(aa / (aa + bb), f, rtol=1e-9)
    assert np.allclose(aa * (aa + 1) / ((aa + bb) * (aa + bb + 1)), e, rtol=1e-9)


# ------------------------------------------------------------------ raycast
def test_raycast_plane_closed_form(oracle_mod):
    """A fused fronto-parallel plane is a linear SDF inside the band, so
    trilinear interpolation and regula falsi are exact: the ray depth equals
    the plane depth up to FP32 rounding."""
    cfg = plane_cfg(width=80, height=60, fx=75.0, fy=75.0, cx=40.0, cy=30.0)
    D = 1.5
    depth = np.full((cfg.height, cfg.width), 1500, np.uint16)
    o = oracle_mod.Oracle(cfg)
    for _ in range(2):
        o.frame(depth, np.eye(4), mesh=False)
    rp = cfg.raycaster_params()
    d, vtx, n = o.raycast(np.eye(4), rp)
    core = (slice(10, -10), slice(10, -10))
    hit = np.isfinite(d[core])
    assert hit.all()
    assert np.abs(d[core][hit] - D).max() < 2e-5
    # the reference negates the SDF gradient (ray_caster.cu:118-119): for
    # sdf = D - z the normal is +z, pointing away from the camera
    assert np.abs(n[core][hit] - np.array([0, 0, 1])).max() < 1e-3


# ------------------------------------------------------------------ primal-dual
def test_pd_constant_field_is_fixed_point(oracle_mod):
    cfg = plane_cfg()
    o = oracle_mod.Oracle(cfg)
    pos = np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.int32)
    sdf = np.full((3, 512), 0.03, np.float32)
    w = np.full((3, 512), 50.0, np.float32)
    o.load_blocks(pos, sdf, w)
    o.collect_all()
    o.pd_init()
    for _ in range(10):
        e = o.pd_iterate(0.5, 0.25, 0.25)
        assert abs(e[0]) < 1e-20 and abs(e[1]) < 1e-20
    d = o.dump_blocks()
    assert np.abs(d["sdf"] - float(np.float32(0.03))).max() < 1e-15
    pd = o.dump_pd()
    assert np.abs(pd["p"]).max() == 0.0
    assert np.allclose(pd["w_pd"], 1.0)


def test_pd_pins_steep_voxels(oracle_mod):
    """w_pd = exp(|grad sdf| / vs) > 2 pins a voxel to its initial value
    (primal_dual.cu:147-148); a ramp of slope 1 per voxel (metric slope
    1/vs > ln 2) is pinned everywhere its gradient exists."""
    cfg = plane_cfg()
    o = oracle_mod.Oracle(cfg)
    pos = np.array([[0, 0, 5]], np.int32)
    x = np.arange(512) % 8
    sdf = (0.02 * x).astype(np.float32)
    w = np.full((1, 512), 50.0, np.float32)
    o.load_blocks(pos, sdf[None], w)
    o.collect_all()
    o.pd_init()
    for _ in range(5):
        o.pd_iterate(0.5, 0.25, 0.25)
    d = o.dump_blocks()
    pd = o.dump_pd()
    i = np.arange(512)
    inner = (x < 7) & ((i % 64) // 8 < 7) & (i // 64 < 7)
    assert np.all(pd["w_pd"][0][inner] > 2.0)
    assert np.allclose(pd["w_pd"][0][~inner], 1.0)
    assert np.array_equal(d["sdf"][0][inner], pd["sdf0"][0][inner])
