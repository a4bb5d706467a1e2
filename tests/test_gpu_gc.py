"""Block recycling on the GPU (mh_collect_low_surfel_blocks,
mh_starve_blocks, mh_collect_garbage_blocks, mh_recycle_garbage_blocks)
against the oracle.

Bars: flags, freed counts, block sets and per-bucket key sets bit-exact;
SDF and weights bit-exact given the same inputs (starving is exact
arithmetic); the mesh (triangles, vertex keys and reference counts)
bit-exact.  The table's slot layout is compared per home bucket as a key
set only: FreeEntry makes a key's bucket-or-list placement depend on which
key of a bucket the (order-free) allocation put in the list.
"""
from dataclasses import replace

import numpy as np
import pytest

from meshhashing_amd.configs import C1
from meshhashing_amd.engine import MainEngine, Volume, synth_depth, synth_pose

pytestmark = pytest.mark.gpu


def small_cfg(**kw):
    return replace(C1, bucket_count=16, value_capacity=256, linked_list_size=150, **kw)


def bucket_keys(rec):
    out = {}
    for r in rec:
        out.setdefault(int(r[1]), set()).add(tuple(int(c) for c in r[3:]))
    return out


def flag_map(rec):
    return {tuple(int(c) for c in r[:3]): int(r[3]) for r in rec}


def test_low_surfel_and_garbage_layouts(hip_device, oracle_mod):
    """The oracle's hand-built layouts: a surface block, an empty block on
    top of it, an isolated block and an isolated pair over five low-surfel
    passes; then the garbage thresholds (|sdf| at the truncation, weights
    at 1e-6f)."""
    cfg = small_cfg()
    vs = cfg.voxel_size
    pos = np.array([(0, 0, 0), (0, 0, 1), (5, 5, 5), (-6, 4, 2), (-5, 4, 2)], np.int32)
    z = np.arange(512) // 64
    sdf = np.zeros((5, 512), np.float32)
    sdf[0] = ((z - 3.5) * vs).astype(np.float32)
    sdf[1] = ((z + 8 - 3.5) * vs).astype(np.float32)
    sdf[2:] = np.float32(3 * vs)
    w = np.ones((5, 512), np.float32)
    for passes, freed_want in ((3, 3), (5, 1)):
        vol, orc = Volume(cfg), oracle_mod.Oracle(cfg)
        vol.load_blocks(pos, sdf, w)
        orc.load_blocks(pos, sdf, w)
        for k in range(passes):
            for x in (vol, orc):
                x.collect_all()
                x.marching_cubes()
                x.collect_low_surfel_blocks()
            fg, fo = flag_map(vol.dump_candidate_flags()), flag_map(orc.dump_cand_flags())
            assert fg == fo, (passes, k)
        assert sum(fg.values()) == (3 if passes == 3 else 1)
        mg = vol.dump_mesh()
        assert len(mg["tri"]) == 98
        assert vol.recycle_garbage_blocks() == orc.recycle_garbage_blocks() == freed_want
        assert np.array_equal(vol.dump_blocks()["pos"], orc.dump_blocks()["pos"])
        assert bucket_keys(vol.dump_hash()) == bucket_keys(orc.dump_hash())
        mg2, mo = vol.dump_mesh(), orc.dump_mesh()
        for k in ("tri", "vkey", "vref"):
            assert np.array_equal(mg2[k], mo[k]) and np.array_equal(mg2[k], mg[k])
        assert vol.status()["block_count"] == 5 - freed_want
        vol.close()
    # garbage thresholds
    t = np.float32(cfg.truncation)
    gpos = np.array([(3 * i, 0, 0) for i in range(7)], np.int32)
    gs = np.zeros((7, 512), np.float32)
    gw = np.ones((7, 512), np.float32)
    gs[0] = t
    gs[1] = t
    gs[1, 17] = -np.nextafter(t, np.float32(0))
    gw[2] = np.float32(1e-6)
    gw[3] = 0.0
    gw[3, 5] = np.float32(2e-6)
    gs[4] = -t
    gw[4, :256] = 0.0
    gs[4, :256] = 0.0
    gw[6] = np.float32(1e-6)
    gw[6, 100] = 1.0
    gs[6, 100] = t
    vol, orc = Volume(cfg), oracle_mod.Oracle(cfg)
    for x in (vol, orc):
        x.load_blocks(gpos, gs, gw)
        x.collect_all()
        x.collect_garbage_blocks()
    fg = flag_map(vol.dump_candidate_flags())
    assert fg == flag_map(orc.dump_cand_flags())
    assert [fg[tuple(p)] for p in gpos.tolist()] == [1, 0, 1, 0, 1, 0, 1]
    assert vol.recycle_garbage_blocks() == orc.recycle_garbage_blocks() == 4
    assert np.array_equal(vol.dump_blocks()["pos"], orc.dump_blocks()["pos"])


def test_free_entry_random_sequence(hip_device, oracle_mod):
    """Inserts and frees crowding three buckets of sixteen (overflow lists
    in use) in an 80-block pool that 300 inserts only fit when freed slots
    are reused; GPU, oracle and a Python set agree after every step."""
    cfg = replace(small_cfg(), value_capacity=80)
    vol, orc = Volume(cfg), oracle_mod.Oracle(cfg)
    rng = np.random.default_rng(11)
    universe = [tuple(int(c) for c in p) for p in rng.integers(-20, 20, size=(4000, 3))]
    universe = [p for p in dict.fromkeys(universe) if oracle_mod.hash_bucket(*p, cfg.bucket_count) < 3]
    live = {}
    lists = 0
    for step in range(12):
        fresh = [p for p in universe if p not in live]
        add = [fresh[i] for i in rng.choice(len(fresh), size=25, replace=False)]
        for p in add:
            live[p] = np.float32(rng.uniform(-0.01, 0.01))
        a = np.array(add, np.int32)
        s = np.repeat(np.array([live[p] for p in add], np.float32)[:, None], 512, axis=1)
        for x in (vol, orc):
            x.load_blocks(a, s, np.ones_like(s))
        lists += int((vol.dump_hash()[:, 2] == 0).sum())
        keys = sorted(live)
        victims = np.array([keys[i] for i in rng.choice(len(keys), size=len(keys) // 3, replace=False)],
                           np.int32)
        z = np.zeros((len(victims), 512), np.float32)
        for x in (vol, orc):
            x.load_blocks(victims, z, z)
            x.collect_all()
            x.collect_garbage_blocks()
        fg, fo = vol.recycle_garbage_blocks(), orc.recycle_garbage_blocks()
        for p in victims.tolist():
            del live[tuple(p)]
        assert fg == fo == len(victims), step
        hg = vol.dump_hash()
        assert len({tuple(r[3:]) for r in hg}) == len(hg) == len(live)
        assert bucket_keys(hg) == bucket_keys(orc.dump_hash())
        d = vol.dump_blocks()
        assert {tuple(p) for p in d["pos"].tolist()} == set(live)
        for p, row in zip(d["pos"].tolist(), d["sdf"]):
            assert (row == live[tuple(p)]).all()
        st = vol.status()
        assert st["flags"] == 0 and st["block_count"] == len(live), st
    assert lists > 50
    vol.close()


def test_engine_schedule_matches_oracle(hip_device, oracle_mod):
    """MainEngine's schedule on the c1 sequence (low-surfel pass every
    frame, recycling at frame counts 10 and 20, starve + garbage at 14),
    with the oracle fed the GPU's fused SDF each frame: flags, freed
    counts, block sets, weights after starving, mesh and bucket key sets
    agree at every step."""
    cfg = C1
    eng = MainEngine(cfg)
    vol = eng.vol
    orc = oracle_mod.Oracle(cfg)
    freed_total = 0
    for f in range(cfg.frames):
        depth, pose = synth_depth(cfg, f), synth_pose(cfg, f)
        vol.upload_depth(depth, pose)
        eng.mapping()
        orc.set_frame(depth, pose)
        orc.alloc()
        orc.collect_visible()
        g = vol.dump_blocks()
        orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
        n = eng.integrated_frame_count
        vol.marching_cubes()
        orc.marching_cubes()
        vol.collect_low_surfel_blocks()
        orc.collect_low_surfel_blocks()
        assert flag_map(vol.dump_candidate_flags()) == flag_map(orc.dump_cand_flags()), f
        if n % 10 == 0:
            a, b = vol.recycle_garbage_blocks(), orc.recycle_garbage_blocks()
            print(f"[gc] frame {f}: low-surfel recycle freed {a}", flush=True)
            assert a == b, f
            freed_total += a
        if n % 15 == 14:
            vol.starve_blocks()
            orc.starve_blocks()
            vol.collect_garbage_blocks()
            orc.collect_garbage_blocks()
            fl = flag_map(vol.dump_candidate_flags())
            assert fl == flag_map(orc.dump_cand_flags()), f
            a, b = vol.recycle_garbage_blocks(), orc.recycle_garbage_blocks()
            print(f"[gc] frame {f}: garbage recycle flagged {sum(fl.values())} freed {a}", flush=True)
            assert a == b, f
            freed_total += a
            gw, ow = vol.dump_blocks()["weight"], orc.dump_blocks()["weight"]
            assert np.array_equal(gw.astype(np.float64), ow)
        if n % 10 == 0 or n % 15 == 14 or n == cfg.frames:
            gb, ob = vol.dump_blocks(), orc.dump_blocks()
            assert np.array_equal(gb["pos"], ob["pos"]), f
            mg, mo = vol.dump_mesh(), orc.dump_mesh()
            for k in ("tri", "vkey", "vref"):
                assert np.array_equal(mg[k], mo[k]), (f, k)
    assert bucket_keys(vol.dump_hash()) == bucket_keys(orc.dump_hash())
    st = vol.status()
    print(f"[gc] freed {freed_total} blocks over {cfg.frames} frames, {st['block_count']} live")
    assert st["flags"] == 0, st
    assert freed_total >= 10
    # MainEngine.frame runs the same schedule by itself
    eng2 = MainEngine(cfg)
    for f in range(cfg.frames):
        eng2.frame(synth_depth(cfg, f), synth_pose(cfg, f))
    assert np.array_equal(eng2.vol.dump_blocks()["pos"], vol.dump_blocks()["pos"])
    m2, m1 = eng2.vol.dump_mesh(), vol.dump_mesh()
    for k in ("tri", "vkey", "vref"):
        assert np.array_equal(m2[k], m1[k]), k
