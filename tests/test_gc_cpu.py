"""Block recycling in the oracle (mapping/recycle.cu, HashTable::FreeEntry),
pinned by what the algorithm fixes rather than by its own formulas:

  * FreeEntry on a table whose keys crowd three buckets of sixteen (every
    key beyond the tenth of a bucket sits in an overflow list): after any sequence of inserts and frees the
    table holds exactly the live key set (a Python set run beside it), each
    key once, with the surviving blocks' data intact, and freed pool slots
    are reused (300 inserts fit an 80-block pool);
  * the garbage test's two triggers at their thresholds: |sdf| equal to the
    truncation at 5 m (garbage) against one ulp below it (kept), and the
    observation threshold compared in double, where 1e-6f < 1e-6;
  * starving: weight - 1 floored at 0, for weights on both sides of 1;
  * the low-surfel life count on a hand-built layout: a surface block stays
    alive, an empty block touching it is never flagged, empty blocks out of
    reach are flagged on the third pass (life 3 -> 0), after which
    recycling frees exactly them and leaves the mesh as it was; one pass
    later an isolated pair keeps itself (a neighbour's life of -1 is not 0,
    recycle.cu:108-111) while a lone block is still flagged.
"""
from dataclasses import replace

import numpy as np
import pytest

from meshhashing_amd.configs import C1


def small_cfg(**kw):
    return replace(C1, bucket_count=16, value_capacity=256, linked_list_size=150, **kw)


def load(orc, pos, sdf=0.0, w=1.0):
    pos = np.asarray(pos, np.int32).reshape(-1, 3)
    n = len(pos)
    s = np.broadcast_to(np.asarray(sdf, np.float32), (n, 512)).copy()
    ww = np.broadcast_to(np.asarray(w, np.float32), (n, 512)).copy()
    orc.load_blocks(pos, s, ww)


def hash_keys(orc):
    return [tuple(int(c) for c in r[3:]) for r in orc.dump_hash()]


def gc_round(orc):
    orc.collect_all()
    orc.collect_garbage_blocks()
    return orc.recycle_garbage_blocks()


def test_free_entry_keeps_the_key_set(oracle_mod):
    # 300 inserts into an 80-block pool (at most 75 live): only works when
    # freed pool slots are handed out again
    cfg = replace(small_cfg(), value_capacity=80)
    orc = oracle_mod.Oracle(cfg)
    rng = np.random.default_rng(11)
    universe = [tuple(int(c) for c in p) for p in rng.integers(-20, 20, size=(4000, 3))]
    universe = [p for p in dict.fromkeys(universe) if oracle_mod.hash_bucket(*p, cfg.bucket_count) < 3]
    assert len(universe) > 300
    live = {}
    lists_seen = 0
    for step in range(12):
        # insert a batch of new keys (marked by their sdf value)
        fresh = [p for p in universe if p not in live]
        add = [fresh[i] for i in rng.choice(len(fresh), size=min(len(fresh), 25), replace=False)]
        for p in add:
            live[p] = float(rng.uniform(-0.01, 0.01))
        load(orc, add, sdf=np.array([live[p] for p in add], np.float32)[:, None])
        rec = orc.dump_hash()
        lists_seen += int((rec[:, 2] == 0).sum())
        # free a random subset: zero weights make them garbage
        keys = sorted(live)
        victims = [keys[i] for i in rng.choice(len(keys), size=len(keys) // 3, replace=False)]
        load(orc, victims, sdf=0.0, w=0.0)
        freed = gc_round(orc)
        for p in victims:
            del live[p]
        assert freed == len(victims), step
        ks = hash_keys(orc)
        assert len(ks) == len(set(ks)), "duplicate key in the table"
        assert set(ks) == set(live), step
        d = orc.dump_blocks()
        got = {tuple(int(c) for c in p): s[0] for p, s in zip(d["pos"], d["sdf"])}
        assert got.keys() == live.keys()
        for p, s in live.items():
            assert got[p] == np.float32(s)
        # a second pass finds nothing new to free
        assert gc_round(orc) == 0
    assert lists_seen > 50, "overflow lists were not exercised"
    assert orc.L.orc_block_count(orc.h) == len(live)


def test_garbage_thresholds(oracle_mod):
    cfg = small_cfg()
    t = np.float32(cfg.truncation)  # truncation at 5 m = scale * 5 + base, scale 0
    below = np.nextafter(t, np.float32(0))
    orc = oracle_mod.Oracle(cfg)
    pos = [(0, 0, 0), (3, 0, 0), (6, 0, 0), (9, 0, 0), (12, 0, 0), (15, 0, 0), (18, 0, 0)]
    sdf = np.zeros((7, 512), np.float32)
    w = np.ones((7, 512), np.float32)
    sdf[0] = t            # every observed |sdf| at the truncation: garbage
    sdf[1] = t
    sdf[1, 17] = -below   # one voxel just inside: kept
    sdf[2] = 0.0
    w[2] = np.float32(1e-6)  # 1e-6f < 1e-6 in double: nothing observed, garbage
    sdf[3] = 0.0
    w[3] = 0.0
    w[3, 5] = np.float32(2e-6)  # one observed voxel near the surface: kept
    sdf[4] = -t
    w[4, :256] = 0.0      # unobserved voxels are ignored; the rest at -t: garbage
    sdf[4, :256] = 0.0
    sdf[5] = 0.0          # ordinary surface block: kept
    w[6] = np.float32(1e-6)  # voxels at 1e-6f are unobserved, so only the
    w[6, 100] = 1.0          # one observed voxel, at the truncation, counts:
    sdf[6, 100] = t          # garbage
    orc.load_blocks(np.array(pos, np.int32), sdf, w)
    orc.collect_all()
    orc.collect_garbage_blocks()
    flags = {tuple(r[:3]): int(r[3]) for r in orc.dump_cand_flags()}
    assert flags == {pos[0]: 1, pos[1]: 0, pos[2]: 1, pos[3]: 0, pos[4]: 1, pos[5]: 0, pos[6]: 1}
    assert orc.recycle_garbage_blocks() == 4
    assert sorted(map(tuple, orc.dump_blocks()["pos"].tolist())) == [pos[1], pos[3], pos[5]]


def test_starve_floors_at_zero(oracle_mod):
    cfg = small_cfg()
    orc = oracle_mod.Oracle(cfg)
    w = np.array([0.0, 0.25, 1.0, 1.5, 7.0, 255.0], np.float32)
    ww = np.resize(w, 512).reshape(1, 512)
    orc.load_blocks(np.zeros((1, 3), np.int32), np.zeros((1, 512), np.float32), ww)
    orc.collect_all()
    orc.starve_blocks()
    got = orc.dump_blocks()["weight"][0]
    assert np.array_equal(got, np.maximum(ww[0].astype(np.float64) - 1.0, 0.0))
    orc.starve_blocks()
    got = orc.dump_blocks()["weight"][0]
    assert np.array_equal(got, np.maximum(ww[0].astype(np.float64) - 2.0, 0.0))


def surfel_layout(oracle_mod):
    cfg = small_cfg()
    vs = cfg.voxel_size
    orc = oracle_mod.Oracle(cfg)
    # A: a horizontal plane at z = 3.5 voxels crosses it (64 z-edge vertices
    # inside the block, 36 of them inner).  B sits on top of A, wholly
    # outside the surface.  C is isolated; D and E are an isolated pair.
    blocks = dict(A=(0, 0, 0), B=(0, 0, 1), C=(5, 5, 5), D=(-6, 4, 2), E=(-5, 4, 2))
    pos = np.array(list(blocks.values()), np.int32)
    z = np.arange(512) // 64
    sdf = np.zeros((5, 512), np.float32)
    sdf[0] = ((z - 3.5) * vs).astype(np.float32)
    sdf[1] = ((z + 8 - 3.5) * vs).astype(np.float32)
    sdf[2:] = np.float32(3 * vs)
    orc.load_blocks(pos, sdf, np.ones((5, 512), np.float32))
    return orc, blocks


def surfel_pass(orc, blocks):
    orc.collect_all()
    orc.marching_cubes()
    orc.collect_low_surfel_blocks()
    name = {v: k for k, v in blocks.items()}
    return "".join(sorted(name[tuple(int(c) for c in r[:3])] for r in orc.dump_cand_flags() if r[3]))


def live_blocks(orc, blocks):
    name = {v: k for k, v in blocks.items()}
    return "".join(sorted(name[tuple(int(c) for c in p)] for p in orc.dump_blocks()["pos"]))


def test_low_surfel_life_and_recycle(oracle_mod):
    orc, blocks = surfel_layout(oracle_mod)
    assert [surfel_pass(orc, blocks) for _ in range(3)] == ["", "", "CDE"]
    mesh0 = orc.dump_mesh()
    assert len(mesh0["tri"]) == 7 * 7 * 2  # the plane in A: 7 x 7 cubes, 2 triangles each
    assert len(mesh0["vkey"]) == 64
    assert orc.recycle_garbage_blocks() == 3
    assert live_blocks(orc, blocks) == "AB"
    assert sorted(hash_keys(orc)) == sorted([blocks["A"], blocks["B"]])
    mesh1 = orc.dump_mesh()
    for k in ("tri", "vkey", "vref"):
        assert np.array_equal(mesh0[k], mesh1[k])


def test_low_surfel_pair_outlives_its_flag(oracle_mod):
    orc, blocks = surfel_layout(oracle_mod)
    assert [surfel_pass(orc, blocks) for _ in range(5)] == ["", "", "CDE", "C", "C"]
    assert orc.recycle_garbage_blocks() == 1
    assert live_blocks(orc, blocks) == "ABDE"
