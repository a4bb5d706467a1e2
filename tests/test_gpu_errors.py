"""Error paths and exhaustion on the GPU, through the C ABI.

* Block-pool exhaustion (hash_table.h:279-286): a key whose pool pop fails
  must not stay in the table, or it would count as allocated forever; once
  recycling has returned a pool slot the key is allocated like any other.
* The primal-dual neighbour table is built over the worklist at init
  (primal_dual.cu:164-223 looks neighbours up each call instead); iterate
  refuses to run on a worklist rebuilt since, rather than read stale rows.
* Loading blocks leaves the current worklist as it is.
"""
from dataclasses import replace

import numpy as np
import pytest

from meshhashing_amd import _capi as A
from meshhashing_amd.configs import C1, C3
from meshhashing_amd.engine import Volume

pytestmark = pytest.mark.gpu


def grid_blocks(n, z0=0):
    return np.array([(i % 8, (i // 8) % 8, z0 + i // 64) for i in range(n)], np.int32)


def test_exhausted_pool_leaves_no_key_behind(hip_device):
    cap = 64
    cfg = replace(C1, bucket_count=32, value_capacity=cap, linked_list_size=40)
    vol = Volume(cfg)
    pos = grid_blocks(cap)
    sdf = np.zeros((cap, 512), np.float32)
    w = np.ones((cap, 512), np.float32)
    w[0] = 0.0  # block 0: nothing observed, so the garbage test frees it
    vol.load_blocks(pos, sdf, w)
    st = vol.status()
    assert st["block_count"] == cap and st["flags"] == 0, st
    extra = np.array([(20, 20, 20)], np.int32)
    vol.load_blocks(extra, np.full((1, 512), 0.5, np.float32), np.full((1, 512), 2.0, np.float32))
    st = vol.status()
    assert st["flags"] & A.STATUS_BITS["BLOCK_HEAP_EXHAUSTED"], st
    assert st["block_count"] == cap
    rec = vol.dump_hash()
    assert len(rec) == cap and not any(tuple(r[3:]) == (20, 20, 20) for r in rec)
    # free block 0 through the reference's garbage path, then retry the key
    vol.collect_all()
    vol.collect_garbage_blocks()
    freed = vol.recycle_garbage_blocks()
    assert freed == 1, freed
    assert vol.status()["block_count"] == cap - 1
    vol.load_blocks(extra, np.full((1, 512), 0.5, np.float32), np.full((1, 512), 2.0, np.float32))
    d = vol.dump_blocks()
    assert len(d["pos"]) == cap
    k = np.nonzero((d["pos"] == extra[0]).all(axis=1))[0]
    assert len(k) == 1, "the key refused while the pool was full was never allocated"
    assert np.all(d["sdf"][k[0]] == np.float32(0.5)) and np.all(d["weight"][k[0]] == 2.0)
    assert not any(tuple(r) == (0, 0, 0) for r in d["pos"])
    vol.close()


def test_primal_dual_refuses_a_rebuilt_worklist(hip_device):
    cfg = replace(C3, value_capacity=1 << 12)
    vol = Volume(cfg)
    n = 40
    rng = np.random.default_rng(3)
    vol.load_blocks(grid_blocks(n), rng.normal(0, 0.02, (n, 512)).astype(np.float32),
                    np.full((n, 512), 3.0, np.float32))
    assert vol.collect_all(sync=True) == n
    vol.pd_init()
    vol.pd_iterate(0.5, 0.25, 0.25)
    vol.collect_all()
    with pytest.raises(A.MHError, match="-4"):
        vol.pd_iterate(0.5, 0.25, 0.25)
    vol.pd_init()
    vol.pd_iterate(0.5, 0.25, 0.25)
    vol.synchronize()
    vol.close()


def test_loading_blocks_keeps_the_worklist(hip_device):
    cfg = replace(C1, value_capacity=1 << 12)
    vol = Volume(cfg)
    n = 30
    vol.load_blocks(grid_blocks(n), np.zeros((n, 512), np.float32), np.ones((n, 512), np.float32))
    assert vol.collect_all(sync=True) == n
    before = vol.dump_candidates()
    vol.load_blocks(grid_blocks(5, z0=9), np.zeros((5, 512), np.float32), np.ones((5, 512), np.float32))
    assert vol.status()["candidate_count"] == n
    assert np.array_equal(vol.dump_candidates(), before)
    assert vol.status()["block_count"] == n + 5
    vol.close()
