"""The allocation-free raycast adjoint (mh_raycast_backward_noalloc) against
mh_raycast_backward through the C ABI: two calls in a row, the second on a
smaller image, must give bit-identical loss, dL/dxi, J^T J, pixel counts and
dL/dsdf (both run the same kernels in the same fixed summation order); an
image beyond the static scratch is refused."""
from dataclasses import replace

import numpy as np
import pytest

from meshhashing_amd import _capi as A
from meshhashing_amd.configs import C5
from meshhashing_amd.engine import Volume, synth_depth, synth_pose

pytestmark = pytest.mark.gpu


def test_noalloc_adjoint_equals_allocating_entry_point(hip_device):
    import torch
    cfg = replace(C5, frames=6)
    vol = Volume(cfg)
    for f in range(0, 6, 2):
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
    cTw = np.linalg.inv(synth_pose(cfg, 3).astype(np.float64)).astype(np.float32)
    big = cfg.raycaster_params()
    small = replace(cfg, width=320, height=240, fx=262.5, fy=262.5, cx=160.0, cy=120.0)
    small_p = small.raycaster_params()
    obs_big = torch.from_numpy((synth_depth(cfg, 3) * np.float32(0.001)).astype(np.float32)).to(hip_device)
    obs_small = torch.from_numpy((synth_depth(small, 3) * np.float32(0.001)).astype(np.float32)).to(hip_device)
    ncap = cfg.value_capacity * 512
    for params, obs in ((big, obs_big), (small_p, obs_small)):
        print(f"[noalloc] {params.width} x {params.height}", flush=True)
        ga = torch.zeros(ncap, dtype=torch.float32, device=hip_device)
        gb = torch.zeros(ncap, dtype=torch.float32, device=hip_device)
        a = vol.raycast_backward(cTw, obs.data_ptr(), ga.data_ptr(), params=params)
        b = vol.raycast_backward(cTw, obs.data_ptr(), gb.data_ptr(), params=params, noalloc=True)
        assert a["pixels"] > 10000 and a["pixels"] == b["pixels"] and a["outliers"] == b["outliers"]
        assert a["loss"] == b["loss"]
        assert np.array_equal(a["grad"], b["grad"]) and np.array_equal(a["hessian"], b["hessian"])
        # dL/dsdf is summed with float atomics: equal up to their order
        da, db = vol.dump_voxel_field(ga.data_ptr()), vol.dump_voxel_field(gb.data_ptr())
        assert np.abs(da - db).max() <= 1e-6 * np.abs(da).max()
    too_big = replace(cfg, width=1920, height=1080).raycaster_params()
    with pytest.raises(A.MHError, match="-3"):
        vol.raycast_backward(cTw, obs_big.data_ptr(), 0, params=too_big, noalloc=True)
    vol.close()
