"""The torch.autograd harness over the raycast adjoint (meshhashing_amd/autograd.py):
loss.backward() must hand torch the adjoint's gradients -- dL/dxi for a left
pose perturbation and dL/dsdf in the volume's own SDF layout -- and they must
equal the oracle's FP32-march adjoint of the same volume (1e-4 relative,
dL/dsdf 1e-4 of its largest entry).  The SDF tensor is a live view of the
volume.  A short step against the gradient lowers the loss."""
import numpy as np
import pytest

from meshhashing_amd.configs import C1
from meshhashing_amd.engine import Volume, synth_depth, synth_pose

pytestmark = pytest.mark.gpu


def test_autograd_harness_matches_the_adjoint_and_oracle(hip_device, oracle_mod):
    import torch
    from meshhashing_amd.autograd import raycast_depth_loss, sdf_tensor, se3_exp
    cfg = C1
    vol = Volume(cfg)
    for f in range(0, 8, 2):
        vol.upload_depth(synth_depth(cfg, f), synth_pose(cfg, f))
        vol.alloc()
        vol.collect_visible()
        vol.integrate_tsdf()
    vol.synchronize()
    sdf = sdf_tensor(vol, hip_device)
    # a live view: read through it, the volume's SDF
    assert np.array_equal(vol.dump_voxel_field(sdf.data_ptr()), vol.dump_blocks()["sdf"])
    sdf.requires_grad_(True)
    wTc0 = se3_exp([0.012, -0.008, 0.01, 0.015, -0.01, 0.008]) @ synth_pose(cfg, 3).astype(np.float64)
    obs_np = (synth_depth(cfg, 3) * np.float32(0.001)).astype(np.float32)
    obs = torch.from_numpy(obs_np).to(hip_device)
    xi = torch.zeros(6, dtype=torch.float64, requires_grad=True)
    loss = raycast_depth_loss(xi, sdf, vol, wTc0, obs)
    loss.backward()
    g = vol.dump_blocks()
    orc = oracle_mod.Oracle(cfg)
    orc.load_blocks(g["pos"], g["sdf"], g["weight"], g["color"])
    ref = orc.raycast_adjoint(False, np.linalg.inv(wTc0).astype(np.float32), obs_np, grad_sdf=True)
    gx = xi.grad.numpy()
    gsdf = vol.dump_voxel_field(sdf.grad.data_ptr())
    ge = np.abs(gx - ref["grad"]).max() / np.abs(ref["grad"]).max()
    se = np.abs(gsdf - ref["grad_sdf"]).max() / np.abs(ref["grad_sdf"]).max()
    print(f"[autograd] loss {loss.item():.6e} vs {ref['loss']:.6e}; grad {ge:.2e}, dL/dsdf {se:.2e}")
    assert abs(loss.item() - ref["loss"]) <= 1e-4 * ref["loss"]
    assert ge < 1e-4 and se < 1e-4
    # a step against the gradient (re-based pose) lowers the loss
    step = -1e-3 * gx / np.linalg.norm(gx)
    x2 = torch.zeros(6, dtype=torch.float64)
    l2 = raycast_depth_loss(x2, sdf.detach(), vol, se3_exp(step) @ wTc0, obs)
    assert l2.item() < loss.item(), (l2.item(), loss.item())
    # evaluation away from xi = 0 is refused (re-base the pose instead)
    with pytest.raises(ValueError):
        raycast_depth_loss(torch.full((6,), 1e-3, dtype=torch.float64), sdf.detach(), vol, wTc0, obs)
    vol.close()
