"""A thin torch.autograd.Function over the raycast adjoint (mh_raycast_backward),
used as a test harness: it lets torch's autograd drive the hand-written
adjoint and checks that the gradients it hands back are the ones an
optimiser would use.  The computation is the library's; this module only
routes tensors.

RaycastDepthLoss(xi, sdf) = L = 1/2 sum_px (d_px - obs_px)^2 over the pixels
where the raycast of the volume hits and the observation is finite and
positive (visualization/ray_caster.cu:16-150, differentiated):

  * xi (float64, 6): a left perturbation of the base pose, wTc = exp(xi^) wTc0,
    evaluated at xi = 0 (the adjoint's tangent space); callers that step the
    pose re-base wTc0 <- exp(xi^) wTc0 and start again from xi = 0;
  * sdf: the volume's SDF storage viewed as a tensor (Volume.field_ptr), so
    dL/dsdf lands in sdf.grad with the volume's own layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as A


def sdf_tensor(vol, device="cuda:0") -> torch.Tensor:
    """The volume's SDF storage as a float32 tensor (a view, not a copy)."""
    ptr, n = vol.field_ptr(A.MH_FIELD_SDF)

    class _Iface:
        __cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (ptr, False),
                                    "version": 2}

    return torch.as_tensor(_Iface(), device=device)


def se3_exp(xi) -> np.ndarray:
    """exp of xi = (omega, upsilon) in se(3) as a 4x4 (float64)."""
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


class RaycastDepthLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, sdf, vol, wTc0, obs, params=None):
        x = xi.detach().cpu().double().numpy()
        if np.any(x != 0.0):
            raise ValueError("evaluate at xi = 0 and re-base the pose (the adjoint's tangent space)")
        wTc = wTc0.astype(np.float64)
        cTw = np.linalg.inv(wTc).astype(np.float32)
        gs = torch.zeros_like(sdf) if sdf.requires_grad else None
        out = vol.raycast_backward(cTw, obs.data_ptr(), 0 if gs is None else gs.data_ptr(), params=params)
        ctx.save_for_backward(torch.as_tensor(out["grad"], dtype=torch.float64),
                              gs if gs is not None else torch.empty(0))
        ctx.has_sdf = gs is not None
        ctx.hessian = out["hessian"]
        ctx.pixels = out["pixels"]
        return torch.tensor(out["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        grad_xi, gs = ctx.saved_tensors
        gx = g.cpu() * grad_xi
        gsdf = (g.to(gs.device).float() * gs) if ctx.has_sdf else None
        return gx, gsdf, None, None, None, None


def raycast_depth_loss(xi, sdf, vol, wTc0, obs, params=None):
    return RaycastDepthLoss.apply(xi, sdf, vol, wTc0, obs, params)
