"""torch.autograd binding of the path-replay adjoint (vkr_adjoint).

    theta = scenes.sphere_params(spheres, light_scale)   # 7 * n_spheres + 1
    radiance = render(theta, ctx, sid, workload)          # (H*W, 3) on cuda:0
    loss = ((radiance - target) ** 2).sum(); loss.backward()

Forward sets the scene's sphere parameters and runs the path kernel, which
also records where each sample's random stream starts (and its clamp bits);
backward runs the adjoint kernel, which replays every path from its RNG state
(the adjoint recomputes the records itself when the scene, its parameters or
the frame changed in between).  Branch
decisions (emitters, Russian roulette, dielectric choice) are held fixed, so
the gradient is that of the per-sample estimate.
"""
from __future__ import annotations

import numpy as np
import torch

import vkr


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, vctx, sid, frame, params):
        host = theta.detach().to("cpu", torch.float32).numpy()
        vctx.set_sphere_params(sid, host)
        n = params.width * params.height
        rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        fwd = params
        if fctx.needs_input_grad[0] and params.max_depth <= 10 and params.kernel == vkr.KERNEL_PATH and not (
                params.flags & (vkr.TRACE_RECORD_DECISIONS | vkr.TRACE_REPLAY_DECISIONS)):
            # Keep the adjoint's first pass: backward then runs only its second.
            fwd = vkr.TraceParams.from_buffer_copy(params)
            fwd.flags |= vkr.TRACE_RECORD_ADJOINT
        vctx.trace(sid, frame, fwd, radiance=rad)
        fctx.vctx, fctx.sid, fctx.frame, fctx.params = vctx, sid, frame, params
        fctx.theta_host = host
        fctx.theta_device = theta.device
        return rad[:, :3].clone()

    @staticmethod
    def backward(fctx, grad_rgb):
        vctx, sid = fctx.vctx, fctx.sid
        vctx.set_sphere_params(sid, fctx.theta_host)
        n = grad_rgb.shape[0]
        g4 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        g4[:, :3] = grad_rgb.to(torch.float32)
        d = torch.zeros(fctx.theta_host.size, dtype=torch.float32, device="cuda")
        vctx.adjoint(sid, fctx.frame, fctx.params, g4, d)
        return d.to(fctx.theta_device), None, None, None, None


def render(theta: torch.Tensor, vctx: "vkr.Context", sid: int, frame: np.ndarray, params) -> torch.Tensor:
    """Mean pre-tonemap radiance per pixel, differentiable in theta."""
    return _Render.apply(theta, vctx, sid, frame, params)
