"""torch.autograd wrapper of the differentiable raycast (vsg_render_adjoint).

The reference's renderer is forward only (RayCastRender.cu:407-620); this module
exposes the library's adjoint to torch: the rendered depth is a function of a pose
perturbation xi (camera-frame, pose = T0 * exp(xi^), xi = (omega, upsilon)) and of
the scene's stored SDF values.  The forward and the backward both run on the GPU
through the C ABI; the depth is downloaded once per forward (the renderer owns its
output buffers).
"""
from __future__ import annotations

import numpy as np
import torch

from . import vsg


def se3_exp(xi) -> np.ndarray:
    """exp of the twist (omega, upsilon) as a 4x4 float64 matrix (Rodrigues' formula
    for the rotation, the left Jacobian of SO(3) for the translation)."""
    xi = np.asarray(xi, np.float64).reshape(6)
    w, u = xi[:3], xi[3:]
    th = float(np.linalg.norm(w))
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    if th < 1e-8:
        a, b, c = 1.0, 0.5, 1.0 / 6.0  # Taylor limits of sin t/t, (1-cos t)/t^2, (t-sin t)/t^3
    else:
        a, b, c = np.sin(th) / th, (1.0 - np.cos(th)) / th**2, (th - np.sin(th)) / th**3
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + a * K + b * (K @ K)
    T[:3, 3] = (np.eye(3) + b * K + c * (K @ K)) @ u
    return T


class _RaycastDepth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, sdf, render, scene, T0, K, min_depth, max_depth):
        T = np.asarray(T0, np.float64) @ se3_exp(xi.detach().cpu().numpy())
        fx, fy, cx, cy = K
        render.render(scene, T[:3, 3].astype(np.float32), T.astype(np.float32), fx, fy, cx, cy, min_depth, max_depth)
        depth = torch.from_numpy(render.downloadDepthFloat()).to(xi.device if xi.is_cuda else "cuda")
        ctx.args = (render, scene, T, K, min_depth, max_depth)
        ctx.sdf_shape = None if sdf is None else sdf.shape
        ctx.sdf_device = None if sdf is None else sdf.device
        ctx.xi_device = xi.device
        return depth

    @staticmethod
    def backward(ctx, gout):
        render, scene, T, K, min_depth, max_depth = ctx.args
        fx, fy, cx, cy = K
        # the adjoint differentiates the renderer's last render: render this pose again
        render.render(scene, T[:3, 3].astype(np.float32), T.astype(np.float32), fx, fy, cx, cy, min_depth, max_depth)
        n = render.getWidth() * render.getHeight()
        g = gout.detach().to(device="cuda", dtype=torch.float32).contiguous().view(n)
        want_xi, want_sdf = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        jac = torch.zeros(6 * n, dtype=torch.float32, device="cuda") if want_xi else None
        gsdf = None
        if want_sdf:
            gsdf = torch.zeros(int(np.prod(ctx.sdf_shape)), dtype=torch.float32, device="cuda")
        torch.cuda.current_stream().synchronize()  # g and the zeroed buffers are ready before the adjoint's stream reads them
        render.adjoint(scene, jac=jac, grad_depth=g if want_sdf else None, grad_sdf=gsdf)
        torch.cuda.synchronize()
        gxi = None
        if want_xi:
            gxi = (jac.view(n, 6).double() * g.double().view(n, 1)).sum(0).to(ctx.xi_device)
        if want_sdf:
            gsdf = gsdf.view(ctx.sdf_shape).to(ctx.sdf_device)
        return gxi, gsdf, None, None, None, None, None, None


def render_depth(render: "vsg.RayCastRender", scene: "vsg.VoxelScene", xi: torch.Tensor, T0, intrinsics,
                 min_depth: float, max_depth: float, sdf: torch.Tensor | None = None) -> torch.Tensor:
    """Rendered depth (float32 [W*H] on the GPU, 1e10 = no hit) at pose T0 * exp(xi^).

    xi: float64 tensor of 6 (omega, upsilon); gradients flow to it through d depth / d xi.
    sdf: optional tensor of sdf_block_num*512 elements standing for the scene's stored
    SDF (its values are not read); its gradient is grad_depth^T d depth / d sdf in the
    layout of VoxelScene.export_voxels (block pointer * 512 + voxel).  Pixels without a
    hit have zero derivatives; the sampled field's jumps at half-voxel planes are not
    differentiated.
    """
    return _RaycastDepth.apply(xi, sdf, render, scene, T0, tuple(float(k) for k in intrinsics), float(min_depth),
                               float(max_depth))
