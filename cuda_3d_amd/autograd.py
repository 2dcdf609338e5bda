"""torch.autograd.Function over the colored-ICP objective adjoint (gca_color_icp_objective_grad).

E(xi, p, I) = sum_i lambda r_g,i^2 + (1 - lambda) r_c,i^2 with the unweighted residuals of the
reference's compute_RMSE_functor (src/registration/cuda_color_icp_build_least_square.cu:82-132),
evaluated at the pose Rz Ry Rx(xi) * T with the correspondences `nn` held fixed (they are
piecewise constant in the pose).  Backward returns dE/dxi (6) and dE/d(x, y, z, intensity) per
source point, both computed on the GPU by one adjoint kernel in the same pass as E.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .core import PointCloud, objective_grad


class ColoredICPObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, src_xyzi, target: PointCloud, nn, T, lam, work: PointCloud):
        if src_xyzi.dtype != torch.float32 or not src_xyzi.is_cuda or src_xyzi.shape[-1] != 4:
            raise ValueError("src_xyzi must be a float32 cuda tensor of shape (N, 4)")
        pts = src_xyzi.detach().contiguous()
        torch.cuda.synchronize()  # torch's writes before the library's stream reads
        L.check(L.lib.gca_cloud_upload_device(work.ctx.handle, work.handle, C.c_void_p(pts.data_ptr()),
                                              None, pts.shape[0]), "gca_cloud_upload_device")
        xi64 = np.asarray(xi.detach().cpu().numpy(), np.float64)
        E, dxi, dpts = objective_grad(work, target, nn, T, xi64, lam, want_points=True)
        ctx.save_for_backward(torch.as_tensor(dxi, dtype=torch.float64, device=xi.device),
                              dpts.clone())
        return torch.tensor(E, dtype=torch.float64, device=xi.device)

    @staticmethod
    def backward(ctx, gE):
        dxi, dpts = ctx.saved_tensors
        g_xi = gE * dxi
        g_pts = (gE.to(dpts.device).to(torch.float32) * dpts) if ctx.needs_input_grad[1] else None
        return g_xi, g_pts, None, None, None, None, None


def colored_icp_objective(xi, src_xyzi, target: PointCloud, nn, T=None, lam=0.968, work=None):
    """Differentiable objective.  xi: (6,) float64 tensor; src_xyzi: (N, 4) float32 cuda tensor
    (x, y, z, intensity); target: prepared PointCloud (normals + colour gradient); nn: int32
    correspondences (cuda tensor or array) into target, -1 for none; T: 4x4 base pose."""
    work = work or PointCloud(target.ctx, max(int(src_xyzi.shape[0]), 1))
    T = np.eye(4) if T is None else np.asarray(T, np.float64)
    return ColoredICPObjective.apply(xi, src_xyzi, target, nn, T, lam, work)
