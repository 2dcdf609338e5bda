"""cuda_3d_amd -- MI355X (gfx950) RGB-D front end and colored-ICP registration.

The compute path is libgca.so (hand-written HIP kernels behind the C ABI in
include/gca/gca.h).  This package is the Python mirror of the reference's host classes
(gca::point_cloud, gca::color_icp) over that ABI, plus a synthetic RGB-D scene source.
"""
from ._lib import (ACCUM_COMPENSATED, ACCUM_FP64, ACCUM_NAIVE_FP32, BILATERAL_FIXED,
                   BILATERAL_NONE, BILATERAL_REFERENCE_ADAPTIVE, RESIDUAL_ABS, RESIDUAL_PAIR,
                   RESIDUAL_RMSE, GcaError, IcpParams, IcpReport, LevelParams)
from .core import (BatchPlan, ColorICP, Context, PointCloud, batch_icp, batch_multiscale_icp, bilateral,
                   build_gauss_newton, compute_residuals, default_context, device_tensor, icp_params, level_params,
                   make_camera, multiscale_icp, objective_grad, solve_gauss_newton)

__all__ = [
    "ACCUM_COMPENSATED", "ACCUM_FP64", "ACCUM_NAIVE_FP32", "BILATERAL_FIXED", "BILATERAL_NONE",
    "BILATERAL_REFERENCE_ADAPTIVE", "RESIDUAL_ABS", "RESIDUAL_PAIR", "RESIDUAL_RMSE", "GcaError",
    "IcpParams", "IcpReport", "LevelParams", "BatchPlan", "ColorICP", "Context", "PointCloud", "batch_icp", "batch_multiscale_icp",
    "bilateral", "build_gauss_newton", "compute_residuals", "default_context", "device_tensor",
    "icp_params", "level_params", "make_camera", "multiscale_icp", "objective_grad",
    "solve_gauss_newton",
]
