"""ctypes binding of libgca.so (declarations in include/gca/gca.h).

The library is the only compute path: if libgca.so is missing or fails to load, importing
this module raises.  torch is imported first when available so that the process uses a
single HIP runtime (torch ships its own libamdhip64.so; libgca.so binds to the one already
loaded, whose soname matches).
"""
from __future__ import annotations

import ctypes as C
import os
import re

try:  # one HIP runtime per process: let torch load its copy first
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is part of the image
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgca.so")
HEADER = os.path.join(HERE, "..", "include", "gca", "gca.h")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libgca.so not built ({LIB_PATH}); run python cuda_3d_amd/build.py")
lib = C.CDLL(LIB_PATH)

GCA_OK, GCA_ERR_INVALID_VALUE, GCA_ERR_OUT_OF_MEMORY, GCA_ERR_DEVICE, GCA_ERR_NOT_SPD, \
    GCA_ERR_NOT_READY = range(6)
BILATERAL_NONE, BILATERAL_FIXED, BILATERAL_REFERENCE_ADAPTIVE = range(3)
ACCUM_COMPENSATED, ACCUM_NAIVE_FP32, ACCUM_FP64 = range(3)
RESIDUAL_PAIR, RESIDUAL_ABS, RESIDUAL_RMSE = range(3)
STAGES = ("frontend", "voxel", "grid", "normals", "gradient", "icp")
TUNE_KEYS = ("icp_lanes_small", "icp_lanes_medium", "icp_lanes_large", "icp_max_units", "normals_lanes",
             "icp_persist_blocks", "icp_launch", "icp_lanes_batch", "frontend", "grid_cell", "voxel")


class GcaError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        msg = lib.gca_status_string(status).decode()
        super().__init__(f"{where}: {msg} (status {status})")


class Intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float)]


class Extrinsics(C.Structure):
    _fields_ = [("rotation", C.c_float * 9), ("translation", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("depth_scale", C.c_float),
                ("depth_intrin", Intrinsics), ("color_intrin", Intrinsics),
                ("depth_to_color", Extrinsics)]


class Bilateral(C.Structure):
    _fields_ = [("mode", C.c_int32), ("radius", C.c_int32), ("sigma_space", C.c_float),
                ("sigma_depth", C.c_float)]


class IcpParams(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("max_correspondence_distance", C.c_float),
                ("lambda_geometric", C.c_float), ("accumulate_mode", C.c_int32),
                ("conv_rotation", C.c_double), ("conv_translation_sq", C.c_double)]


class IcpReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("solver_failures", C.c_int32), ("n_correspondences", C.c_int32),
                ("rmse", C.c_double), ("fitness", C.c_double), ("latency_us", C.c_double),
                ("active_us", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LevelParams(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("normal_radius", C.c_float),
                ("normal_max_nn", C.c_int32), ("gradient_radius", C.c_float),
                ("gradient_max_nn", C.c_int32), ("icp", IcpParams)]


vp, i32, i64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
P = C.POINTER
_SIGS = {
    "gca_status_string": (C.c_char_p, [i32]),
    "gca_abi_version": (i32, []),
    "gca_ctx_create": (i32, [i32, P(vp)]),
    "gca_ctx_destroy": (i32, [vp]),
    "gca_ctx_synchronize": (i32, [vp]),
    "gca_ctx_get_stream": (i32, [vp, P(vp)]),
    "gca_ctx_profile": (i32, [vp, i32]),
    "gca_ctx_set_tuning": (i32, [vp, i32, i32]),
    "gca_ctx_get_tuning": (i32, [vp, i32, P(i32)]),
    "gca_ctx_profile_read": (i32, [vp, vp, vp]),
    "gca_cloud_create": (i32, [vp, i64, P(vp)]),
    "gca_cloud_destroy": (i32, [vp]),
    "gca_cloud_size": (i32, [vp, P(i64)]),
    "gca_cloud_capacity": (i32, [vp, P(i64)]),
    "gca_cloud_has_normals": (i32, [vp, P(i32)]),
    "gca_cloud_upload": (i32, [vp, vp, vp, i64]),
    "gca_cloud_upload_device": (i32, [vp, vp, vp, vp, i64]),
    "gca_cloud_upload_normals": (i32, [vp, vp, i64]),
    "gca_cloud_upload_gradient": (i32, [vp, vp, i64]),
    "gca_cloud_download": (i32, [vp, vp, vp, vp, vp, vp]),
    "gca_cloud_device_arrays": (i32, [vp, P(vp), P(vp), P(vp), P(vp), P(vp)]),
    "gca_cloud_bounds": (i32, [vp, vp, vp]),
    "gca_cloud_copy": (i32, [vp, vp, vp]),
    "gca_cloud_append": (i32, [vp, vp, vp]),
    "gca_cloud_from_rgbd": (i32, [vp, vp, vp, P(Camera), P(Bilateral), f32, f32, vp]),
    "gca_voxel_downsample": (i32, [vp, vp, f32, vp]),
    "gca_radius_outlier_removal": (i32, [vp, vp, f32, i32, vp]),
    "gca_euclidean_clustering": (i32, [vp, vp, f32, i32, i32, vp, vp, vp]),
    "gca_euclidean_clustering_host": (i32, [vp, vp, f32, i32, i32, vp, i64, vp, vp]),
    "gca_residual_segmentation": (i32, [vp, vp, f32, vp, i32, i32, vp, vp, vp]),
    "gca_residual_segmentation_host": (i32, [vp, vp, f32, vp, i32, i32, vp, i64, vp, vp]),
    "gca_abs_residual_host": (i32, [vp, vp, vp, f32, f32, f32, vp, i64, P(i64), P(C.c_double)]),
    "gca_nn_search": (i32, [vp, vp, vp, f32, vp]),
    "gca_nn_search_host": (i32, [vp, vp, vp, f32, vp, i64, P(i64)]),
    "gca_radius_search": (i32, [vp, vp, vp, f32, i32, vp, vp, i64, P(i64)]),
    "gca_grid_cells": (i32, [vp, vp, f32, vp, vp, vp, vp]),
    "gca_estimate_normals": (i32, [vp, vp, f32, i32]),
    "gca_compute_color_gradient": (i32, [vp, vp, f32, i32]),
    "gca_transform": (i32, [vp, vp, vp]),
    "gca_build_gauss_newton": (i32, [vp, vp, vp, vp, f32, vp, i32, vp]),
    "gca_build_gauss_newton_noalloc": (i32, [vp, vp, vp, vp, f32, vp, i32, vp]),
    "gca_compute_residuals": (i32, [vp, vp, vp, vp, f32, vp, i32, vp, vp]),
    "gca_solve_gauss_newton": (i32, [vp, vp, vp, vp]),
    "gca_color_icp_align": (i32, [vp, vp, vp, P(IcpParams), f32, vp, vp, P(IcpReport)]),
    "gca_color_icp_objective_grad": (i32, [vp, vp, vp, vp, f32, vp, vp, vp, vp, vp]),
    "gca_color_icp_batch": (i32, [vp, i32, vp, vp, P(IcpParams), f32, vp, vp, vp]),
    "gca_color_icp_batch_multiscale": (i32, [vp, i32, i32, vp, vp, vp, vp, vp, vp, vp]),
    "gca_icp_batch_plan_create": (i32, [vp, i32, i32, vp, vp, vp, vp, vp, i32, P(vp)]),
    "gca_icp_batch_plan_enqueue": (i32, [vp]),
    "gca_icp_batch_plan_set_initial": (i32, [vp, vp]),
    "gca_icp_batch_plan_results": (i32, [vp, vp, vp]),
    "gca_icp_batch_plan_captures": (i32, [vp, P(i64)]),
    "gca_icp_batch_plan_destroy": (i32, [vp]),
    "gca_multiscale_icp": (i32, [vp, vp, vp, i32, P(LevelParams), vp, vp, vp]),
    "gca_pipeline_create": (i32, [vp, P(Camera), P(Bilateral), f32, f32, i32, vp, P(vp)]),
    "gca_pipeline_destroy": (i32, [vp]),
    "gca_pipeline_step": (i32, [vp, vp, vp]),
    "gca_pipeline_result": (i32, [vp, vp, vp, P(i32)]),
    "gca_pipeline_set_graph": (i32, [vp, i32]),
    "gca_pipeline_ctx": (i32, [vp, P(vp)]),
    "gca_pipeline_sizes": (i32, [vp, i32, P(i64), vp]),
    "gca_pipeline_graph_stats": (i32, [vp, P(i64), P(i64)]),
    "gca_pipeline_set_overlap": (i32, [vp, i32]),
    "gca_pipeline_flush": (i32, [vp]),
    "gca_pipeline_slots": (i32, [vp, P(i32)]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def declared_symbols() -> list[str]:
    """Entry points declared in include/gca/gca.h."""
    with open(HEADER) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:const char \*|int32_t )\s*(gca_[a-z0-9_]+)\(", text, re.M)))


def check(status: int, where: str) -> None:
    if status != GCA_OK:
        raise GcaError(status, where)
