"""Python mirror of the reference's host API over libgca.so.

``PointCloud`` follows ``gca::point_cloud`` (src/geometry/point_cloud.hpp:22-129) and
``ColorICP`` follows ``gca::color_icp`` (src/registration/color_icp.hpp:16-63): same method
names, same argument meaning, and the reference's failure conventions (invalid arguments ->
``None`` / ``False`` where the reference returns nullptr / false; device errors raise).
Everything computes in libgca.so on the GPU; there is no CPU path.

Device memory for frames and index buffers is torch tensors on ``cuda``: torch only
allocates them.  The library runs on its own HIP stream, so the wrappers synchronise at the
torch <-> library hand-over points.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from ._lib import check, lib

try:
    import torch
except Exception:  # pragma: no cover
    torch = None


def _torch():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("a HIP device (torch.cuda) is required")
    return torch


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, cols)


def _T(T):
    return np.ascontiguousarray(np.asarray(T, dtype=np.float64).reshape(16))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """A device + HIP stream (reference: the implicit CUDA context, cuda_warm_up_gpu)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib.gca_ctx_create(device, C.byref(h)), "gca_ctx_create")
        self.handle = h
        self.device = device

    def synchronize(self):
        check(lib.gca_ctx_synchronize(self.handle), "gca_ctx_synchronize")

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        check(lib.gca_ctx_get_stream(self.handle, C.byref(s)), "gca_ctx_get_stream")
        return s.value or 0

    def set_tuning(self, **kw):
        """Launch-shape knobs (names in _lib.TUNE_KEYS), 0 = built-in heuristic."""
        for k, v in kw.items():
            check(lib.gca_ctx_set_tuning(self.handle, L.TUNE_KEYS.index(k), int(v)), f"set_tuning {k}")

    def get_tuning(self) -> dict:
        out = {}
        for i, k in enumerate(L.TUNE_KEYS):
            v = C.c_int32()
            check(lib.gca_ctx_get_tuning(self.handle, i, C.byref(v)), "get_tuning")
            out[k] = v.value
        return out

    def profile(self, enable: bool = True):
        """Bracket every stage launch with hipEvents on the ctx stream (clears old records)."""
        check(lib.gca_ctx_profile(self.handle, 1 if enable else 0), "gca_ctx_profile")

    def profile_read(self) -> dict:
        """{stage: (launches, total_ms)} since profiling was enabled (synchronises)."""
        n = np.zeros(len(L.STAGES), np.int64)
        ms = np.zeros(len(L.STAGES), np.float64)
        check(lib.gca_ctx_profile_read(self.handle, _ptr(n), _ptr(ms)), "gca_ctx_profile_read")
        return {s: (int(n[i]), float(ms[i])) for i, s in enumerate(L.STAGES)}

    @classmethod
    def borrow(cls, handle, device: int = 0) -> "Context":
        """A non-owning view of a ctx owned elsewhere (e.g. a frame pipeline's own ctx)."""
        c = cls.__new__(cls)
        c.handle = handle
        c.device = device
        c._borrowed = True
        return c

    def close(self):
        if getattr(self, "handle", None) and not getattr(self, "_borrowed", False):
            lib.gca_ctx_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def make_camera(width, height, fx, fy, cx, cy, depth_scale=0.001, color_intrin=None,
                rotation=None, translation=None) -> L.Camera:
    """gca::cuda_camera_param equivalent (cuda_camera_param_container.cu:35-63)."""
    cam = L.Camera()
    cam.width, cam.height, cam.depth_scale = width, height, depth_scale
    cam.depth_intrin = L.Intrinsics(fx, fy, cx, cy)
    cam.color_intrin = L.Intrinsics(*(color_intrin if color_intrin is not None else (fx, fy, cx, cy)))
    R = rotation if rotation is not None else (1, 0, 0, 0, 1, 0, 0, 0, 1)
    t = translation if translation is not None else (0, 0, 0)
    cam.depth_to_color = L.Extrinsics((C.c_float * 9)(*R), (C.c_float * 3)(*t))
    return cam


def bilateral(mode=L.BILATERAL_FIXED, radius=2, sigma_space=2.0, sigma_depth=30.0) -> L.Bilateral:
    return L.Bilateral(mode, radius, sigma_space, sigma_depth)


def device_tensor(a, dtype=None):
    """Host array -> torch cuda tensor, synchronised so the library sees the data."""
    t = _torch()
    x = t.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()
    t.cuda.synchronize()
    return x


class PointCloud:
    """Device point cloud (SoA float4 arrays owned by libgca)."""

    def __init__(self, ctx: Optional[Context] = None, capacity: int = 1, _handle=None):
        self.ctx = ctx or default_context()
        if _handle is None:
            h = C.c_void_p()
            check(lib.gca_cloud_create(self.ctx.handle, max(int(capacity), 1), C.byref(h)),
                  "gca_cloud_create")
            self.handle = h
        else:
            self.handle = _handle

    def __del__(self):
        try:
            if getattr(self, "handle", None) and self.ctx.handle:
                lib.gca_cloud_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    # ---- construction ------------------------------------------------------------------------
    @staticmethod
    def from_numpy(xyz, rgb, ctx: Optional[Context] = None) -> "PointCloud":
        """point_cloud::create_from_pcl (point_cloud.cu:481-504): rgb in [0,1]."""
        xyz, rgb = _f32(xyz, 3), _f32(rgb, 3)
        pc = PointCloud(ctx, len(xyz))
        check(lib.gca_cloud_upload(pc.handle, _ptr(xyz), _ptr(rgb), len(xyz)), "gca_cloud_upload")
        return pc

    @staticmethod
    def create_from_ply(path: str, ctx: Optional[Context] = None, with_normals: bool = False):
        """loadPointCloud + removeNaNPoints + create_from_pcl (test_color_icp.cpp:298-343,
        point_cloud.cu:481-504).  The reference goes through pcl::PointXYZRGB and so drops the
        file's normals; ``with_normals`` uploads them instead."""
        from .ply import read_ply

        v = read_ply(path)
        pc = PointCloud.from_numpy(v.xyz, v.rgb(), ctx)
        if with_normals and v.normals is not None and len(v.xyz):
            pc.set_normals(v.normals)
        return pc

    @staticmethod
    def create_from_rgbd(depth, bgr, camera: L.Camera, filter: Optional[L.Bilateral] = None,
                         threshold_min_in_meter: float = 0.0, threshold_max_in_meter: float = 10.0,
                         ctx: Optional[Context] = None, out: Optional["PointCloud"] = None):
        """point_cloud::create_from_rgbd (point_cloud.cu:455-479).  depth/bgr: torch cuda
        tensors (uint16 HxW, uint8 HxWx3) or host arrays (uploaded first)."""
        t = _torch()
        if not (isinstance(depth, t.Tensor) and depth.is_cuda):
            depth = device_tensor(np.asarray(depth, np.uint16).view(np.int16)).view(t.int16)
        if not (isinstance(bgr, t.Tensor) and bgr.is_cuda):
            bgr = device_tensor(np.asarray(bgr, np.uint8))
        pc = out or PointCloud(ctx, camera.width * camera.height)
        f = filter if filter is not None else L.Bilateral(L.BILATERAL_NONE, 0, 1.0, 1.0)
        st = lib.gca_cloud_from_rgbd(pc.ctx.handle, C.c_void_p(depth.data_ptr()),
                                     C.c_void_p(bgr.data_ptr()), C.byref(camera), C.byref(f),
                                     threshold_min_in_meter, threshold_max_in_meter, pc.handle)
        check(st, "gca_cloud_from_rgbd")
        return pc

    # ---- properties -----------------------------------------------------------------------------
    def points_number(self) -> int:
        n = C.c_int64()
        check(lib.gca_cloud_size(self.handle, C.byref(n)), "gca_cloud_size")
        return n.value

    def __len__(self):
        return self.points_number()

    def has_normals(self) -> bool:
        f = C.c_int32()
        check(lib.gca_cloud_has_normals(self.handle, C.byref(f)), "gca_cloud_has_normals")
        return bool(f.value)

    def _bounds(self):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        check(lib.gca_cloud_bounds(self.handle, _ptr(mn), _ptr(mx)), "gca_cloud_bounds")
        return mn, mx

    def get_min_bound(self):
        return self._bounds()[0]

    def get_max_bound(self):
        return self._bounds()[1]

    def device_arrays(self):
        ptrs = [C.c_void_p() for _ in range(5)]
        check(lib.gca_cloud_device_arrays(self.handle, *[C.byref(p) for p in ptrs]),
              "gca_cloud_device_arrays")
        return [p.value for p in ptrs]

    # ---- download -----------------------------------------------------------------------------
    def download(self):
        """point_cloud::download: (xyz [n,3], rgb [n,3], intensity [n])."""
        n = self.points_number()
        xyz = np.zeros((n, 3), np.float32)
        rgb = np.zeros((n, 3), np.float32)
        inten = np.zeros(n, np.float32)
        check(lib.gca_cloud_download(self.handle, _ptr(xyz), _ptr(rgb), _ptr(inten), None, None),
              "gca_cloud_download")
        return xyz, rgb, inten

    def download_xyzi(self):
        xyz, _, inten = self.download()
        return np.concatenate([xyz, inten[:, None]], 1)

    def download_normals(self):
        n = self.points_number()
        out = np.zeros((n, 3), np.float32)
        check(lib.gca_cloud_download(self.handle, None, None, None, _ptr(out), None),
              "gca_cloud_download")
        return out

    def download_gradient(self):
        n = self.points_number()
        out = np.zeros((n, 3), np.float32)
        check(lib.gca_cloud_download(self.handle, None, None, None, None, _ptr(out)),
              "gca_cloud_download")
        return out

    def set_normals(self, normals):
        nrm = _f32(normals, 3)
        check(lib.gca_cloud_upload_normals(self.handle, _ptr(nrm), len(nrm)), "upload_normals")

    def set_gradient(self, grad):
        g = _f32(grad, 3)
        check(lib.gca_cloud_upload_gradient(self.handle, _ptr(g), len(g)), "upload_gradient")

    # ---- operations ---------------------------------------------------------------------------
    def copy(self) -> "PointCloud":
        out = PointCloud(self.ctx, 1)
        check(lib.gca_cloud_copy(self.ctx.handle, self.handle, out.handle), "gca_cloud_copy")
        return out

    def __iadd__(self, other: "PointCloud") -> "PointCloud":
        """point_cloud::operator+= (point_cloud.cu:55-73): other's points appended; normals kept
        only when both clouds have them; bounds recomputed."""
        check(lib.gca_cloud_append(self.ctx.handle, self.handle, other.handle), "gca_cloud_append")
        return self

    def voxel_grid_down_sample(self, voxel_size: float, out: Optional["PointCloud"] = None):
        """point_cloud::voxel_grid_down_sample (point_cloud.cu:198-229); None on bad size."""
        pc = out or PointCloud(self.ctx, 1)
        st = lib.gca_voxel_downsample(self.ctx.handle, self.handle, voxel_size, pc.handle)
        if st == L.GCA_ERR_INVALID_VALUE:
            return None
        check(st, "gca_voxel_downsample")
        return pc

    def radius_outlier_removal(self, radius: float, min_neighbors_in_radius: int):
        """point_cloud::radius_outlier_removal (point_cloud.cu:238-282)."""
        pc = PointCloud(self.ctx, 1)
        st = lib.gca_radius_outlier_removal(self.ctx.handle, self.handle, radius,
                                            min_neighbors_in_radius, pc.handle)
        if st == L.GCA_ERR_INVALID_VALUE:
            return pc  # reference: empty cloud on a non-positive radius
        check(st, "gca_radius_outlier_removal")
        return pc

    def euclidean_cluster_labels(self, cluster_tolerance: float, min_cluster_size: int,
                                 max_cluster_size: int):
        """Per-point cluster numbers (device int32 tensor, -1 = dropped) from
        gca_euclidean_clustering, with the cluster count and the number of components cut at
        max_cluster_size.  None when the reference returns nothing (no normals, bad sizes)."""
        t = _torch()
        n = self.points_number()
        labels = t.empty(max(n, 1), dtype=t.int32, device="cuda")
        nc, ncut = C.c_int32(), C.c_int32()
        t.cuda.synchronize()
        st = lib.gca_euclidean_clustering(self.ctx.handle, self.handle, cluster_tolerance, min_cluster_size,
                                          max_cluster_size, C.c_void_p(labels.data_ptr()), C.byref(nc),
                                          C.byref(ncut))
        if st in (L.GCA_ERR_INVALID_VALUE, L.GCA_ERR_NOT_READY):
            return None
        check(st, "gca_euclidean_clustering")
        return labels[:n], nc.value, ncut.value

    def residual_segment_labels(self, cluster_tolerance: float, abs_residual, min_cluster_size: int,
                                max_cluster_size: int):
        """gca_residual_segmentation: per-point segment numbers (device int32 tensor, -1 =
        dropped), the segment count and the number of components cut at max_cluster_size.
        abs_residual: one float per point (numpy or a cuda tensor).  None where the reference
        returns nothing (bad tolerance or sizes)."""
        t = _torch()
        n = self.points_number()
        res = abs_residual if isinstance(abs_residual, t.Tensor) else device_tensor(np.asarray(abs_residual, np.float32))
        if res.numel() < n:
            raise ValueError("abs_residual needs one value per point")
        res = res.to(device="cuda", dtype=t.float32).contiguous()
        labels = t.empty(max(n, 1), dtype=t.int32, device="cuda")
        nc, ncut = C.c_int32(), C.c_int32()
        t.cuda.synchronize()
        st = lib.gca_residual_segmentation(self.ctx.handle, self.handle, cluster_tolerance, C.c_void_p(res.data_ptr()),
                                           min_cluster_size, max_cluster_size, C.c_void_p(labels.data_ptr()),
                                           C.byref(nc), C.byref(ncut))
        if st == L.GCA_ERR_INVALID_VALUE:
            return None
        check(st, "gca_residual_segmentation")
        return labels[:n], nc.value, ncut.value

    @staticmethod
    def _clusters_from_labels(r):
        if r is None or r[1] == 0:
            return []
        lab = r[0].cpu().numpy()
        idx = np.argsort(lab, kind="stable").astype(np.int32)
        lab_sorted = lab[idx]
        cuts = np.searchsorted(lab_sorted, np.arange(r[1] + 1))
        return [idx[cuts[k]:cuts[k + 1]] for k in range(r[1])]

    def euclidean_clustering(self, cluster_tolerance: float, min_cluster_size: int, max_cluster_size: int):
        """point_cloud::euclidean_clustering (point_cloud.cu:284-335): a list of clusters in seed
        order, each an ascending int32 array of point indices (the reference lists a cluster in
        its search-queue order); [] where the reference returns an empty result."""
        return self._clusters_from_labels(
            self.euclidean_cluster_labels(cluster_tolerance, min_cluster_size, max_cluster_size))

    def icp_residual_segmentation(self, cluster_tolerance: float, abs_residual, min_cluster_size: int,
                                  max_cluster_size: int):
        """point_cloud::icp_residual_segmentation (point_cloud.cu:391-435): segments in seed order,
        each an ascending int32 array of point indices; abs_residual as from
        ColorICP.get_abs_residual; [] where the reference returns an empty result."""
        return self._clusters_from_labels(
            self.residual_segment_labels(cluster_tolerance, abs_residual, min_cluster_size, max_cluster_size))

    def estimate_normals(self, search_radius: float, max_nn: int = 0) -> bool:
        """point_cloud::estimate_normals (point_cloud.cu:145-181)."""
        st = lib.gca_estimate_normals(self.ctx.handle, self.handle, search_radius, max_nn)
        if st == L.GCA_ERR_INVALID_VALUE:
            return False
        check(st, "gca_estimate_normals")
        return True

    def compute_color_gradient(self, search_radius: float, max_nn: int = 0) -> bool:
        """color_icp::compute_color_gradient_target (color_icp.cu:50-87)."""
        st = lib.gca_compute_color_gradient(self.ctx.handle, self.handle, search_radius, max_nn)
        if st in (L.GCA_ERR_INVALID_VALUE, L.GCA_ERR_NOT_READY):
            return False
        check(st, "gca_compute_color_gradient")
        return True

    def transform(self, T):
        """point_cloud::transform (point_cloud.cu:93-105)."""
        check(lib.gca_transform(self.ctx.handle, self.handle, _ptr(_T(T))), "gca_transform")

    def grid_cells(self, radius: float):
        t = _torch()
        n = self.points_number()
        cells = t.empty((max(n, 1), 3), dtype=t.int32, device="cuda")
        origin = np.zeros(3, np.float32)
        cell = C.c_float()
        dims = np.zeros(3, np.int32)
        check(lib.gca_grid_cells(self.ctx.handle, self.handle, radius,
                                 C.c_void_p(cells.data_ptr()), _ptr(origin), C.byref(cell),
                                 _ptr(dims)), "gca_grid_cells")
        self.ctx.synchronize()
        return cells[:n].cpu().numpy(), origin, cell.value, dims

    @staticmethod
    def nn_search(query_pc: "PointCloud", reference_pc: "PointCloud", radius: float,
                  as_tensor: bool = False):
        """point_cloud::nn_search (point_cloud.cu:506-557): index in reference or -1."""
        t = _torch()
        n = query_pc.points_number()
        out = t.empty(max(n, 1), dtype=t.int32, device="cuda")
        st = lib.gca_nn_search(query_pc.ctx.handle, query_pc.handle, reference_pc.handle, radius,
                               C.c_void_p(out.data_ptr()))
        if st == L.GCA_ERR_INVALID_VALUE:
            return np.zeros(0, np.int32)
        check(st, "gca_nn_search")
        query_pc.ctx.synchronize()
        return out[:n] if as_tensor else out[:n].cpu().numpy()

    @staticmethod
    def radius_search(query_pc: "PointCloud", reference_pc: "PointCloud", radius: float,
                      max_nn: int = 0):
        """cuda_search_radius_neighbors (cuda_nn_search.cu:622-723): CSR (offsets, indices)."""
        t = _torch()
        n = query_pc.points_number()
        offs = t.empty(n + 1, dtype=t.int32, device="cuda")
        total = C.c_int64()
        ctx = query_pc.ctx.handle
        st = lib.gca_radius_search(ctx, query_pc.handle, reference_pc.handle, radius, max_nn,
                                   C.c_void_p(offs.data_ptr()), None, 0, C.byref(total))
        check(st, "gca_radius_search")
        idx = t.empty(max(total.value, 1), dtype=t.int32, device="cuda")
        st = lib.gca_radius_search(ctx, query_pc.handle, reference_pc.handle, radius, max_nn,
                                   C.c_void_p(offs.data_ptr()), C.c_void_p(idx.data_ptr()),
                                   total.value, C.byref(total))
        check(st, "gca_radius_search")
        query_pc.ctx.synchronize()
        return offs.cpu().numpy(), idx[:total.value].cpu().numpy()


def icp_params(maximum_iterations=30, max_correspondence_distance=0.05, lambda_geometric=0.968,
               accumulate_mode=L.ACCUM_COMPENSATED, conv_rotation=1e-6,
               conv_translation_sq=1e-6) -> L.IcpParams:
    return L.IcpParams(maximum_iterations, max_correspondence_distance, lambda_geometric,
                       accumulate_mode, conv_rotation, conv_translation_sq)


class ColorICP:
    """gca::color_icp (src/registration/color_icp.hpp:16-63, color_icp.cu:17-261)."""

    m_color_icp_lambda = 0.968  # color_icp.hpp:46

    def __init__(self, maximum_iterations: int, max_correspondence_distance: float,
                 search_radius_color_gradient: float, accumulate_mode=L.ACCUM_COMPENSATED,
                 conv_rotation=1e-6, conv_translation_sq=1e-6):
        self.params = icp_params(maximum_iterations, max_correspondence_distance,
                                 self.m_color_icp_lambda, accumulate_mode, conv_rotation,
                                 conv_translation_sq)
        self.grad_radius = search_radius_color_gradient
        self.source: Optional[PointCloud] = None
        self.target: Optional[PointCloud] = None
        self.T = np.eye(4)
        self.rmse = 0.0
        self.done = False
        self.report = L.IcpReport()

    def set_source_point_cloud(self, pc: PointCloud):
        self.done, self.T, self.rmse = False, np.eye(4), 0.0
        self.source = pc.copy()  # color_icp.cu:37 copies the source

    def set_target_point_cloud(self, pc: PointCloud):
        self.done, self.T, self.rmse = False, np.eye(4), 0.0
        self.target = pc

    def align(self, T0=None) -> bool:
        """color_icp::align (color_icp.cu:151-234); the loop runs on the device."""
        if self.done:
            return True
        if self.source is None or self.target is None or not self.target.has_normals():
            return False
        T = np.zeros(16)
        st = lib.gca_color_icp_align(self.source.ctx.handle, self.source.handle, self.target.handle,
                                     C.byref(self.params), self.grad_radius,
                                     _ptr(_T(np.eye(4) if T0 is None else T0)), _ptr(T),
                                     C.byref(self.report))
        if st in (L.GCA_ERR_INVALID_VALUE, L.GCA_ERR_NOT_READY):
            return False
        check(st, "gca_color_icp_align")
        self.T = T.reshape(4, 4)
        self.rmse = self.report.rmse
        self.done = True
        return True

    def get_final_transformation_matrix(self):
        return self.T.copy()

    def get_RSME(self) -> float:  # (sic) reference spelling, color_icp.hpp:33
        return self.rmse

    def get_transformed_source_point_cloud(self) -> PointCloud:
        pc = self.source.copy()
        pc.transform(self.T)
        return pc

    def get_abs_residual(self):
        """color_icp::get_abs_residual (color_icp.cu:90-137) at the final pose."""
        t = _torch()
        moved = self.get_transformed_source_point_cloud()
        nn = PointCloud.nn_search(moved, self.target, self.params.max_correspondence_distance,
                                  as_tensor=True)
        n = moved.points_number()
        out = t.empty(max(n, 1), dtype=t.float32, device="cuda")
        s = C.c_double()
        check(lib.gca_compute_residuals(moved.ctx.handle, moved.handle, self.target.handle,
                                        C.c_void_p(nn.data_ptr()), self.params.lambda_geometric,
                                        _ptr(_T(np.eye(4))), L.RESIDUAL_ABS,
                                        C.c_void_p(out.data_ptr()), C.byref(s)),
              "gca_compute_residuals")
        moved.ctx.synchronize()
        return out[:n].cpu().numpy(), (s.value / n if n else 0.0)


# ---- free functions over the C ABI ------------------------------------------------------------

def build_gauss_newton(src: PointCloud, tgt: PointCloud, nn, lambda_geometric=0.968, T=None,
                       accumulate_mode=L.ACCUM_COMPENSATED):
    """cuda_build_gauss_newton_color_icp: returns (JTJ 6x6, JTr 6, sum r^2, n_corr)."""
    t = _torch()
    nn_t = nn if isinstance(nn, t.Tensor) else device_tensor(np.asarray(nn, np.int32))
    out = np.zeros(29)
    check(lib.gca_build_gauss_newton(src.ctx.handle, src.handle, tgt.handle,
                                     C.c_void_p(nn_t.data_ptr()), lambda_geometric,
                                     _ptr(_T(np.eye(4) if T is None else T)), accumulate_mode,
                                     _ptr(out)), "gca_build_gauss_newton")
    JTJ = np.zeros((6, 6))
    k = 0
    for a in range(6):
        for b in range(a, 6):
            JTJ[a, b] = JTJ[b, a] = out[k]
            k += 1
    return JTJ, out[21:27].copy(), out[27], int(out[28])


def compute_residuals(src: PointCloud, tgt: PointCloud, nn, kind, lambda_geometric=0.968, T=None):
    t = _torch()
    nn_t = nn if isinstance(nn, t.Tensor) else device_tensor(np.asarray(nn, np.int32))
    n = src.points_number()
    cols = 2 if kind == L.RESIDUAL_PAIR else 1
    out = t.empty(max(n * cols, 1), dtype=t.float32, device="cuda")
    s = C.c_double()
    check(lib.gca_compute_residuals(src.ctx.handle, src.handle, tgt.handle,
                                    C.c_void_p(nn_t.data_ptr()), lambda_geometric,
                                    _ptr(_T(np.eye(4) if T is None else T)), kind,
                                    C.c_void_p(out.data_ptr()), C.byref(s)), "gca_compute_residuals")
    src.ctx.synchronize()
    vals = out[:n * cols].cpu().numpy()
    return (vals.reshape(n, 2) if cols == 2 else vals), s.value


def solve_gauss_newton(JTJ, JTr):
    """solve_JTJ_JTr<true> (eigen_solver.hpp:44-85): (status, xi, T_step)."""
    JTJ = np.asarray(JTJ, np.float64)
    up = np.ascontiguousarray([JTJ[a, b] for a in range(6) for b in range(a, 6)], np.float64)
    jtr = np.ascontiguousarray(JTr, np.float64)
    xi = np.zeros(6)
    T = np.zeros(16)
    st = lib.gca_solve_gauss_newton(_ptr(up), _ptr(jtr), _ptr(xi), _ptr(T))
    return st, xi, T.reshape(4, 4)


def objective_grad(src: PointCloud, tgt: PointCloud, nn, T, xi, lambda_geometric=0.968,
                   want_points=True):
    """Colored-ICP objective at dT(xi) T and its adjoint: (E, dE/dxi, dE/d(xyz,I) or None)."""
    t = _torch()
    nn_t = nn if isinstance(nn, t.Tensor) else device_tensor(np.asarray(nn, np.int32))
    n = src.points_number()
    dp = t.empty((max(n, 1), 4), dtype=t.float32, device="cuda") if want_points else None
    E = C.c_double()
    dxi = np.zeros(6)
    check(lib.gca_color_icp_objective_grad(src.ctx.handle, src.handle, tgt.handle,
                                           C.c_void_p(nn_t.data_ptr()), lambda_geometric,
                                           _ptr(_T(T)), _ptr(np.ascontiguousarray(xi, np.float64)),
                                           C.byref(E), _ptr(dxi),
                                           C.c_void_p(dp.data_ptr()) if dp is not None else None),
          "gca_color_icp_objective_grad")
    src.ctx.synchronize()
    return E.value, dxi, (dp[:n] if dp is not None else None)


def batch_icp(srcs, tgts, params: L.IcpParams, T0s, gradient_radius=0.0):
    """gca_color_icp_batch: all pairs in one persistent launch.  Returns (T [n,4,4], reports)."""
    n = len(srcs)
    ctx = srcs[0].ctx
    hs = (C.c_void_p * n)(*[s.handle.value for s in srcs])
    ht = (C.c_void_p * n)(*[t_.handle.value for t_ in tgts])
    T0 = np.ascontiguousarray(np.asarray(T0s, np.float64).reshape(n, 16))
    T = np.zeros((n, 16))
    reps = (L.IcpReport * n)()
    check(lib.gca_color_icp_batch(ctx.handle, n, hs, ht, C.byref(params), gradient_radius,
                                  _ptr(T0), _ptr(T), reps), "gca_color_icp_batch")
    return T.reshape(n, 4, 4), list(reps)


def batch_multiscale_icp(srcs, tgts, params, T0s, gradient_radii):
    """gca_color_icp_batch_multiscale: srcs[k][l], tgts[k][l] = pair k at level l (coarse to
    fine); params / gradient_radii one per level.  All pairs and levels run in one persistent
    launch.  Returns (T [n,4,4], reports [n][levels])."""
    n, nl = len(srcs), len(params)
    if any(len(s) != nl for s in srcs) or any(len(t_) != nl for t_ in tgts) or len(tgts) != n:
        raise ValueError("srcs / tgts must be n_pairs x n_levels")
    ctx = srcs[0][0].ctx
    hs = (C.c_void_p * (n * nl))(*[c.handle.value for row in srcs for c in row])
    ht = (C.c_void_p * (n * nl))(*[c.handle.value for row in tgts for c in row])
    prm = (L.IcpParams * nl)(*params)
    gr = (C.c_float * nl)(*[float(r) for r in gradient_radii])
    T0 = np.ascontiguousarray(np.asarray(T0s, np.float64).reshape(n, 16))
    T = np.zeros((n, 16))
    reps = (L.IcpReport * (n * nl))()
    check(lib.gca_color_icp_batch_multiscale(ctx.handle, n, nl, hs, ht, prm, gr, _ptr(T0), _ptr(T), reps),
          "gca_color_icp_batch_multiscale")
    return T.reshape(n, 4, 4), [list(reps[k * nl:(k + 1) * nl]) for k in range(n)]


class BatchPlan:
    """gca_icp_batch_plan: batch_multiscale_icp set up once, run many times.  srcs[k][l],
    tgts[k][l], params, gradient_radii and T0s as for batch_multiscale_icp.  enqueue() queues a
    run without a host round trip (graph=True: the plan replays a graph of its own after the
    first run); results() synchronises and returns (T [n,4,4], reports [n][levels])."""

    def __init__(self, srcs, tgts, params, T0s, gradient_radii, graph: bool = True):
        n, nl = len(srcs), len(params)
        if any(len(s) != nl for s in srcs) or any(len(t_) != nl for t_ in tgts) or len(tgts) != n:
            raise ValueError("srcs / tgts must be n_pairs x n_levels")
        self.ctx = srcs[0][0].ctx
        self.n, self.nl = n, nl
        self._clouds = (srcs, tgts)   # the plan reads them: keep them alive
        hs = (C.c_void_p * (n * nl))(*[c.handle.value for row in srcs for c in row])
        ht = (C.c_void_p * (n * nl))(*[c.handle.value for row in tgts for c in row])
        prm = (L.IcpParams * nl)(*params)
        gr = (C.c_float * nl)(*[float(r) for r in gradient_radii])
        T0 = np.ascontiguousarray(np.asarray(T0s, np.float64).reshape(n, 16))
        h = C.c_void_p()
        check(lib.gca_icp_batch_plan_create(self.ctx.handle, n, nl, hs, ht, prm, gr, _ptr(T0), int(graph),
                                            C.byref(h)), "gca_icp_batch_plan_create")
        self.handle = h

    def enqueue(self) -> None:
        check(lib.gca_icp_batch_plan_enqueue(self.handle), "gca_icp_batch_plan_enqueue")

    def set_initial(self, T0s) -> None:
        T0 = np.ascontiguousarray(np.asarray(T0s, np.float64).reshape(self.n, 16))
        check(lib.gca_icp_batch_plan_set_initial(self.handle, _ptr(T0)), "gca_icp_batch_plan_set_initial")

    def results(self):
        T = np.zeros((self.n, 16))
        reps = (L.IcpReport * (self.n * self.nl))()
        check(lib.gca_icp_batch_plan_results(self.handle, _ptr(T), reps), "gca_icp_batch_plan_results")
        nl = self.nl
        return T.reshape(self.n, 4, 4), [list(reps[k * nl:(k + 1) * nl]) for k in range(self.n)]

    @property
    def captures(self) -> int:
        v = C.c_int64()
        check(lib.gca_icp_batch_plan_captures(self.handle, C.byref(v)), "gca_icp_batch_plan_captures")
        return v.value

    def close(self) -> None:
        if self.handle:
            check(lib.gca_icp_batch_plan_destroy(self.handle), "gca_icp_batch_plan_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def level_params(voxel, normal_radius=None, normal_max_nn=30, gradient_radius=None,
                 gradient_max_nn=0, icp: Optional[L.IcpParams] = None) -> L.LevelParams:
    """One level of the coarse-to-fine schedule; radii default to 2.5 v (normals,
    correspondences) and 2 v (colour gradient), the reference's 0.02 setting scaled
    (apps/test_color_icp.cpp:406-410)."""
    nr = 2.5 * voxel if normal_radius is None else normal_radius
    gr = 2.0 * voxel if gradient_radius is None else gradient_radius
    p = icp or icp_params(30, 2.5 * voxel)
    return L.LevelParams(voxel, nr, normal_max_nn, gr, gradient_max_nn, p)


def multiscale_icp(src_full: PointCloud, tgt_full: PointCloud, levels, T0=None):
    arr = (L.LevelParams * len(levels))(*levels)
    reps = (L.IcpReport * len(levels))()
    T = np.zeros(16)
    check(lib.gca_multiscale_icp(src_full.ctx.handle, src_full.handle, tgt_full.handle, len(levels),
                                 arr, _ptr(_T(np.eye(4) if T0 is None else T0)), _ptr(T), reps),
          "gca_multiscale_icp")
    return T.reshape(4, 4), list(reps)
