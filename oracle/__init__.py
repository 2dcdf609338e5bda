"""FP64 host oracle of the RGB-D front end + colored-ICP path (ctypes over oracle/_build).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import this package, and only as the checker / CPU baseline.  The
product (``cuda_3d_amd`` / ``libgca.so``) never imports or links it.

Each wrapper cites the reference function it restates (see ``gca_oracle.c`` for the
file:line of every step).  Arrays are numpy, float4 SoA as in the product
(``xyzi``: x, y, z, intensity; ``rgb``: r, g, b, 0; normals/gradients: x, y, z, 0).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgca_oracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


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
    """gca_icp_report (include/gca/gca.h), which the oracle fills; latency_us stays 0 here."""
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("solver_failures", C.c_int32), ("n_correspondences", C.c_int32),
                ("rmse", C.c_double), ("fitness", C.c_double), ("latency_us", C.c_double),
                ("active_us", C.c_double)]


class LevelParams(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("normal_radius", C.c_float),
                ("normal_max_nn", C.c_int32), ("gradient_radius", C.c_float),
                ("gradient_max_nn", C.c_int32), ("icp", IcpParams)]


def build():
    """Compile the oracle and the CPU baseline (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_variant(plant: int, out_dir: str) -> str:
    """Compile the oracle with a planted mistake (-DORC_PLANT=plant) into out_dir."""
    out = os.path.join(out_dir, f"libgca_oracle_plant{plant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", f"-DORC_PLANT={plant}",
                    "-shared", "-o", out, os.path.join(_HERE, "gca_oracle.c"), "-lm"], check=True)
    return out


_path = _LIB_PATH


@contextlib.contextmanager
def using(path: str):
    """Temporarily route every wrapper through another build of the oracle."""
    global _lib, _path
    old = (_lib, _path)
    _lib, _path = None, path
    try:
        yield
    finally:
        _lib, _path = old


def lib():
    global _lib
    if _lib is None:
        if _path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_path)
        vp = C.c_void_p
        L.orc_cloud_from_rgbd.restype = C.c_int64
        L.orc_cloud_from_rgbd.argtypes = [u16p, u8p, C.POINTER(Camera), C.POINTER(Bilateral),
                                          C.c_float, C.c_float, f32p, f32p, i32p, f64p]
        L.orc_bounds.argtypes = [f32p, C.c_int64, f32p, f32p]
        L.orc_intensity.restype = C.c_float
        L.orc_intensity.argtypes = [C.c_float] * 3
        L.orc_voxel_downsample.restype = C.c_int64
        L.orc_voxel_downsample.argtypes = [f32p, f32p, C.c_int64, f32p, f32p, C.c_float, f32p,
                                           f32p, i32p]
        L.orc_nn_search.argtypes = [f32p, C.c_int64, f32p, C.c_int64, f32p, f32p, C.c_float, i32p]
        L.orc_radius_search.restype = C.c_int64
        L.orc_radius_search.argtypes = [f32p, C.c_int64, f32p, C.c_int64, f32p, f32p, C.c_float,
                                        C.c_int32, vp, vp]
        L.orc_estimate_normals.argtypes = [f32p, C.c_int64, f32p, f32p, C.c_float, C.c_int32, f32p]
        L.orc_eigen3.argtypes = [f64p, f64p, f64p]
        L.orc_abi_sizes.argtypes = [i64p]
        L.orc_color_gradient.argtypes = [f32p, f32p, C.c_int64, f32p, f32p, C.c_float, C.c_int32,
                                         f32p]
        L.orc_build_gn.argtypes = [f32p, C.c_int64, f32p, f32p, f32p, i32p, C.c_float, f64p, f64p]
        L.orc_residuals.restype = C.c_double
        L.orc_residuals.argtypes = [f32p, C.c_int64, f32p, f32p, f32p, i32p, C.c_float, f64p,
                                    C.c_int, f64p]
        L.orc_solve6.restype = C.c_int
        L.orc_solve6.argtypes = [f64p, f64p, f64p, f64p]
        L.orc_euler_to_T.argtypes = [f64p, f64p]
        L.orc_icp_align.argtypes = [f32p, C.c_int64, f32p, f32p, f32p, C.c_int64, f32p, f32p,
                                    C.POINTER(IcpParams), f64p, f64p, C.POINTER(IcpReport)]
        L.orc_objective_grad.restype = C.c_double
        L.orc_objective_grad.argtypes = [f32p, C.c_int64, f32p, f32p, f32p, i32p, C.c_float, f64p,
                                         f64p, f64p, vp]
        L.orc_radius_outlier_removal.restype = C.c_int64
        L.orc_radius_outlier_removal.argtypes = [f32p, f32p, C.c_int64, f32p, f32p, C.c_float,
                                                 C.c_int32, f32p, f32p]
        L.orc_multiscale_icp.restype = C.c_int
        L.orc_multiscale_icp.argtypes = [f32p, f32p, C.c_int64, f32p, f32p, C.c_int64, C.c_int,
                                         C.POINTER(LevelParams), f64p, f64p, C.POINTER(IcpReport)]
        L.orc_grid_build.argtypes = [vp, f32p, C.c_int64, f32p, f32p, C.c_float]
        _lib = L
    return _lib


def _f4(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 4)


def _b3(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(3)


def make_camera(width, height, fx, fy, cx, cy, depth_scale=0.001, color_intrin=None,
                rotation=None, translation=None) -> Camera:
    cam = Camera()
    cam.width, cam.height, cam.depth_scale = width, height, depth_scale
    cam.depth_intrin = Intrinsics(fx, fy, cx, cy)
    ci = color_intrin if color_intrin is not None else (fx, fy, cx, cy)
    cam.color_intrin = Intrinsics(*ci)
    R = rotation if rotation is not None else (1, 0, 0, 0, 1, 0, 0, 0, 1)
    t = translation if translation is not None else (0, 0, 0)
    cam.depth_to_color = Extrinsics((C.c_float * 9)(*R), (C.c_float * 3)(*t))
    return cam


def cloud_from_rgbd(depth, bgr, cam: Camera, bil: Bilateral, tmin, tmax):
    """cuda_make_point_cloud (cuda_point_cloud_factory.cu:269-313). Returns (xyzi, rgb, pix, filtered)."""
    depth = np.ascontiguousarray(depth, dtype=np.uint16)
    bgr = np.ascontiguousarray(bgr, dtype=np.uint8)
    P = cam.width * cam.height
    xyzi = np.zeros((P, 4), np.float32)
    rgb = np.zeros((P, 4), np.float32)
    pix = np.zeros(P, np.int32)
    filt = np.zeros(P, np.float64)
    n = lib().orc_cloud_from_rgbd(depth.reshape(-1), bgr.reshape(-1), C.byref(cam), C.byref(bil),
                                  tmin, tmax, xyzi, rgb, pix, filt)
    if n < 0:
        raise ValueError("invalid camera parameters")
    return xyzi[:n].copy(), rgb[:n].copy(), pix[:n].copy(), filt


def bounds(xyzi):
    """cuda_compute_min_max_bound (geometry_util.cuh:88-105)."""
    xyzi = _f4(xyzi)
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    lib().orc_bounds(xyzi, len(xyzi), mn, mx)
    return mn, mx


def intensity(r, g, b):
    return lib().orc_intensity(r, g, b)


def voxel_downsample(xyzi, rgb, voxel, bmin=None, bmax=None):
    """cuda_voxel_grid_downsample (cuda_voxel_grid_down_sample.cu:123-161). Returns (xyzi, rgb,
    keys) with keys int32 [M, 3] (kx, ky, kz) in the reference's lexicographic order."""
    xyzi, rgb = _f4(xyzi), _f4(rgb)
    if bmin is None:
        bmin, bmax = bounds(xyzi)
    n = len(xyzi)
    ox = np.zeros((max(n, 1), 4), np.float32)
    orgb = np.zeros((max(n, 1), 4), np.float32)
    keys = np.zeros((max(n, 1), 3), np.int32)
    m = lib().orc_voxel_downsample(xyzi, rgb, n, _b3(bmin), _b3(bmax), voxel, ox, orgb, keys)
    if m < 0:
        raise ValueError("invalid voxel size")
    return ox[:m].copy(), orgb[:m].copy(), keys[:m].copy()


class _Grid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("cell", C.c_float), ("dims", C.c_int32 * 3),
                ("ncells", C.c_int64), ("start", C.POINTER(C.c_int32)),
                ("order", C.POINTER(C.c_int32)), ("n", C.c_int32)]


def abi_sizes():
    """sizeof of (gca_camera, gca_bilateral, gca_icp_params, gca_icp_report, gca_level_params,
    orc_grid) as compiled into the oracle."""
    out = np.zeros(6, np.int64)
    lib().orc_abi_sizes(out)
    return [int(v) for v in out]


def grid_cells(xyzi, radius, bmin=None, bmax=None):
    """compute_grid_cell_functor (cuda_nn_search.cu:36-59) for our canonical grid.
    Returns (cells[n,3], origin[3], cell_size, dims[3])."""
    xyzi = _f4(xyzi)
    if bmin is None:
        bmin, bmax = bounds(xyzi)
    L = lib()
    g = _Grid()
    L.orc_grid_build(C.byref(g), xyzi, len(xyzi), _b3(bmin), _b3(bmax), radius)
    cells = np.zeros((max(len(xyzi), 1), 3), np.int32)
    L.orc_grid_cells.argtypes = [C.c_void_p, f32p, C.c_int64, i32p]
    L.orc_grid_cells(C.byref(g), xyzi, len(xyzi), cells)
    origin = np.array(g.origin[:], np.float32)
    cell = float(np.float32(g.cell))
    dims = np.array(g.dims[:], np.int32)
    L.orc_grid_free.argtypes = [C.c_void_p]
    L.orc_grid_free(C.byref(g))
    return cells[:len(xyzi)].copy(), origin, cell, dims


def nn_search(q_xyzi, r_xyzi, radius, rbmin=None, rbmax=None):
    """cuda_nn_search (cuda_nn_search.cu:545-619)."""
    q, r = _f4(q_xyzi), _f4(r_xyzi)
    if rbmin is None:
        rbmin, rbmax = bounds(r)
    out = np.zeros(max(len(q), 1), np.int32)
    lib().orc_nn_search(q, len(q), r, len(r), _b3(rbmin), _b3(rbmax), radius, out)
    return out[:len(q)].copy()


def radius_search(q_xyzi, r_xyzi, radius, max_nn=0, rbmin=None, rbmax=None):
    """cuda_search_radius_neighbors (cuda_nn_search.cu:622-723). Returns (offsets, idx)."""
    q, r = _f4(q_xyzi), _f4(r_xyzi)
    if rbmin is None:
        rbmin, rbmax = bounds(r)
    L = lib()
    offs = np.zeros(len(q) + 1, np.int32)
    total = L.orc_radius_search(q, len(q), r, len(r), _b3(rbmin), _b3(rbmax), radius, max_nn,
                                offs.ctypes.data, None)
    idx = np.zeros(max(total, 1), np.int32)
    L.orc_radius_search(q, len(q), r, len(r), _b3(rbmin), _b3(rbmax), radius, max_nn,
                        offs.ctypes.data, idx.ctypes.data)
    return offs, idx[:total].copy()


def euclidean_clustering(xyzi, tolerance, min_cluster_size, max_cluster_size, bmin=None, bmax=None,
                         residual=None):
    """cuda_euclidean_clustering (cuda_clustering.cu:31-110): for each unvisited point in index
    order, a breadth-first search over the radius-neighbour lists (cuda_search_radius_neighbors,
    here radius_search, so each list in (squared distance, index) order) that stops expanding once
    its queue holds max_cluster_size points (:75-96); the queue is kept as a cluster when it has at
    least min_cluster_size points (:98-103).  With ``residual`` (one float per point) this is
    cuda_residual_based_segmentation (:251-341): a neighbour is taken only when
    |residual[p] - residual[j]| (FP32) < 0.003 for the point p being expanded (:291-293, :316-319),
    and is otherwise left unvisited.  Returns the clusters as int32 arrays in queue order.  Pure
    Python: small inputs."""
    if min_cluster_size < 0 or max_cluster_size <= 0 or max_cluster_size < min_cluster_size:
        raise ValueError("invalid cluster sizes")  # :38-42
    xyzi = _f4(xyzi)
    n = len(xyzi)
    if n == 0:
        return []
    res = None if residual is None else np.ascontiguousarray(residual, np.float32).reshape(-1)
    if res is not None and len(res) < n:
        raise ValueError("one residual per point")
    offs, idx = radius_search(xyzi, xyzi, tolerance, 0, bmin, bmax)
    visited = np.zeros(n, bool)
    out = []
    for i in range(n):
        if visited[i]:
            continue
        queue = [i]
        visited[i] = True
        k = 0
        while k < len(queue) and len(queue) < max_cluster_size:
            p = queue[k]
            for j in idx[offs[p]:offs[p + 1]].tolist():
                if visited[j]:
                    continue
                if res is not None and not float(abs(res[p] - res[j])) < 0.003:
                    continue
                visited[j] = True
                queue.append(j)
            k += 1
        if len(queue) >= min_cluster_size:
            out.append(np.array(queue, np.int32))
    return out


def estimate_normals(xyzi, radius, max_nn=0, bmin=None, bmax=None):
    """cuda_estimate_normals (cuda_estimate_normals.cu:153-199)."""
    xyzi = _f4(xyzi)
    if bmin is None:
        bmin, bmax = bounds(xyzi)
    out = np.zeros((max(len(xyzi), 1), 4), np.float32)
    lib().orc_estimate_normals(xyzi, len(xyzi), _b3(bmin), _b3(bmax), radius, max_nn, out)
    return out[:len(xyzi)].copy()


def eigen3(A):
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    ev = np.zeros(3)
    V = np.zeros(9)
    lib().orc_eigen3(A, ev, V)
    return ev, V.reshape(3, 3)


def color_gradient(xyzi, normals, radius, max_nn=0, bmin=None, bmax=None):
    """cuda_compute_color_gradient (cuda_compute_color_gradient.cu:92-127)."""
    xyzi = _f4(xyzi)
    if bmin is None:
        bmin, bmax = bounds(xyzi)
    out = np.zeros((max(len(xyzi), 1), 4), np.float32)
    lib().orc_color_gradient(xyzi, _f4(normals), len(xyzi), _b3(bmin), _b3(bmax), radius, max_nn,
                             out)
    return out[:len(xyzi)].copy()


def _T(T):
    return np.ascontiguousarray(np.asarray(T, np.float64).reshape(16))


def transform(xyzi, T, vectors=None):
    """point_cloud::transform (point_cloud.cu:93-105): transform_point_functor / transform_normal_functor
    (geometry_util.cuh:133-165) apply R p + t to the points and R v to normals.  The reference
    multiplies in FP32; this restatement (like the product) forms ((r0 x + r1 y) + r2 z) + t in FP64
    and rounds once to FP32, so it sits within an FP32 ulp or two of the reference's arithmetic.
    Returns (xyzi', vectors')."""
    T = np.asarray(T, np.float64).reshape(4, 4)
    x = _f4(xyzi).astype(np.float64)
    out = _f4(xyzi).copy()
    for r in range(3):
        out[:, r] = (((T[r, 0] * x[:, 0] + T[r, 1] * x[:, 1]) + T[r, 2] * x[:, 2]) + T[r, 3]).astype(np.float32)
    vout = None
    if vectors is not None:
        v = _f4(vectors).astype(np.float64)
        vout = _f4(vectors).copy()
        for r in range(3):
            vout[:, r] = ((T[r, 0] * v[:, 0] + T[r, 1] * v[:, 1]) + T[r, 2] * v[:, 2]).astype(np.float32)
    return out, vout


def accumulate(model_xyzi, model_rgb, frame_xyzi, frame_rgb, T, voxel=None):
    """One reconstruction step of apps/app.cpp:336-348: the model is moved by T
    (point_cloud::transform), the new frame is appended (point_cloud::operator+=,
    point_cloud.cu:55-73: plain concatenation, model first) and, when ``voxel`` is given, the sum
    is voxel-downsampled (app.cpp:345-349).  Returns (xyzi, rgb)."""
    moved, _ = transform(model_xyzi, T)
    xyzi = np.concatenate([moved, _f4(frame_xyzi)])
    rgb = np.concatenate([_f4(model_rgb), _f4(frame_rgb)])
    if voxel is None:
        return xyzi, rgb
    ox, orgb, _ = voxel_downsample(xyzi, rgb, voxel)
    return ox, orgb


def build_gn(src, tgt, tn, tg, nn, lam, T=None):
    """cuda_build_gauss_newton_color_icp (.cu:431-474) in FP64. Returns 29 doubles."""
    T = np.eye(4) if T is None else T
    out = np.zeros(29)
    src = _f4(src)
    lib().orc_build_gn(src, len(src), _f4(tgt), _f4(tn), _f4(tg),
                       np.ascontiguousarray(nn, np.int32), lam, _T(T), out)
    return out


def residuals(src, tgt, tn, tg, nn, lam, kind, T=None):
    T = np.eye(4) if T is None else T
    src = _f4(src)
    out = np.zeros(2 * max(len(src), 1))
    s = lib().orc_residuals(src, len(src), _f4(tgt), _f4(tn), _f4(tg),
                            np.ascontiguousarray(nn, np.int32), lam, _T(T), kind, out)
    k = 2 if kind == 0 else 1
    return out[:k * len(src)].reshape(len(src), k).squeeze(-1) if k == 1 else \
        out[:2 * len(src)].reshape(len(src), 2), s


def solve6(jtj21, jtr):
    """solve_JTJ_JTr<true> (eigen_solver.hpp:44-85)."""
    xi = np.zeros(6)
    T = np.zeros(16)
    st = lib().orc_solve6(np.ascontiguousarray(jtj21, np.float64),
                          np.ascontiguousarray(jtr, np.float64), xi, T)
    return st, xi, T.reshape(4, 4)


def euler_to_T(xi):
    T = np.zeros(16)
    lib().orc_euler_to_T(np.ascontiguousarray(xi, np.float64), T)
    return T.reshape(4, 4)


def icp_params(max_iterations=30, max_corr=0.05, lam=0.968, conv_rotation=1e-6,
               conv_translation_sq=1e-6, accumulate_mode=0) -> IcpParams:
    return IcpParams(max_iterations, max_corr, lam, accumulate_mode, conv_rotation,
                     conv_translation_sq)


def icp_align(src, tgt, tn, tg, params: IcpParams, T0=None, tbmin=None, tbmax=None):
    """color_icp::align (color_icp.cu:151-234). Returns (T, report)."""
    src, tgt = _f4(src), _f4(tgt)
    if tbmin is None:
        tbmin, tbmax = bounds(tgt)
    T0 = np.eye(4) if T0 is None else T0
    T = np.zeros(16)
    rep = IcpReport()
    lib().orc_icp_align(src, len(src), tgt, _f4(tn), _f4(tg), len(tgt), _b3(tbmin), _b3(tbmax),
                        C.byref(params), _T(T0), T, C.byref(rep))
    return T.reshape(4, 4), rep


def objective_grad(src, tgt, tn, tg, nn, lam, T, xi, want_points=True):
    """Objective of compute_RMSE_functor (.cu:82-132) and its adjoint. Returns (E, dxi, dpts)."""
    src = _f4(src)
    dxi = np.zeros(6)
    dp = np.zeros((max(len(src), 1), 4)) if want_points else None
    E = lib().orc_objective_grad(src, len(src), _f4(tgt), _f4(tn), _f4(tg),
                                 np.ascontiguousarray(nn, np.int32), lam, _T(T),
                                 np.ascontiguousarray(xi, np.float64), dxi,
                                 dp.ctypes.data if dp is not None else None)
    return E, dxi, (dp[:len(src)] if dp is not None else None)


def radius_outlier_removal(xyzi, rgb, radius, min_nn, bmin=None, bmax=None):
    """cuda_grid_radius_outliers_removal (cuda_nn_search.cu:725-800)."""
    xyzi, rgb = _f4(xyzi), _f4(rgb)
    if bmin is None:
        bmin, bmax = bounds(xyzi)
    ox = np.zeros((max(len(xyzi), 1), 4), np.float32)
    orgb = np.zeros_like(ox)
    m = lib().orc_radius_outlier_removal(xyzi, rgb, len(xyzi), _b3(bmin), _b3(bmax), radius,
                                         min_nn, ox, orgb)
    if m < 0:
        raise ValueError("invalid radius")
    return ox[:m].copy(), orgb[:m].copy()


def multiscale_icp(src_xyzi, src_rgb, tgt_xyzi, tgt_rgb, levels, T0=None):
    """Coarse-to-fine colored ICP (apps/test_color_icp.cpp:378-414). Returns (T, reports)."""
    s, sr, t, tr = _f4(src_xyzi), _f4(src_rgb), _f4(tgt_xyzi), _f4(tgt_rgb)
    arr = (LevelParams * len(levels))(*levels)
    reps = (IcpReport * len(levels))()
    T = np.zeros(16)
    T0 = np.eye(4) if T0 is None else T0
    rc = lib().orc_multiscale_icp(s, sr, len(s), t, tr, len(t), len(levels), arr, _T(T0), T, reps)
    if rc != 0:
        raise ValueError("invalid level parameters")
    return T.reshape(4, 4), list(reps)


# ---- CPU baseline (oracle/cpu_pipeline.cpp): the benchmark's frame step, FP32, OpenMP ----------

class CpuPipelineParams(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float),
                ("cy", C.c_float), ("scale", C.c_float), ("tmin", C.c_float), ("tmax", C.c_float),
                ("radius", C.c_int), ("sigma_space", C.c_float), ("sigma_depth", C.c_float),
                ("n_levels", C.c_int), ("voxel", C.c_float * 8), ("max_iter", C.c_int), ("lam", C.c_float),
                ("conv_rot", C.c_double), ("conv_t2", C.c_double)]


_cpb = None


def _cpu_lib():
    global _cpb
    if _cpb is None:
        path = os.path.join(_HERE, "_build", "libgca_cpu_pipeline.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.cpb_run.restype = C.c_int
        L.cpb_run.argtypes = [u16p, u8p, u16p, u8p, C.POINTER(CpuPipelineParams), C.c_int, C.c_int, C.c_int,
                              f64p, f64p, i32p, i64p]
        _cpb = L
    return _cpb


def cpu_pipeline_run(frame_a, frame_b, W, H, fx, fy, cx, cy, levels, max_iter, conv=(1e-6, 1e-6),
                     threads=0, warmup=3, reps=20, lam=0.968, bilateral=(2, 2.0, 30.0), tmin=0.2, tmax=3.5):
    """Times the frame step of frame_b ((depth, bgr)) registered onto frame_a with the
    multithreaded FP32 CPU pipeline: front end, per level voxel / grid / normals (r 2.5 v, K 30) /
    colour gradient (r 2 v), colored ICP (corr 2.5 v).  threads = 0: OpenMP's default.  Returns
    dict(ms=[reps], threads, T, iterations, points)."""
    (da, ca), (db, cb) = frame_a, frame_b
    prm = CpuPipelineParams(W, H, fx, fy, cx, cy, 0.001, tmin, tmax, bilateral[0], bilateral[1], bilateral[2],
                            len(levels), (C.c_float * 8)(*levels), max_iter, lam, conv[0], conv[1])
    ms = np.zeros(max(reps, 1))
    T = np.zeros(16)
    its = np.zeros(8, np.int32)
    pts = np.zeros(9, np.int64)
    args = [np.ascontiguousarray(a, dt).reshape(-1) for a, dt in ((da, np.uint16), (ca, np.uint8), (db, np.uint16),
                                                                   (cb, np.uint8))]
    used = _cpu_lib().cpb_run(*args, C.byref(prm), threads, warmup, reps, ms, T, its, pts)
    return dict(ms=ms[:reps].tolist(), threads=int(used), T=T.reshape(4, 4), iterations=its[:len(levels)].tolist(),
                points=pts[:1 + len(levels)].tolist())
