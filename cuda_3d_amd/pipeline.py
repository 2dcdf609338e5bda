"""Per-frame RGB-D odometry pipeline over libgca.so.

One step takes a new depth + colour frame that is already resident in device memory and runs:
  front end     bilateral filter, back-projection, compaction (gca_cloud_from_rgbd)
  per level     voxel downsample, PCA normals (2.5 v, K nearest), colour gradient (2 v):
                the target-side data this frame needs when the next frame registers to it
  registration  coarse-to-fine colored ICP of the new frame onto the previous frame: all
                levels in one persistent launch, each level starting from the pose the
                previous one ends with (gca_color_icp_batch_multiscale, one pair)
This is the loop of the reference's demo (apps/app.cpp:111-372: create_from_rgbd ->
voxel_grid_down_sample -> estimate_normals -> color_icp.align against the previous cloud) with
the coarse-to-fine schedule of apps/test_color_icp.cpp:378-414.  All working clouds are
allocated once (two frame slots) and reused, so a step allocates nothing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .core import Context, PointCloud, _ptr, icp_params


class FramePipeline:
    def __init__(self, ctx: Context, camera: L.Camera, bilateral: L.Bilateral,
                 levels=(0.04, 0.02, 0.01), iters=30, normal_max_nn=30, tmin=0.2, tmax=3.5,
                 conv_rotation=1e-6, conv_translation_sq=1e-6, accumulate_mode=L.ACCUM_COMPENSATED):
        """conv_*: the stopping rule; the defaults are the reference's fixed thresholds
        (color_icp.hpp:45-47), 0, 0 runs every level to `iters`."""
        self.ctx, self.cam, self.bil = ctx, camera, bilateral
        self.levels, self.normal_max_nn, self.tmin, self.tmax = tuple(levels), normal_max_nn, tmin, tmax
        cap = camera.width * camera.height
        self.full = [PointCloud(ctx, cap) for _ in range(2)]
        self.lv = [[PointCloud(ctx, cap) for _ in self.levels] for _ in range(2)]
        nl = len(self.levels)
        self.params = (L.IcpParams * nl)(*[icp_params(iters, 2.5 * v, 0.968, accumulate_mode, conv_rotation,
                                                      conv_translation_sq) for v in self.levels])
        self.grad_r = (C.c_float * nl)(*[2.0 * v for v in self.levels])
        self.reports = (L.IcpReport * nl)()
        # level-cloud handle tables for (source slot, target slot) = (0, 1) and (1, 0)
        self._tables = {}
        for a in (0, 1):
            b = 1 - a
            self._tables[a] = ((C.c_void_p * nl)(*[c.handle.value for c in self.lv[a]]),
                               (C.c_void_p * nl)(*[c.handle.value for c in self.lv[b]]))
        self.slot = 0
        self.have_prev = False
        self._T = np.zeros(16)
        self._I = np.ascontiguousarray(np.eye(4).reshape(16))

    def frontend(self, slot: int, depth_ptr: int, bgr_ptr: int):
        lib, h = L.lib, self.ctx.handle
        full = self.full[slot]
        L.check(lib.gca_cloud_from_rgbd(h, C.c_void_p(depth_ptr), C.c_void_p(bgr_ptr), C.byref(self.cam),
                                        C.byref(self.bil), self.tmin, self.tmax, full.handle), "from_rgbd")
        for l, v in enumerate(self.levels):
            c = self.lv[slot][l]
            L.check(lib.gca_voxel_downsample(h, full.handle, v, c.handle), "voxel")
            L.check(lib.gca_estimate_normals(h, c.handle, 2.5 * v, self.normal_max_nn), "normals")
            L.check(lib.gca_compute_color_gradient(h, c.handle, 2.0 * v, 0), "gradient")

    def register(self, src_slot: int, tgt_slot: int, T0=None) -> np.ndarray:
        if tgt_slot != 1 - src_slot:
            raise ValueError("register() pairs the two frame slots")
        T = self._T
        T[:] = self._I if T0 is None else np.asarray(T0, np.float64).reshape(16)
        hs, ht = self._tables[src_slot]
        L.check(L.lib.gca_color_icp_batch_multiscale(self.ctx.handle, 1, len(self.levels), hs, ht, self.params,
                                                     self.grad_r, _ptr(T), _ptr(T), self.reports), "register")
        return T.reshape(4, 4).copy()

    def iterations(self):
        """Gauss-Newton iterations of each level in the last registration."""
        return [r.iterations for r in self.reports]

    def step(self, depth_ptr: int, bgr_ptr: int):
        """Process a new frame; returns the pose mapping it onto the previous frame (or None)."""
        new = self.slot
        self.frontend(new, depth_ptr, bgr_ptr)
        T = self.register(new, 1 - new) if self.have_prev else None
        self.have_prev = True
        self.slot = 1 - new
        return T

    def level_sizes(self, slot: int):
        return [c.points_number() for c in self.lv[slot]]
