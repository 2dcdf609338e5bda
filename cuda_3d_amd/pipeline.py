"""Per-frame RGB-D odometry pipeline: the native gca_pipeline (include/gca/gca.h).

One step takes a new depth + colour frame that is already resident in device memory and runs,
without a host round trip:
  front end     bilateral filter, back-projection, compaction
  per level     voxel downsample, PCA normals (2.5 v, K nearest), colour gradient (2 v): the
                target-side data this frame needs when the next frame registers to it
  registration  coarse-to-fine colored ICP of the new frame onto the previous frame: all
                levels in one persistent launch, each level starting from the pose the
                previous one ends with
This is the loop of the reference's demo (apps/app.cpp:111-372: create_from_rgbd ->
voxel_grid_down_sample -> estimate_normals -> color_icp.align against the previous cloud) with
the coarse-to-fine schedule of apps/test_color_icp.cpp:378-414.  Every buffer is allocated
once (two frame slots; `slots` when overlapped); with graph=True each slot's step is replayed
as a hipGraph.

overlap=True: a step builds the new frame and, beside it on a second stream, registers the
previous frame (built by the previous step); with up to four levels that use the joint normals
+ gradient walk, the levels are built in lockstep and consecutive steps alternate between two
lanes of streams, so two frames' steps also run side by side (four slots).  The poses are the
sequential pipeline's, one step later.  result() then reports the frame before the one just
stepped, and flush() registers the last stepped frame on its own.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .core import Context, _ptr, icp_params, level_params


class FramePipeline:
    def __init__(self, ctx: Context, camera: L.Camera, bilateral: L.Bilateral,
                 levels=(0.04, 0.02, 0.01), iters=30, normal_max_nn=30, tmin=0.2, tmax=3.5,
                 conv_rotation=1e-6, conv_translation_sq=1e-6, accumulate_mode=L.ACCUM_COMPENSATED,
                 graph=True, overlap=False):
        """conv_*: the stopping rule; the defaults are the reference's fixed thresholds
        (color_icp.hpp:45-47), 0, 0 runs every level to `iters`."""
        self.levels = tuple(levels)
        nl = len(self.levels)
        lv = (L.LevelParams * nl)(*[
            level_params(v, normal_max_nn=normal_max_nn,
                         icp=icp_params(iters, 2.5 * v, 0.968, accumulate_mode, conv_rotation, conv_translation_sq))
            for v in self.levels])
        h = C.c_void_p()
        L.check(L.lib.gca_pipeline_create(ctx.handle, C.byref(camera), C.byref(bilateral), tmin, tmax, nl, lv,
                                          C.byref(h)), "gca_pipeline_create")
        self.handle = h
        c = C.c_void_p()
        L.check(L.lib.gca_pipeline_ctx(h, C.byref(c)), "gca_pipeline_ctx")
        self.ctx = Context.borrow(c, ctx.device)  # the pipeline's own stream, profiler and tuning
        self.set_graph(graph)
        self.frames = 0
        self.overlap = False
        if overlap:
            self.set_overlap(True)
        self.reports = (L.IcpReport * nl)()
        self._T = np.zeros(16)

    def set_graph(self, enable: bool):
        L.check(L.lib.gca_pipeline_set_graph(self.handle, 1 if enable else 0), "gca_pipeline_set_graph")

    def set_overlap(self, enable: bool):
        """Overlapped or sequential steps; synchronises and starts a new sequence."""
        L.check(L.lib.gca_pipeline_set_overlap(self.handle, 1 if enable else 0), "gca_pipeline_set_overlap")
        self.overlap = bool(enable)
        self.frames = 0

    def flush(self):
        """Overlapped steps: register the last stepped frame now (its pose is otherwise computed
        by the next step)."""
        L.check(L.lib.gca_pipeline_flush(self.handle), "gca_pipeline_flush")

    @property
    def slots(self) -> int:
        """Frame slots of the current mode (step k uses slot k % slots)."""
        n = C.c_int32()
        L.check(L.lib.gca_pipeline_slots(self.handle, C.byref(n)), "gca_pipeline_slots")
        return int(n.value)

    @property
    def slot(self) -> int:
        """Frame slot the next step uses."""
        return self.frames % self.slots

    def step(self, depth_ptr: int, bgr_ptr: int):
        """Enqueue a new frame (asynchronous)."""
        L.check(L.lib.gca_pipeline_step(self.handle, C.c_void_p(depth_ptr), C.c_void_p(bgr_ptr)),
                "gca_pipeline_step")
        self.frames += 1

    def result(self):
        """Waits for the last step: the pose mapping the last registered frame onto the one before
        (sequential: the frame just stepped; overlapped: the one before it, or the last stepped
        frame after flush()), None before any registration, and the per-level reports."""
        valid = C.c_int32()
        L.check(L.lib.gca_pipeline_result(self.handle, _ptr(self._T), self.reports, C.byref(valid)),
                "gca_pipeline_result")
        return (self._T.reshape(4, 4).copy() if valid.value else None), list(self.reports)

    def iterations(self):
        """Gauss-Newton iterations of each level in the last registration."""
        return [r.iterations for r in self.result()[1]]

    def graph_stats(self):
        """(graph captures, front-end node updates) so far."""
        cap, upd = C.c_int64(), C.c_int64()
        L.check(L.lib.gca_pipeline_graph_stats(self.handle, C.byref(cap), C.byref(upd)), "gca_pipeline_graph_stats")
        return int(cap.value), int(upd.value)

    def sizes(self, slot: int):
        full = C.c_int64()
        lv = np.zeros(len(self.levels), np.int64)
        L.check(L.lib.gca_pipeline_sizes(self.handle, slot, C.byref(full), _ptr(lv)), "gca_pipeline_sizes")
        return int(full.value), [int(x) for x in lv]

    def level_sizes(self, slot: int):
        return self.sizes(slot)[1]

    def full_size(self, slot: int) -> int:
        return self.sizes(slot)[0]

    def close(self):
        if getattr(self, "handle", None):
            L.lib.gca_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
