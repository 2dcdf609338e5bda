"""Seeded workloads shared by bench.py and the autotuner.

frame_workload   the 640x480 odometry step: front end, 3 voxel levels with normals and colour
                 gradients, 3-level colored ICP against the previous frame (FramePipeline),
                 over the two frames of a seeded pair, alternating.
batch_pairs      uneven registration pairs for batched ICP: point counts log-uniform in
                 [5e3, 2e5], each pair its own seeded rigid transform (rotation <= 15 deg,
                 translation <= 0.2 m), a 3-level voxel pyramid per pair with the targets'
                 normals and colour gradients ready.  The scene is scaled 2x (make_scene) so
                 that the finest level (voxel 0.01) has room for 2e5 points.
"""
from __future__ import annotations

import math
import time

import numpy as np

LEVELS = (0.04, 0.02, 0.01)
LAMBDA = 0.968


def frame_workload(ctx, steps: int = 6, W: int = 640, H: int = 480, seed: int = 12, levels=LEVELS, iters=30,
                   overlap: bool = False):
    """Returns run() -> None that pushes `steps` frames through a FramePipeline and syncs.
    overlap=False: sequential eager steps (the stage profiler brackets each launch); overlap=True:
    overlapped, graph-replayed steps (their throughput is what counts: run() returns nothing, the
    caller times it).  run.pipe is the pipeline, run.ctx its own ctx (where its tuning and
    profiling live)."""
    from . import core as g
    from . import scene
    from ._lib import BILATERAL_FIXED
    from .pipeline import FramePipeline

    (td, tc), (sd, sc), _, cm = scene.frame_pair(seed, W, H)
    cam = g.make_camera(W, H, cm.fx, cm.fy, cm.cx, cm.cy)
    bil = g.bilateral(BILATERAL_FIXED, 2, 2.0, 30.0)
    dev = [(g.device_tensor(td.view(np.int16)), g.device_tensor(tc)),
           (g.device_tensor(sd.view(np.int16)), g.device_tensor(sc))]
    ptrs = [(d.data_ptr(), c.data_ptr()) for d, c in dev]
    pipe = FramePipeline(ctx, cam, bil, levels, iters, graph=overlap, overlap=overlap)
    state = {"k": 0}

    def run():
        for _ in range(steps):
            d, c = ptrs[state["k"] % 2]
            pipe.step(d, c)
            state["k"] += 1
        pipe.ctx.synchronize()

    run.pipe = pipe
    run.ctx = pipe.ctx
    run.frames = ((td, tc), (sd, sc))
    run.camera = cam
    run._keep = dev  # device buffers live as long as the closure
    return run


class BatchPairs:
    """n_pairs x len(levels) clouds ready for gca_color_icp_batch_multiscale."""

    def __init__(self, ctx, n_pairs: int = 64, levels=LEVELS, iters: int = 30, seed: int = 2026,
                 scale: float = 2.0, conv=(1e-6, 1e-6), log=None, keep_host: bool = False):
        from . import core as g
        from . import scene

        rng = np.random.default_rng(np.random.PCG64(seed))
        self.counts = np.rint(np.exp(rng.uniform(math.log(5e3), math.log(2e5), n_pairs))).astype(int)
        t0 = time.perf_counter()
        self.src, self.tgt = [], []   # [pair][level]
        self.truth = []
        self.host = [] if keep_host else None   # [pair] (src_xyz, src_rgb, tgt_xyz, tgt_rgb)
        for k, n in enumerate(self.counts):
            sx, sr, tx, tr, G = scene.sample_surface_pair(seed + 17 * k, int(n), scale=scale)
            if keep_host:
                self.host.append((sx, sr, tx, tr))
            s_full = g.PointCloud.from_numpy(sx, sr, ctx=ctx)
            t_full = g.PointCloud.from_numpy(tx, tr, ctx=ctx)
            srow, trow = [], []
            for v in levels:
                s2 = s_full.voxel_grid_down_sample(v)
                t2 = t_full.voxel_grid_down_sample(v)
                t2.estimate_normals(2.5 * v, 30)
                t2.compute_color_gradient(2.0 * v, 0)
                srow.append(s2)
                trow.append(t2)
            self.src.append(srow)
            self.tgt.append(trow)
            self.truth.append(G)
        ctx.synchronize()
        self.levels = tuple(levels)
        self.iters, self.conv = iters, tuple(conv)
        self.params = [g.icp_params(iters, 2.5 * v, LAMBDA, 0, conv[0], conv[1]) for v in levels]
        self.grad_r = [2.0 * v for v in levels]
        self.T0 = np.tile(np.eye(4), (n_pairs, 1, 1))
        self.sizes = [[self.src[k][l].points_number() for k in range(n_pairs)] for l in range(len(levels))]
        self.tgt_sizes = [[self.tgt[k][l].points_number() for k in range(n_pairs)] for l in range(len(levels))]
        if log:
            log(f"batch: {n_pairs} pairs, sampled {int(self.counts.sum())} points per side, finest level "
                f"{min(self.sizes[-1])}..{max(self.sizes[-1])} ({time.perf_counter() - t0:.1f}s)")

    def batched(self):
        """All pairs and levels in one persistent launch: (T [n,4,4], reports [n][levels])."""
        from . import core as g

        return g.batch_multiscale_icp(self.src, self.tgt, self.params, self.T0, self.grad_r)

    def plain(self, ctx):
        """Pair by pair, level by level (one gca_color_icp_align per level): (T, per-pair seconds)."""
        import ctypes as C

        from . import _lib as L

        n = len(self.src)
        T = np.zeros((n, 4, 4))
        lat = []
        rep = L.IcpReport()
        for k in range(n):
            t1 = time.perf_counter()
            Tk = np.ascontiguousarray(np.eye(4).reshape(16))
            for l in range(len(self.levels)):
                L.check(L.lib.gca_color_icp_align(ctx.handle, self.src[k][l].handle, self.tgt[k][l].handle,
                                                  C.byref(self.params[l]), self.grad_r[l],
                                                  Tk.ctypes.data_as(C.c_void_p), Tk.ctypes.data_as(C.c_void_p),
                                                  C.byref(rep)), "align")
            lat.append(time.perf_counter() - t1)
            T[k] = Tk.reshape(4, 4)
        return T, lat
