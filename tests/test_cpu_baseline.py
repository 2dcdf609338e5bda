"""The CPU baseline (oracle/cpu_pipeline.cpp, the multithreaded FP32 frame step that bench.py
times beside the GPU) computes the same registration as the FP64 oracle on the 320x240 frame
pair: FP32 front end, voxel means, normals and gradients move the converged pose by a few 1e-4
(the FP32 bilateral flips the colour pixel of some points), far below the motion itself, and the
thread count changes nothing but the order of the per-thread Gauss-Newton sums."""
import numpy as np

import oracle as O


def _pose_err(T, U):
    dR = T[:3, :3] @ U[:3, :3].T
    return float(np.arccos(np.clip((np.trace(dR) - 1) / 2, -1, 1))), float(np.linalg.norm(T[:3, 3] - U[:3, 3]))


def test_cpu_baseline_registers_like_the_oracle():
    from cuda_3d_amd import scene

    W, H = 320, 240
    (td, tc), (sd, sc), G, cm = scene.frame_pair(11, W, H)
    levels = (0.04, 0.02, 0.01)
    runs = {th: O.cpu_pipeline_run((td, tc), (sd, sc), W, H, cm.fx, cm.fy, cm.cx, cm.cy, levels, 30, conv=(0.0, 0.0),
                                   threads=th, warmup=0, reps=1) for th in (1, 4)}
    camo = O.make_camera(W, H, cm.fx, cm.fy, cm.cx, cm.cy)
    bo = O.Bilateral(1, 2, 2.0, 30.0)
    tx, tr, _, _ = O.cloud_from_rgbd(td, tc, camo, bo, 0.2, 3.5)
    sx, sr, _, _ = O.cloud_from_rgbd(sd, sc, camo, bo, 0.2, 3.5)
    lo = [O.LevelParams(v, 2.5 * v, 30, 2 * v, 0, O.icp_params(30, 2.5 * v, 0.968, 0.0, 0.0)) for v in levels]
    To, _ = O.multiscale_icp(sx, sr, tx, tr, lo)
    for th, r in runs.items():
        assert r["threads"] == th and r["iterations"] == [30, 30, 30]
        assert r["points"][0] == len(sx)  # same valid pixels as the oracle's front end
        rot, tr_ = _pose_err(r["T"], To)
        assert rot < 2e-3 and tr_ < 2e-3, (th, rot, tr_)
        moved, _ = _pose_err(np.linalg.inv(G), np.eye(4))
        assert rot < 0.05 * moved
    d = _pose_err(runs[1]["T"], runs[4]["T"])
    assert d[0] < 1e-5 and d[1] < 1e-5
