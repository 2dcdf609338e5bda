#!/usr/bin/env python3
"""Benchmark: 640x480 depth -> normals -> 3-level colored ICP, ms/frame on one MI355X.

A step is one new frame through FramePipeline: front end with 5x5 bilateral, voxel 0.04 / 0.02
/ 0.01 with normals (r 2.5 v, K 30) and colour gradients (r 2 v) for each level, then 3-level
colored ICP against the previous frame, at most 30 Gauss-Newton iterations per level with the
reference's fixed stopping rule (color_icp.hpp:45-47: cos(theta_step) >= 1 - 1e-6 and
|t_step|^2 <= 1e-6).  The two frames of a seeded synthetic scene alternate, so every step
registers a real frame pair.  Depth and colour are resident in HBM before the timed region.

The JSON line also carries:
  roofline      the dominant kernel (the resident ICP launch, one per pyramid level),
                algorithmic bytes per launch = iterations x (24 M_src + 48 M_tgt) (source
                x,y,z,I read + previous match read and written; target x,y,z,I, normal and
                gradient read once) over its average launch time, from hipEvents on the
                launching stream in a second, identical pass;
  cpu_baseline  the same frame step as a multithreaded FP32 C++ program (oracle/cpu_pipeline.cpp)
                on the host cores: median of 20 steps after 3 warm-ups, at all threads (value)
                and at one thread, with the core count and CPU model;
  fixed_iterations  the same workload with the stop thresholds at 0: exactly 30 Gauss-Newton
                iterations per level;
  sequence      8 distinct frames of one scene (poses within the motion bounds of each other)
                stepped in order, each registered onto the one before;
  h2d           host -> device upload of one frame (depth + BGR), pinned and pageable, and the
                PCIe-inclusive frame time (value + pageable upload: the reference uploads
                pageable frames synchronously before each step);
  batch         64 uneven pairs x 3 levels in one persistent launch vs per-pair, per-level
                plain calls, same stopping rule.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per ICP launch measured with rocprofv3 --pmc (committed summary)."""
    path = os.path.join(ROOT, "profiles", "icp_resident_traffic_640x480.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(frames, cm, levels, iters, reps=20, warmup=3):
    """The frame step (front end, levels, 3-level ICP onto the previous frame) of the multithreaded
    FP32 CPU pipeline: median ms at all threads and at one thread."""
    import numpy as np

    import oracle as O

    W, H = cm.width, cm.height
    out = {}
    for key, th in (("all", 0), ("one", 1)):
        r = O.cpu_pipeline_run(frames[0], frames[1], W, H, cm.fx, cm.fy, cm.cx, cm.cy, levels, iters,
                               threads=th, warmup=warmup, reps=reps)
        out[key] = (float(np.median(r["ms"])), r["threads"], r["iterations"])
    return out


def h2d_upload(td, tc, reps=20):
    """One frame's upload (uint16 depth + BGR8) with synchronous copies: from pinned host memory
    and from pageable numpy arrays.  Median ms of each."""
    import numpy as np
    import torch

    d_dev = torch.empty(td.shape, dtype=torch.int16, device="cuda")
    c_dev = torch.empty(tc.shape, dtype=torch.uint8, device="cuda")
    d_pin = torch.from_numpy(td.view(np.int16)).pin_memory()
    c_pin = torch.from_numpy(tc).pin_memory()
    d_pag = torch.from_numpy(td.view(np.int16).copy())
    c_pag = torch.from_numpy(tc.copy())
    res = {}
    for name, (d, c) in (("pinned", (d_pin, c_pin)), ("pageable", (d_pag, c_pag))):
        ts = []
        for k in range(reps + 2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d_dev.copy_(d)
            c_dev.copy_(c)
            torch.cuda.synchronize()
            if k >= 2:
                ts.append((time.perf_counter() - t0) * 1e3)
        res[name] = float(np.median(ts))
    res["bytes"] = int(td.nbytes + tc.nbytes)
    return res


def level_iteration_us(reports):
    """Per level: device time of one Gauss-Newton iteration (active time / iterations)."""
    return [r.active_us / r.iterations if r.iterations else 0.0 for r in reports]


def level_timeline_us(reports):
    """Per level: (start, finish) of its registration in device-clock us after the step's
    registration set-up (which follows the front end)."""
    return [(r.latency_us - r.active_us, r.latency_us) for r in reports]


def run_fixed(g, ctx, cam, bil, ptrs, levels, iters, steps, graph, overlap):
    """The headline workload with the stop thresholds at 0 (every level runs `iters`), in the
    headline's mode and sequentially."""
    import numpy as np

    from cuda_3d_amd import autotune
    from cuda_3d_amd.pipeline import FramePipeline

    out = {}
    for ov in ((True, False) if overlap else (False,)):
        pipe = FramePipeline(ctx, cam, bil, levels, iters, conv_rotation=0.0, conv_translation_sq=0.0, graph=graph,
                             overlap=ov)
        autotune.apply(pipe.ctx, workload="frame_overlapped" if ov else "frame")
        for k in range(12):  # eager steps of every frame slot, then graph capture
            pipe.step(*ptrs[k % 2])
        pipe.ctx.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            pipe.step(*ptrs[k % 2])
        pipe.ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out["ms_per_frame" if ov or not overlap else "sequential_ms_per_frame"] = ms
        if not ov:  # per-iteration device times from a sequential pipeline's reports
            per_it = []
            for k in range(6):
                pipe.step(*ptrs[k % 2])
                per_it.append(level_iteration_us(pipe.result()[1]))
            out["iterations_per_level"] = pipe.iterations()
            out["icp_us_per_iteration_per_level"] = np.median(np.array(per_it), axis=0).tolist()
        pipe.close()
    out["steps"] = steps
    return out


def run_sequence(g, ctx, cam, bil, levels, iters, W, H, n_frames, passes, graph, overlap):
    """n_frames distinct frames stepped in order (each onto the one before), `passes` times: the
    pose error against the generating poses from a sequential pass with a readback per frame, then
    the timed passes in the headline's mode."""
    import numpy as np

    from cuda_3d_amd import autotune, scene
    from cuda_3d_amd.pipeline import FramePipeline

    frames, poses, _ = scene.frame_sequence(31, n_frames, W, H)
    dev = [(g.device_tensor(d.view(np.int16)), g.device_tensor(c)) for d, c in frames]
    ptrs = [(d.data_ptr(), c.data_ptr()) for d, c in dev]
    errs, its = [], []
    seq = FramePipeline(ctx, cam, bil, levels, iters, graph=graph)
    autotune.apply(seq.ctx, workload="frame")
    for k in range(n_frames + 1):
        seq.step(*ptrs[k % n_frames])
        T, reps = seq.result()
        if T is None:
            continue
        a, b = poses[(k - 1) % n_frames], poses[k % n_frames]
        truth = np.linalg.inv(a) @ b
        dR = T[:3, :3] @ truth[:3, :3].T
        errs.append(float(np.arccos(np.clip((np.trace(dR) - 1) / 2, -1, 1))))
        its.append([r.iterations for r in reps])
    seq.close()
    pipe = FramePipeline(ctx, cam, bil, levels, iters, graph=graph, overlap=overlap)
    autotune.apply(pipe.ctx, workload="frame_overlapped" if overlap else "frame")
    for k in range(max(n_frames + 1, 12)):  # eager steps of every slot, then graph capture
        pipe.step(*ptrs[k % n_frames])
    pipe.ctx.synchronize()
    t0 = time.perf_counter()
    steps = n_frames * passes
    for k in range(steps):
        pipe.step(*ptrs[k % n_frames])
    pipe.ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    captures, updates = pipe.graph_stats()
    pipe.close()
    return {"frames": n_frames, "steps": steps, "ms_per_frame": ms, "overlap": overlap, "graph_captures": captures,
            "graph_node_updates": updates, "iterations_per_level": its,
            "rotation_err_vs_truth_rad_median": float(np.median(errs)), "rotation_err_vs_truth_rad_max": float(max(errs))}


def run_batch(g, ctx, n_pairs, levels, iters, seed=2026):
    import numpy as np

    from cuda_3d_amd.workloads import BatchPairs

    bp = BatchPairs(ctx, n_pairs, levels, iters, seed, log=log)
    bp.batched()
    bp.plain(ctx)  # warm both paths
    ctx.synchronize()
    reps = 3
    tb, blat = [], None
    for _ in range(reps):
        t1 = time.perf_counter()
        Tb, breps = bp.batched()
        ctx.synchronize()
        tb.append(time.perf_counter() - t1)
        blat = [row[-1].latency_us * 1e-3 for row in breps]
    # the same batch as a plan: set up once, its run captured as a graph and replayed (one
    # graph launch and the results read per run)
    plan = g.BatchPlan(bp.src, bp.tgt, bp.params, bp.T0, bp.grad_r, graph=True)
    plan.enqueue()
    plan.results()  # capture + warm
    tg = []
    for _ in range(reps):
        t1 = time.perf_counter()
        plan.enqueue()
        Tg, _ = plan.results()
        tg.append(time.perf_counter() - t1)
    plan.close()
    tp = []
    for _ in range(reps):
        t1 = time.perf_counter()
        Tp, lat = bp.plain(ctx)
        ctx.synchronize()
        tp.append(time.perf_counter() - t1)
    tb_s, tp_s, tg_s = min(tb), min(tp), min(tg)
    diff = float(np.abs(Tb - Tp).max())
    lat = np.array(lat) * 1e3
    sizes = bp.sizes
    iters_run = [[r.iterations for r in row] for row in breps]
    # algorithmic bytes of the batched launch: per pair, level and iteration 24 B per source
    # point (x, y, z, I read; previous match read and written) + 48 B per target point
    # (x, y, z, I, normal, gradient read once), as for the frame's launches
    algo = sum(iters_run[k][l] * (24.0 * sizes[l][k] + 48.0 * bp.tgt_sizes[l][k])
               for k in range(n_pairs) for l in range(len(levels)))
    achieved = algo / tb_s / 1e9
    # registration quality against the generating transforms (not a parity bar)
    truth_err = []
    for k, G in enumerate(bp.truth):
        dR = Tb[k][:3, :3] @ G[:3, :3].T
        truth_err.append(float(np.arccos(np.clip((np.trace(dR) - 1) / 2, -1, 1))))
    return {
        "pairs": n_pairs, "points_sampled_per_side": int(bp.counts.sum()),
        "points_sampled_min": int(bp.counts.min()), "points_sampled_max": int(bp.counts.max()),
        "points_per_level_min": [int(min(x)) for x in sizes], "points_per_level_max": [int(max(x)) for x in sizes],
        "points_per_level_total": [int(sum(x)) for x in sizes],
        "levels": list(levels), "max_iters_per_level": iters, "stopping_rule": "reference (1e-6, 1e-6)",
        "iterations_total": int(sum(map(sum, iters_run))),
        "batched_pairs_per_s": n_pairs / tb_s, "plain_pairs_per_s": n_pairs / tp_s,
        "speedup": tp_s / tb_s, "batched_ms": tb_s * 1e3, "plain_ms": tp_s * 1e3,
        "batched_latency_ms_p50": float(np.percentile(blat, 50)), "batched_latency_ms_p99": float(np.percentile(blat, 99)),
        "plain_latency_ms_p50": float(np.percentile(lat, 50)), "plain_latency_ms_p99": float(np.percentile(lat, 99)),
        "graph_plan_pairs_per_s": n_pairs / tg_s, "graph_plan_ms": tg_s * 1e3, "graph_plan_speedup": tp_s / tg_s,
        "max_abs_pose_diff_batched_vs_plain": diff,
        "max_abs_pose_diff_graph_plan_vs_batched": float(np.abs(Tg - Tb).max()),
        "batched_roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": algo},
        "rotation_err_vs_truth_rad_median": float(np.median(truth_err)),
    }


def copy_calibration(nbytes: int, reps: int = 5):
    """Plain device copies of known size: what the HBM counters read per copy is the scale for
    the ICP kernel's counter readings."""
    import torch

    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes": nbytes, "copies": reps, "ms_per_copy": dt * 1e3,
                      "GB_per_s": 2 * nbytes / dt / 1e9}), flush=True)


def run_clustering(g, ctx, depth, bgr, cam, bil, tolerance=0.01, calls=20):
    """Euclidean clustering (point_cloud::euclidean_clustering) of the full-resolution target
    frame cloud: median host wall time of gca_euclidean_clustering over ``calls`` calls, the
    cloud's grid at the tolerance already built (the reference's call also rebuilds its radius
    lists each time; here the grid persists on the cloud)."""
    import numpy as np

    pc = g.PointCloud.create_from_rgbd(depth, bgr, cam, bil, 0.2, 3.5, ctx=ctx)
    pc.set_normals(np.tile([0.0, 0.0, 1.0], (pc.points_number(), 1)))  # the call requires normals
    r = pc.euclidean_cluster_labels(tolerance, 20, 10 ** 7)
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        r = pc.euclidean_cluster_labels(tolerance, 20, 10 ** 7)
        ts.append(time.perf_counter() - t0)
    return {"points": pc.points_number(), "tolerance": tolerance, "min_cluster_size": 20,
            "clusters": r[1], "ms_per_call": float(np.median(ts) * 1e3), "calls": calls}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-pairs", type=int, default=64)
    ap.add_argument("--no-clustering", action="store_true")
    ap.add_argument("--no-fixed", action="store_true", help="skip the fixed-iterations run")
    ap.add_argument("--no-sequence", action="store_true", help="skip the 8-frame sequence run")
    ap.add_argument("--no-roofline-pass", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch every step eagerly (no hipGraph replay)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="headline in sequential steps (default: each step registers the previous frame beside "
                         "the new frame's build)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (reference traffic for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    import numpy as np
    import torch

    import cuda_3d_amd as g
    from cuda_3d_amd import autotune, scene
    from cuda_3d_amd.pipeline import FramePipeline

    W, H, levels, iters = 640, 480, (0.04, 0.02, 0.01), 30
    (td, tc), (sd, sc), G, cm = scene.frame_pair(12, W, H)
    ctx = g.Context(0)
    cam = g.make_camera(W, H, cm.fx, cm.fy, cm.cx, cm.cy)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    dev = [(g.device_tensor(td.view(np.int16)), g.device_tensor(tc)), (g.device_tensor(sd.view(np.int16)), g.device_tensor(sc))]
    ptrs = [(d.data_ptr(), c.data_ptr()) for d, c in dev]
    overlap = not args.no_overlap
    # the headline pipeline, and a sequential one for the per-frame latency, the registration
    # timeline and the roofline pass (which read each frame's own registration after its step)
    pipe = FramePipeline(ctx, cam, bil, levels, iters, graph=not args.no_graph, overlap=overlap)
    seqp = FramePipeline(ctx, cam, bil, levels, iters, graph=not args.no_graph)
    pctx = seqp.ctx  # the sequential pipeline's own stream, profiler and tuning
    # the launch shapes autotuned on MI355X (cuda_3d_amd/tuning_mi355x.json), per workload
    tuning = {"table": os.path.relpath(autotune.DEFAULT_TABLE, ROOT),
              "frame": autotune.apply(pctx, workload="frame"), "batch": autotune.apply(ctx, workload="batch")}
    if overlap:  # overlapped steps have their own shapes (tuned on throughput, not one frame's stages)
        tuning["frame_overlapped"] = autotune.apply(pipe.ctx, workload="frame_overlapped")
    else:
        autotune.apply(pipe.ctx, workload="frame")

    counters = {}

    def step(p):
        i = counters.get(id(p), 0)
        d, c = ptrs[i % 2]
        p.step(d, c)
        counters[id(p)] = i + 1

    def timed(p, n):
        p.ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            step(p)
        p.ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    for _ in range(max(args.warmup, 12)):  # eager steps of every frame slot, then graph capture
        step(pipe)
        step(seqp)
    log(f"warmup done; timing {args.steps} steps ({'graph replay' if not args.no_graph else 'eager'}, "
        f"{'overlapped' if overlap else 'sequential'})")
    ms = timed(pipe, args.steps)
    log(f"{ms:.3f} ms/frame")
    seq_ms = timed(seqp, args.steps) if overlap else ms
    if overlap:
        log(f"{seq_ms:.3f} ms/frame in sequential steps")
    # latency of one frame when the host waits for each pose (sequential steps)
    lat, per_it, tl = [], [], []
    for _ in range(min(args.steps, 20)):
        t2 = time.perf_counter()
        step(seqp)
        _, reps = seqp.result()
        lat.append((time.perf_counter() - t2) * 1e3)
        per_it.append(level_iteration_us(reps))
        tl.append(level_timeline_us(reps))
    lat_ms = float(np.median(lat))
    it_us = np.median(np.array(per_it), axis=0).tolist()
    timeline = np.median(np.array(tl), axis=0).round(1).tolist()
    log(f"registration start / finish by level (us after its set-up): {timeline}")
    log(f"{lat_ms:.3f} ms per frame with a pose readback after each frame; "
        f"ICP us per iteration by level {[round(x, 1) for x in it_us]}")

    sizes = [seqp.level_sizes(s) for s in (0, 1)]
    full_n = [seqp.full_size(s) for s in (0, 1)]
    roof = None
    stages = None
    iters_seen = []
    if not args.no_roofline_pass:
        # the same kernels launched eagerly, each stage bracketed by hipEvents on the stream
        seqp.set_graph(False)
        pctx.profile(True)
        algo = 0.0
        for _ in range(args.steps):
            new = seqp.slot
            step(seqp)
            it = seqp.iterations()
            iters_seen.append(it)
            src, tgt = new, 1 - new
            algo += sum(it[l] * (24.0 * sizes[src][l] + 48.0 * sizes[tgt][l]) for l in range(len(levels)))
        prof = pctx.profile_read()
        pctx.profile(False)
        seqp.set_graph(not args.no_graph)
        n_icp, icp_ms = prof["icp"]
        per_launch_bytes = algo / max(n_icp, 1)
        avg_launch_ms = icp_ms / max(n_icp, 1)
        achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_icp_resident", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                "algorithmic_bytes_per_launch": per_launch_bytes, "avg_launch_us": avg_launch_ms * 1e3,
                "launches": n_icp}
        stages = {s: {"launches_per_frame": n / args.steps, "ms_per_frame": t / args.steps} for s, (n, t) in prof.items()}
        log(f"roofline pass: icp {n_icp} launches avg {avg_launch_ms * 1e3:.2f} us, {achieved:.1f} GB/s; "
            f"iterations per level {iters_seen[-2:]}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP32 C++ pipeline, OpenMP) on the same frame step ...")
        cb = cpu_baseline([(td, tc), (sd, sc)], cm, levels, iters)
        cpu = {"value": cb["all"][0], "unit": "ms/frame", "cores": cb["all"][1], "kind": "port",
               "value_1_thread": cb["one"][0], "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "iterations_per_level": cb["all"][2],
               "sample": "the benchmark's frame step (640x480 front end, 3 levels of voxel / normals / gradient, "
                         "3-level colored ICP onto the previous frame, reference stopping rule) as a multithreaded "
                         "FP32 C++ program (oracle/cpu_pipeline.cpp): median of 20 steps after 3 warm-ups"}
        log(f"cpu baseline {cb['all'][0]:.1f} ms/frame at {cb['all'][1]} threads, {cb['one'][0]:.1f} ms at 1 thread")

    fixed = None
    if not args.no_fixed:
        fixed = run_fixed(g, ctx, cam, bil, ptrs, levels, iters, max(10, args.steps // 4), not args.no_graph, overlap)
        log(f"fixed {iters} iterations per level: {fixed['ms_per_frame']:.3f} ms/frame"
            + (f" ({fixed['sequential_ms_per_frame']:.3f} sequential)" if overlap else "")
            + f"; ICP us per iteration by level {[round(x, 1) for x in fixed['icp_us_per_iteration_per_level']]}")

    seq = None
    if not args.no_sequence:
        seq = run_sequence(g, ctx, cam, bil, levels, iters, W, H, 8, max(2, args.steps // 8), not args.no_graph,
                           overlap)
        log(f"sequence of {seq['frames']} distinct frames: {seq['ms_per_frame']:.3f} ms/frame "
            f"({seq['graph_captures']} captures, {seq['graph_node_updates']} node updates); rotation error vs "
            f"truth median {seq['rotation_err_vs_truth_rad_median']:.2e} rad")

    h2d = h2d_upload(td, tc)
    h2d["pcie_inclusive_ms_per_frame"] = ms + h2d["pageable"]
    log(f"upload of one frame ({h2d['bytes']} B): {h2d['pinned']:.3f} ms pinned, {h2d['pageable']:.3f} ms pageable")

    batch = None
    if not args.no_batch and args.batch_pairs > 0:
        batch = run_batch(g, ctx, args.batch_pairs, levels, iters)
        log(f"batch: {batch['batched_pairs_per_s']:.1f} pairs/s batched vs {batch['plain_pairs_per_s']:.1f} plain "
            f"({batch['speedup']:.2f}x); graph-replayed plan {batch['graph_plan_pairs_per_s']:.1f} pairs/s "
            f"({batch['graph_plan_speedup']:.2f}x); batched launch {batch['batched_roofline']['achieved']:.0f} GB/s "
            f"algorithmic")

    clustering = None
    if not args.no_clustering:
        clustering = run_clustering(g, ctx, td, tc, cam, bil)
        log(f"clustering: {clustering['points']} points, tolerance {clustering['tolerance']} m: "
            f"{clustering['clusters']} clusters, {clustering['ms_per_call']:.3f} ms per call (host wall)")

    out = {
        "metric": "ms/frame 640x480 depth->ICP; batched ICP pairs/s vs plain at equal pose error",
        "value": ms, "unit": "ms/frame", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic seeded RGB-D scene (3 planes + 4 spheres, 2 mm noise, 5% holes), frames resident in HBM",
        "config": {"workload": "640x480 depth+RGB -> bilateral 5x5 -> voxel 0.04/0.02/0.01 -> normals r=2.5v K=30 "
                               "+ colour gradient r=2v -> 3-level colored ICP, max 30 it/level, reference "
                               "stopping rule" + ("; overlapped steps: frame k built while frame k-1 registers"
                                                  if overlap else ""),
                   "width": W, "height": H, "levels": list(levels), "max_iters_per_level": iters,
                   "iterations_per_level": iters_seen[-2:] if iters_seen else None,
                   "hip_graph": not args.no_graph, "overlapped_steps": overlap,
                   "sequential_ms_per_frame": seq_ms, "latency_ms_per_frame_with_readback": lat_ms,
                   "icp_us_per_iteration_per_level": it_us,
                   "icp_start_finish_us_per_level": timeline,
                   "points_full": full_n, "points_per_level": sizes, "tuning": tuning},
        "roofline": roof, "cpu_baseline": cpu, "fixed_iterations": fixed, "sequence": seq, "h2d": h2d,
        "batch": batch, "clustering": clustering, "stages": stages,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
