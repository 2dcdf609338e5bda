"""Launch-shape autotuner for libgca.so (no reference counterpart: the reference takes the
shapes Thrust picks).

The knobs are the ctx tuning keys (GCA_TUNE_*): lanes per neighbourhood query for normals and
colour gradients, lanes per correspondence query of the ICP by cloud size, the cap on ICP work
units per iteration (one FP64 partial each) and the workgroup count of the persistent ICP
launch.  None of them changes a result beyond the last bits of the ICP's FP64 partial sums
(neighbour sets, moments, voxel and grid output are exact whatever the shape), so the tuner
only measures time.

Method: coordinate descent.  Starting from the current values, each knob in turn is swept over
its candidates with the others held, and the value with the lowest median time of the stages
it drives (ctx stage profiler: hipEvents on the library's stream) is kept; `rounds` passes.

    python -m cuda_3d_amd.autotune --out cuda_3d_amd/tuning_mi355x.json     # on the GPU

writes {"tuning": {...}, "workload": ..., "measurements": [...]}; `apply(ctx, path)` loads it.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from typing import Callable, Dict, List, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning_mi355x.json")


def space(n_cu: int) -> Dict[str, List[int]]:
    """Candidate values per knob, and the profiler stages each knob drives."""
    return {
        "normals_lanes": [2, 4, 8, 16, 32, 64],
        "icp_lanes_small": [1, 2, 4, 8, 16, 32],
        "icp_lanes_medium": [1, 2, 4, 8, 16],
        "icp_lanes_large": [1, 2, 4, 8, 16],
        "icp_max_units": [64, 128, 256, 512, 1024],
        "icp_persist_blocks": [n_cu // 2, n_cu, 2 * n_cu, 4 * n_cu, 8 * n_cu],
    }


STAGES_OF = {
    "normals_lanes": ("normals", "gradient"),
    "icp_lanes_small": ("icp",),
    "icp_lanes_medium": ("icp",),
    "icp_lanes_large": ("icp",),
    "icp_max_units": ("icp",),
    "icp_persist_blocks": ("icp",),
}


def measure(ctx, workload: Callable[[], None], stages: Sequence[str], reps: int) -> float:
    """Median over `reps` runs of the summed GPU time (ms) of `stages` in one workload run."""
    times = []
    for _ in range(reps):
        ctx.profile(True)
        workload()
        prof = ctx.profile_read()
        ctx.profile(False)
        times.append(sum(prof[s][1] for s in stages))
    return statistics.median(times)


def tune(ctx, workload: Callable[[], None], knobs: Dict[str, List[int]], rounds: int = 2, reps: int = 5,
         log: Callable[[str], None] = print) -> tuple:
    """Coordinate descent; returns (best settings, list of measurements)."""
    best = {k: v for k, v in ctx.get_tuning().items() if k in knobs}
    rows = []
    workload()  # warm
    for rnd in range(rounds):
        changed = False
        for knob, cands in knobs.items():
            stages = STAGES_OF[knob]
            scores = {}
            for v in cands:
                ctx.set_tuning(**{knob: v})
                t = measure(ctx, workload, stages, reps)
                scores[v] = t
                rows.append({"round": rnd, "knob": knob, "value": v, "ms": t, "settings": dict(best, **{knob: v})})
            v_best = min(scores, key=scores.get)
            if best.get(knob) != v_best:
                changed = True
            best[knob] = v_best
            ctx.set_tuning(**{knob: v_best})
            log(f"[autotune] round {rnd} {knob}: " + " ".join(f"{v}:{scores[v]:.3f}" for v in cands) +
                f" -> {v_best}")
        if not changed:
            break
    return best, rows


def apply(ctx, path: str = DEFAULT_TABLE) -> dict:
    """Apply a saved table to a ctx (keys unknown to this build are ignored)."""
    from . import _lib as L

    with open(path) as fh:
        table = json.load(fh)["tuning"]
    settings = {k: int(v) for k, v in table.items() if k in L.TUNE_KEYS}
    ctx.set_tuning(**settings)
    return settings


def frame_workload(ctx, steps: int = 6, W: int = 640, H: int = 480, seed: int = 12):
    """The 640x480 frame pipeline (front end, 3 voxel levels with normals and gradients,
    3-level colored ICP) over a seeded frame pair, `steps` frames per run."""
    import numpy as np

    from . import core as g
    from . import scene
    from ._lib import BILATERAL_FIXED
    from .pipeline import FramePipeline

    (td, tc), (sd, sc), _, cm = scene.frame_pair(seed, W, H)
    cam = g.make_camera(W, H, cm.fx, cm.fy, cm.cx, cm.cy)
    bil = g.bilateral(BILATERAL_FIXED, 2, 2.0, 30.0)
    dev = [(g.device_tensor(td.view(np.int16)), g.device_tensor(tc)),
           (g.device_tensor(sd.view(np.int16)), g.device_tensor(sc))]
    pipe = FramePipeline(ctx, cam, bil)
    state = {"k": 0}

    def run():
        for _ in range(steps):
            d, c = dev[state["k"] % 2]
            pipe.step(d.data_ptr(), c.data_ptr())
            state["k"] += 1
        ctx.synchronize()

    run._keep = (dev, pipe)  # keep device buffers alive with the closure
    return run


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--steps", type=int, default=6, help="frames per measured run")
    ap.add_argument("--knobs", default="", help="comma-separated subset of knobs")
    args = ap.parse_args(argv)

    from . import core as g

    ctx = g.Context(0)
    import torch  # noqa: E402  (device properties)

    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    knobs = space(n_cu)
    if args.knobs:
        keep = set(args.knobs.split(","))
        knobs = {k: v for k, v in knobs.items() if k in keep}
    wl = frame_workload(ctx, steps=args.steps)
    t0 = time.perf_counter()
    best, rows = tune(ctx, wl, knobs, rounds=args.rounds, reps=args.reps,
                      log=lambda m: print(m, file=sys.stderr, flush=True))
    out = {"tuning": best, "device": torch.cuda.get_device_name(0), "compute_units": n_cu,
           "workload": "640x480 frame pipeline (3 levels, normals K30, 3-level colored ICP), "
                       f"{args.steps} frames per run, median of {args.reps}",
           "seconds": time.perf_counter() - t0, "measurements": rows}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps({"tuning": best}), flush=True)


if __name__ == "__main__":
    main()
