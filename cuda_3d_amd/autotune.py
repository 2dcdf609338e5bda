"""Launch-shape autotuner for libgca.so (no reference counterpart: the reference takes the
shapes Thrust picks).

The knobs are the ctx tuning keys (GCA_TUNE_*): lanes per neighbourhood query for normals and
colour gradients, lanes per correspondence query of the ICP by cloud size (the size of a query's
lane group within the 256-thread workgroup), the cap on ICP work units per iteration (one FP64
partial each, one workgroup each in a resident launch), the workgroup count of the persistent ICP
launch, and the neighbour-grid cell edge (in quarters of the search radius).  None of them
changes a result beyond the last bits of the ICP's FP64 partial sums (neighbour sets, moments,
voxel and grid output are exact whatever the shape or cell), so the tuner only measures time.

Method: coordinate descent.  Starting from the current values, each knob in turn is swept over
its candidates with the others held, and the value with the lowest median time of the stages
it drives (ctx stage profiler: hipEvents on the library's stream) is kept; `rounds` passes.  The
overlapped frame workload is scored by wall time per frame instead: its steps run side by side
(two lanes, graph replay), so the device time a shape takes from the other frames in flight
counts, not one frame's stage latency (fewer lanes per query use fewer workgroup slots).

    python -m cuda_3d_amd.autotune --workload frame --out cuda_3d_amd/tuning_mi355x.json  # on the GPU
    python -m cuda_3d_amd.autotune --workload frame_overlapped --out cuda_3d_amd/tuning_mi355x.json
    python -m cuda_3d_amd.autotune --workload batch --out cuda_3d_amd/tuning_mi355x.json

each writes its workload's entry {"tuning": {...}, "workload": ..., "measurements": [[round, knob,
value, ms], ...]} into the table (other workloads' entries are kept); `apply(ctx, path, workload)`
loads one.
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
        "icp_lanes_batch": [1, 2, 4, 8],
        "grid_cell": [4, 5, 6, 8],
    }


STAGES_OF = {
    "normals_lanes": ("normals", "gradient"),
    "icp_lanes_small": ("icp",),
    "icp_lanes_medium": ("icp",),
    "icp_lanes_large": ("icp",),
    "icp_max_units": ("icp",),
    "icp_persist_blocks": ("icp",),
    "icp_lanes_batch": ("icp",),
    "grid_cell": ("grid", "normals", "gradient", "icp"),
}
# knobs a workload does not exercise (one-pair calls never read the batch lanes, the batch
# never reads the one-pair lanes or the normals shape)
UNUSED_BY = {
    "frame": ("icp_lanes_batch",),
    "frame_overlapped": ("icp_lanes_batch", "icp_persist_blocks"),
    "batch": ("normals_lanes", "icp_lanes_small", "icp_lanes_medium", "icp_lanes_large", "grid_cell"),
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


def measure_wall(ctx, workload: Callable[[], None], stages: Sequence[str], reps: int, warm: int = 2) -> float:
    """Median over `reps` runs of the wall time (ms) of one workload run, after `warm` unmeasured
    runs (a tuning change drops a pipeline's graphs: the first steps after it are eager)."""
    for _ in range(warm):
        workload()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        workload()
        times.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(times)


def tune(ctx, workload: Callable[[], None], knobs: Dict[str, List[int]], rounds: int = 2, reps: int = 5,
         log: Callable[[str], None] = print, score=measure) -> tuple:
    """Coordinate descent; returns (best settings, list of measurements).  score(ctx, workload,
    stages, reps) -> ms is the quantity minimised."""
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
                t = score(ctx, workload, stages, reps)
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


def load(path: str = DEFAULT_TABLE, workload: str = "frame") -> dict:
    """The saved settings of one workload ({} when the table or the entry is missing)."""
    from . import _lib as L

    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        entry = json.load(fh).get(workload)
    if not entry:
        return {}
    return {k: int(v) for k, v in entry["tuning"].items() if k in L.TUNE_KEYS}


def apply(ctx, path: str = DEFAULT_TABLE, workload: str = "frame") -> dict:
    """Apply a saved table's settings for `workload` to a ctx (keys the library does not know are
    ignored)."""
    settings = load(path, workload)
    if settings:
        ctx.set_tuning(**settings)
    return settings


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--steps", type=int, default=6, help="frames per measured run")
    ap.add_argument("--knobs", default="", help="comma-separated subset of knobs")
    ap.add_argument("--workload", choices=("frame", "frame_overlapped", "batch"), default="frame",
                    help="frame: 640x480 odometry step (stage times); frame_overlapped: the same in overlapped "
                         "steps (wall time per frame); batch: 64 uneven pairs x 3 levels, one launch")
    args = ap.parse_args(argv)

    from . import core as g
    from . import workloads

    ctx = g.Context(0)
    import torch  # noqa: E402  (device properties)

    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    knobs = space(n_cu)
    knobs = {k: v for k, v in knobs.items() if k not in UNUSED_BY[args.workload]}
    if args.knobs:
        keep = set(args.knobs.split(","))
        knobs = {k: v for k, v in knobs.items() if k in keep}
    say = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    score = measure
    if args.workload == "frame":
        wl = workloads.frame_workload(ctx, steps=args.steps)
        ctx = wl.ctx  # the pipeline runs (and is tuned) on its own ctx
        what = (f"640x480 frame pipeline (3 levels, normals K30, 3-level colored ICP), {args.steps} frames "
                f"per run, median of {args.reps}")
    elif args.workload == "frame_overlapped":
        steps = max(args.steps, 24)
        wl = workloads.frame_workload(ctx, steps=steps, overlap=True)
        ctx = wl.ctx
        score = measure_wall
        what = (f"640x480 frame pipeline in overlapped steps (two lanes, graph replay), wall time of {steps} "
                f"frames per run, median of {args.reps}")
    else:
        bp = workloads.BatchPairs(ctx, log=say)

        def wl():
            bp.batched()
            ctx.synchronize()
        what = f"64 uneven pairs x 3 levels in one persistent launch, median of {args.reps}"
    t0 = time.perf_counter()
    best, rows = tune(ctx, wl, knobs, rounds=args.rounds, reps=args.reps, log=say, score=score)
    table = {}
    if os.path.exists(args.out):
        with open(args.out) as fh:
            table = json.load(fh)
    table[args.workload] = {
        "tuning": best, "device": torch.cuda.get_device_name(0), "compute_units": n_cu, "workload": what,
        "seconds": round(time.perf_counter() - t0, 1),
        "measurements": [[r["round"], r["knob"], r["value"], round(r["ms"], 4)] for r in rows]}
    with open(args.out, "w") as fh:
        json.dump(table, fh, indent=1)
    print(json.dumps({"tuning": best}), flush=True)


if __name__ == "__main__":
    main()
