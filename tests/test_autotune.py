"""Autotuner logic on the CPU: coordinate descent over a fake ctx whose stage times are a known
function of the launch shape."""
import json
import time

from cuda_3d_amd import autotune


class FakeCtx:
    """Stage times: normals minimal at lanes 16, ICP separable with minimum at lanes (4, 8, 2),
    256 units, batch lanes 2."""

    def __init__(self):
        self.t = {"normals_lanes": 0, "icp_lanes_small": 0, "icp_lanes_medium": 0, "icp_lanes_large": 0,
                  "icp_max_units": 0, "icp_persist_blocks": 0, "icp_launch": 0, "icp_lanes_batch": 0,
                  "frontend": 0, "grid_cell": 0, "voxel": 0}
        self.on = False
        self.calls = 0

    def get_tuning(self):
        return dict(self.t)

    def set_tuning(self, **kw):
        for k, v in kw.items():
            assert k in self.t
            self.t[k] = v

    def profile(self, on=True):
        self.on = on

    def profile_read(self):
        import math

        t = self.t
        lg = lambda v: math.log2(max(v, 1))  # noqa: E731
        normals = 1.0 + (lg(t["normals_lanes"]) - 4) ** 2
        icp = (3.0 + (lg(t["icp_lanes_small"]) - 2) ** 2 + (lg(t["icp_lanes_medium"]) - 3) ** 2
               + (lg(t["icp_lanes_large"]) - 1) ** 2 + abs(lg(t["icp_max_units"]) - 8) + 0.01 * t["icp_persist_blocks"]
               + (lg(t["icp_lanes_batch"]) - 1) ** 2)
        grid = 0.5 + 0.1 * (t["grid_cell"] - 5) ** 2  # cell edge minimal at 5 quarters of the radius
        return {"normals": (3, normals), "gradient": (3, 0.5), "icp": (1, icp), "grid": (3, grid)}


def test_coordinate_descent_finds_separable_minimum():
    ctx = FakeCtx()
    ran = []
    best, rows = autotune.tune(ctx, lambda: ran.append(1), autotune.space(256), rounds=3, reps=3, log=lambda m: None)
    assert best["normals_lanes"] == 16
    assert (best["icp_lanes_small"], best["icp_lanes_medium"], best["icp_lanes_large"]) == (4, 8, 2)
    assert best["icp_max_units"] == 256
    assert best["icp_persist_blocks"] == 128  # smallest candidate: time grows with it
    assert best["icp_lanes_batch"] == 2
    assert best["grid_cell"] == 5
    # the ctx is left at the best settings, and every measurement is recorded
    assert all(ctx.t[k] == v for k, v in best.items())
    assert len(rows) >= sum(len(v) for v in autotune.space(256).values())
    assert len(ran) == 1 + 3 * len(rows)  # one warm-up run, reps per measurement


def test_table_round_trip(tmp_path):
    ctx = FakeCtx()
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"frame": {"tuning": {"normals_lanes": 8, "icp_max_units": 128, "not_a_knob": 3}},
                                "batch": {"tuning": {"icp_lanes_batch": 2}}}))
    applied = autotune.apply(ctx, str(path))
    assert applied == {"normals_lanes": 8, "icp_max_units": 128}
    assert ctx.t["normals_lanes"] == 8 and ctx.t["icp_max_units"] == 128
    assert autotune.apply(ctx, str(path), "batch") == {"icp_lanes_batch": 2} and ctx.t["icp_lanes_batch"] == 2
    assert autotune.load(str(tmp_path / "missing.json")) == {}


def test_committed_table_is_loadable():
    """The MI355X table holds every workload's settings, every key a knob of this build."""
    from cuda_3d_amd import _lib as L

    for w in ("frame", "frame_overlapped", "batch"):
        s = autotune.load(autotune.DEFAULT_TABLE, w)
        assert s and all(k in L.TUNE_KEYS for k in s)


def test_wall_score_times_runs_after_warm_ones():
    """The overlapped workload's score: wall time of a run, after unmeasured warm runs (a tuning
    change drops a pipeline's graphs, so its first steps after it are eager)."""
    ran = []

    def run():
        ran.append(1)
        time.sleep(0.01 if len(ran) <= 2 else 0.001)

    ms = autotune.measure_wall(FakeCtx(), run, ("icp",), reps=3, warm=2)
    assert len(ran) == 5 and 0.5 <= ms < 8.0
