"""GPU parity of the benchmark workloads at their stated sizes, against the FP64 oracle.

The 64-pair batch: 64 uneven surface-sample pairs (log-uniform 5e3..2e5 points per side, seed
2026, each with its own rigid motion of up to 15 deg / 0.2 m), voxel pyramid 0.04 / 0.02 / 0.01 m,
normals r = 2.5 v (K <= 30), colour gradient r = 2 v, colored ICP up to 30 iterations per level
with the reference's stopping rule.  All 64 pairs x 3 levels run in one persistent launch
(gca_color_icp_batch_multiscale); each pose must be within 1e-5 rad / 1e-5 m of the oracle's
coarse-to-fine registration of the same pair, which follows color_icp::align
(color_icp.cu:151-234) per level.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402


def pose_err(T, U):
    dR = T[:3, :3] @ U[:3, :3].T
    return (float(np.arctan2(np.linalg.norm([dR[2, 1] - dR[1, 2], dR[0, 2] - dR[2, 0], dR[1, 0] - dR[0, 1]]),
                             np.trace(dR) - 1)), float(np.linalg.norm(T[:3, 3] - U[:3, 3])))


def host_f4(xyz, rgb):
    """xyzi / rgb float4 rows as gca_cloud_upload stores them (intensity as color3::to_intensity)."""
    c = rgb.astype(np.float64)
    inten = (0.2126 * c[:, 0] + 0.7152 * c[:, 1] + 0.0722 * c[:, 2]).astype(np.float32)
    z = np.zeros((len(xyz), 1), np.float32)
    return (np.concatenate([xyz.astype(np.float32), inten[:, None]], 1),
            np.concatenate([rgb.astype(np.float32), z], 1))


def oracle_pair(host, levels, iters, conv, lam):
    sx, sr, tx, tr = host
    a, b = host_f4(sx, sr)
    c, d = host_f4(tx, tr)
    lo = [O.LevelParams(v, 2.5 * v, 30, 2.0 * v, 0, O.icp_params(iters, 2.5 * v, lam, conv[0], conv[1]))
          for v in levels]
    T, reps = O.multiscale_icp(a, b, c, d, lo)
    return T, [r.iterations for r in reps]


@pytest.fixture(scope="module")
def batch64(ctx):
    from cuda_3d_amd.workloads import BatchPairs

    print("[gpu-workloads] batch64: building 64 pairs (seed 2026)", flush=True)
    return BatchPairs(ctx, 64, seed=2026, keep_host=True)


def test_batch64_matches_oracle(batch64):
    from cuda_3d_amd.workloads import LAMBDA

    bp = batch64
    T, reps = bp.batched()
    print("[gpu-workloads] batch64: batched launch done; oracle per pair", flush=True)
    O.lib()
    workers = min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(workers) as ex:  # ctypes releases the GIL inside the oracle
        res = list(ex.map(lambda h: oracle_pair(h, bp.levels, bp.iters, bp.conv, LAMBDA), bp.host))
    errs = np.array([pose_err(T[k], res[k][0]) for k in range(64)])
    its_gpu = [[r.iterations for r in reps[k]] for k in range(64)]
    its_orc = [res[k][1] for k in range(64)]
    worst = int(np.argmax(errs.max(1)))
    print(f"[gpu-workloads] batch64 vs oracle: worst {errs[:, 0].max():.2e} rad {errs[:, 1].max():.2e} m "
          f"(pair {worst}, {bp.sizes[-1][worst]} finest source points); median {np.median(errs[:, 0]):.2e} rad; "
          f"iterations equal in {sum(a == b for a, b in zip(its_gpu, its_orc))}/64 pairs", flush=True)
    assert errs[:, 0].max() <= 1e-5 and errs[:, 1].max() <= 1e-5
    assert its_gpu == its_orc


def test_batch64_plan_graph_matches_batched(batch64):
    """The 64-pair batch as a plan whose run is captured once as a graph and replayed: each run
    gives the one-shot batch call's poses and reports bit for bit (and so the oracle's, to the
    tolerance above)."""
    import cuda_3d_amd as g

    bp = batch64
    T, reps = bp.batched()
    plan = g.BatchPlan(bp.src, bp.tgt, bp.params, bp.T0, bp.grad_r, graph=True)
    key = lambda rs: [[(r.iterations, r.converged, r.n_correspondences, r.rmse) for r in row] for row in rs]
    for run in range(3):
        print(f"[gpu-workloads] batch64 plan: graph run {run}", flush=True)
        plan.enqueue()
        Tp, rp = plan.results()
        np.testing.assert_array_equal(Tp, T)
        assert key(rp) == key(reps)
    assert plan.captures == 1
    plan.close()
