"""Pins of the FP64 oracle (CPU only).

The reference ships no tests or golden vectors for this path, so the oracle is pinned by
fixtures computed independently with numpy/scipy brute force (tests/golden/make_golden.py),
by closed forms and invariants, and by planted-mistake checks that a plausible error (sign,
index, dropped term) would break.
"""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))


def f4(xyz, w=None):
    w = np.zeros(len(xyz), np.float32) if w is None else np.asarray(w, np.float32)
    return np.concatenate([np.asarray(xyz, np.float32), w[:, None]], 1)


def angle(a, b, signed=False):
    """Angle between row vectors, atan2(|a x b|, a . b) in FP64 (robust near 0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = np.linalg.norm(np.cross(a, b), axis=-1)
    d = (a * b).sum(-1)
    return np.arctan2(c, d if signed else np.abs(d))


# ---- front end -----------------------------------------------------------------------------

def small_cam():
    return O.make_camera(16, 12, 20.0, 20.0, 8.0, 6.0)


def test_binding_structs_match_compiled_oracle():
    """The ctypes mirrors have the sizes the oracle was compiled with (a shorter mirror would let
    the oracle write past the caller's report arrays)."""
    import ctypes as C

    want = [C.sizeof(O.Camera), C.sizeof(O.Bilateral), C.sizeof(O.IcpParams), C.sizeof(O.IcpReport),
            C.sizeof(O.LevelParams), C.sizeof(O._Grid)]
    assert O.abi_sizes() == want


def test_bilateral_matches_direct_fp64_loop():
    d, bgr = GOLD["bil_depth"], GOLD["bil_bgr"]
    xyzi, rgb, pix, filt = O.cloud_from_rgbd(d, bgr, small_cam(), O.Bilateral(1, 2, 2.0, 30.0), 0.2, 3.5)
    ref = GOLD["bil_filtered"].reshape(-1)
    m = d.reshape(-1) > 0
    np.testing.assert_allclose(filt[m], ref[m], rtol=1e-14, atol=0)
    # validity mask, raster order and float32 back-projection (closed form)
    assert np.array_equal(np.sort(pix), np.nonzero(GOLD["bil_valid"].reshape(-1))[0])
    assert np.array_equal(pix, np.sort(pix))
    np.testing.assert_array_equal(xyzi[:, :3], GOLD["bil_xyz"])
    np.testing.assert_array_equal(rgb[:, :3], GOLD["bil_rgb"])


def test_bilateral_invariants():
    cam = O.make_camera(20, 15, 30.0, 30.0, 10.0, 7.5)
    bgr = np.zeros((15, 20, 3), np.uint8)
    # constant depth -> unchanged
    d = np.full((15, 20), 1234, np.uint16)
    _, _, _, f = O.cloud_from_rgbd(d, bgr, cam, O.Bilateral(1, 3, 2.0, 30.0), 0.2, 3.5)
    np.testing.assert_allclose(f, 1234.0, rtol=1e-14)
    # linear ramp + huge range sigma -> symmetric normalised kernel reproduces the ramp inside
    yy, xx = np.mgrid[0:15, 0:20]
    d = (1000 + 3 * xx + 2 * yy).astype(np.uint16)
    _, _, _, f = O.cloud_from_rgbd(d, bgr, cam, O.Bilateral(1, 2, 2.0, 1e6), 0.2, 3.5)
    f = f.reshape(15, 20)
    np.testing.assert_allclose(f[2:-2, 2:-2], d[2:-2, 2:-2], rtol=0, atol=1e-9)
    # step edge + small range sigma -> no mixing across the edge
    d = np.where(xx < 10, 1000, 2000).astype(np.uint16)
    _, _, _, f = O.cloud_from_rgbd(d, bgr, cam, O.Bilateral(1, 2, 2.0, 30.0), 0.2, 3.5)
    f = f.reshape(15, 20)
    assert np.abs(f[:, :10] - 1000).max() < 1e-9 and np.abs(f[:, 10:] - 2000).max() < 1e-9
    # holes are skipped, a centre hole is invalid
    d = np.full((15, 20), 1500, np.uint16)
    d[7, 10] = 0
    d[3, 3] = 0
    x, _, pix, f = O.cloud_from_rgbd(d, bgr, cam, O.Bilateral(1, 2, 2.0, 30.0), 0.2, 3.5)
    assert 7 * 20 + 10 not in set(pix.tolist()) and len(x) == 15 * 20 - 2
    np.testing.assert_allclose(f.reshape(15, 20)[d > 0], 1500.0, rtol=1e-14)


def test_backprojection_identity_extrinsics_samples_own_pixel():
    from cuda_3d_amd import scene

    (td, tc), _, _, cam = scene.frame_pair(5, 64, 48)
    c = O.make_camera(64, 48, cam.fx, cam.fy, cam.cx, cam.cy)
    # unique colour per pixel to read back which pixel was sampled
    ids = np.arange(64 * 48)
    bgr = np.stack([ids % 256, (ids // 256) % 256, np.zeros_like(ids)], -1).astype(np.uint8).reshape(48, 64, 3)
    x, rgb, pix, _ = O.cloud_from_rgbd(td, bgr, c, O.Bilateral(0, 0, 1, 1), 0.2, 3.5)
    b = np.rint(rgb[:, 2] * 255).astype(int)
    g = np.rint(rgb[:, 1] * 255).astype(int)
    sampled = b + 256 * g
    # int(u + 0.5) of an FP32 re-projection: own pixel, or one to the left / above when the
    # rounding lands just below the integer (reference quirk, cuda_point_cloud_factory.cu:242)
    dx, dy = (pix % 64) - (sampled % 64), (pix // 64) - (sampled // 64)
    assert set(np.unique(dx)) <= {0, 1} and set(np.unique(dy)) <= {0, 1}
    assert np.mean(sampled == pix) > 0.5
    z = x[:, 2].astype(np.float64)
    u = (pix % 64) - 0.5
    np.testing.assert_allclose(x[:, 0], (u - cam.cx) * z / cam.fx, rtol=2e-6, atol=1e-7)
    c = rgb.astype(np.float64)
    inten = 0.2126 * c[:, 0] + 0.7152 * c[:, 1] + 0.0722 * c[:, 2]
    np.testing.assert_array_equal(x[:, 3], inten.astype(np.float32))


def test_invalid_camera_rejected():
    cam = small_cam()
    cam.depth_scale = 0.0
    with pytest.raises(ValueError):
        O.cloud_from_rgbd(GOLD["bil_depth"], GOLD["bil_bgr"], cam, O.Bilateral(0, 0, 1, 1), 0.2, 3.5)


def test_bounds_keep_reference_max_init():
    x = f4(-np.abs(np.random.default_rng(0).normal(size=(10, 3))) - 1.0)
    mn, mx = O.bounds(x)
    assert np.all(mn == x[:, :3].min(0))
    assert np.all(mx == np.float32(np.finfo(np.float32).tiny))  # FLT_MIN, geometry_util.cuh:92


# ---- voxel --------------------------------------------------------------------------------

def test_voxel_matches_numpy_unique():
    xyz, rgb, v = GOLD["vox_xyz"], GOLD["vox_rgb"], float(GOLD["vox_size"])
    ox, orgb, keys = O.voxel_downsample(f4(xyz), f4(rgb), v)
    mean = GOLD["vox_mean"]
    assert len(ox) == len(mean)
    # 2^-30 fixed-point sums: within one float ulp of the FP64 mean plus the 2^-31 per-point
    # quantisation (dominant only for means near zero)
    for got, want in ((ox[:, :3], mean[:, :3]), (orgb[:, :3], mean[:, 3:])):
        tol = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64) + 2.0 ** -30
        assert np.all(np.abs(got.astype(np.float64) - want) <= tol)
    assert np.mean(ox[:, :3] == mean[:, :3].astype(np.float32)) > 0.95
    np.testing.assert_array_equal(keys, GOLD["vox_keys"])
    c = orgb.astype(np.float64)
    inten = (0.2126 * c[:, 0] + 0.7152 * c[:, 1] + 0.0722 * c[:, 2]).astype(np.float32)
    np.testing.assert_array_equal(ox[:, 3], inten)


def test_voxel_edge_cases():
    one = f4([[0.1, 0.2, 0.3]])
    ox, orgb, _ = O.voxel_downsample(one, f4([[0.5, 0.5, 0.5]]), 0.01)
    assert len(ox) == 1 and np.all(ox[0, :3] == one[0, :3])
    same = f4(np.tile([[0.5, 0.5, 1.0]], (7, 1)))
    ox, _, _ = O.voxel_downsample(same, same, 0.02)
    assert len(ox) == 1
    with pytest.raises(ValueError):
        O.voxel_downsample(same, same, 0.0)
    with pytest.raises(ValueError):
        O.voxel_downsample(f4([[0, 0, 0], [1e4, 1e4, 1e4]]), f4(np.zeros((2, 3))), 1e-9)


def test_voxel_huge_lattice_is_not_refused():
    """The reference guards only each axis against INT_MAX voxels (point_cloud.cu:213-218): a
    lattice of 2^50 cells (2 km at 1 mm per axis, far beyond any dense table) is valid, and its
    voxels come out in (kx, ky, kz) order."""
    rng = np.random.default_rng(11)
    xyz = np.concatenate([rng.uniform(0, 0.01, (40, 3)), [[0, 0, 0], [2000.0, 1500.0, 1800.0]]])
    v = np.float32(0.001)
    ox, _, keys = O.voxel_downsample(f4(xyz), f4(np.zeros_like(xyz)), v)
    mn = xyz.astype(np.float32).min(0)
    want = np.unique(np.floor((xyz.astype(np.float32) - mn) / v).astype(np.int64), axis=0)
    np.testing.assert_array_equal(keys, want)
    assert len(ox) == len(want) and keys[-1, 0] > 10 ** 6


# ---- neighbours -----------------------------------------------------------------------------

def test_transform_closed_forms_and_reference_fp32():
    """transform: identity and a quarter turn about z are exact, T then T^-1 returns the points,
    normals stay unit, and the result is within a few FP32 ulps of the reference's own FP32
    product R p + t (geometry_util.cuh:133-149)."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-3, 3, (2000, 3)), rng.uniform(0, 1, (2000, 1))], 1).astype(np.float32)
    same, _ = O.transform(x, np.eye(4))
    np.testing.assert_array_equal(same, x)
    Tz = np.eye(4)
    Tz[:3, :3] = Rotation.from_euler("z", 90, degrees=True).as_matrix()
    q, _ = O.transform(x, Tz)
    np.testing.assert_array_equal(q[:, :3], np.stack([-x[:, 1], x[:, 0], x[:, 2]], 1))
    np.testing.assert_array_equal(q[:, 3], x[:, 3])
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec([0.3, -0.2, 0.5]).as_matrix()
    T[:3, 3] = [0.4, -1.2, 2.0]
    v = rng.normal(size=(2000, 3))
    v = np.concatenate([v / np.linalg.norm(v, axis=1, keepdims=True), np.zeros((2000, 1))], 1)
    y, vy = O.transform(x, T, v)
    back, vb = O.transform(y, np.linalg.inv(T), vy)
    assert np.abs(back[:, :3] - x[:, :3]).max() < 4e-6
    assert np.abs(np.linalg.norm(vy[:, :3], axis=1) - 1).max() < 3e-7
    assert np.abs(vb[:, :3] - v[:, :3]).max() < 3e-7
    R32, t32, p32 = T[:3, :3].astype(np.float32), T[:3, 3].astype(np.float32), x[:, :3]
    ref = np.stack([((R32[r, 0] * p32[:, 0] + R32[r, 1] * p32[:, 1]) + R32[r, 2] * p32[:, 2]) + t32[r]
                    for r in range(3)], 1)
    scale = np.abs(p32).sum(1, keepdims=True) + np.abs(t32)
    assert (np.abs(y[:, :3] - ref) <= 4 * np.finfo(np.float32).eps * scale).all()


def test_accumulate_is_transform_then_concatenation():
    rng = np.random.default_rng(6)
    m = rng.uniform(-1, 1, (500, 4)).astype(np.float32)
    f = rng.uniform(-1, 1, (300, 4)).astype(np.float32)
    mr, fr = rng.uniform(0, 1, (500, 4)).astype(np.float32), rng.uniform(0, 1, (300, 4)).astype(np.float32)
    T = np.eye(4)
    T[:3, 3] = [0.5, 0.0, -0.25]  # dyadic shifts of values in [-1, 1]: exact
    x, r = O.accumulate(m, mr, f, fr, T)
    np.testing.assert_array_equal(x[:500, :3], m[:, :3] + np.float32([0.5, 0.0, -0.25]))
    np.testing.assert_array_equal(x[500:], f)
    np.testing.assert_array_equal(r, np.concatenate([mr, fr]))
    vx, vr = O.accumulate(m, mr, f, fr, T, 0.1)
    ox, orgb, _ = O.voxel_downsample(x, r, 0.1)
    np.testing.assert_array_equal(vx, ox)
    np.testing.assert_array_equal(vr, orgb)


def cluster_scene(seed, tol):
    """Blobs of uneven sizes, a long chain (spacing 0.6 tol: one neighbour each way), two blobs
    bridged by a chain, and isolated points; shuffled so indices are unrelated to geometry."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(c, 0.3 * tol * s, (m, 3)) for c, m, s in
             (((0, 0, 0), 400, 3), ((1, 0, 0), 60, 1), ((0, 1, 0), 1500, 5), ((1, 1, 1), 12, 1),
              ((2, 0, 0), 250, 3), ((2, 0.5, 0), 250, 3))]
    parts.append(np.stack([np.linspace(0, 0.6 * tol * 299, 300), np.full(300, 3.0), np.zeros(300)], 1))
    parts.append(np.stack([np.full(60, 2.0), np.linspace(0.03, 0.47, 60), np.zeros(60)], 1))
    parts.append(rng.uniform(4, 6, (40, 3)))
    x = np.concatenate(parts)
    x = x[rng.permutation(len(x))]
    return np.concatenate([x, rng.uniform(0, 1, (len(x), 1))], 1).astype(np.float32)


def test_clustering_matches_connected_components():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    tol = 0.02
    x = cluster_scene(3, tol)
    pairs = cKDTree(x[:, :3].astype(np.float64)).query_pairs(tol, output_type="ndarray")
    d = np.linalg.norm(x[pairs[:, 0], :3].astype(np.float64) - x[pairs[:, 1], :3], axis=1)
    pairs = pairs[d < tol]
    A = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(len(x), len(x)))
    nc, lab = connected_components(A, directed=False)
    sizes = np.bincount(lab)
    for mn in (0, 10, 100):
        cl = O.euclidean_clustering(x, tol, mn, len(x))
        want = [np.flatnonzero(lab == c) for c in np.unique(lab) if sizes[c] >= max(mn, 1)]
        want.sort(key=lambda a: a[0])
        assert len(cl) == len(want)
        for a, b in zip(cl, want):
            np.testing.assert_array_equal(np.sort(a), b)
            assert a[0] == b[0]  # the seed is the lowest index


def test_residual_segmentation_matches_gated_components():
    """cuda_residual_based_segmentation (cuda_clustering.cu:251-341) without truncation is the
    connected components of the radius graph restricted to edges whose residuals differ by less
    than 0.003 (FP32 difference): pinned by scipy on that graph, with residuals laid out so that
    many edges are cut (steps of 0.002 and 0.004 along the chain, random residuals in blobs)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    tol = 0.02
    x = cluster_scene(5, tol)
    rng = np.random.default_rng(9)
    res = rng.choice(np.float32([0.0, 0.002, 0.004, 0.0055, 0.01]), len(x)).astype(np.float32)
    res += rng.uniform(0, 1e-4, len(x)).astype(np.float32)
    pairs = cKDTree(x[:, :3].astype(np.float64)).query_pairs(tol, output_type="ndarray")
    d = np.linalg.norm(x[pairs[:, 0], :3].astype(np.float64) - x[pairs[:, 1], :3], axis=1)
    gate = np.abs(res[pairs[:, 0]] - res[pairs[:, 1]]).astype(np.float64) < 0.003
    keep = (d < tol) & gate
    assert 0.2 < gate[d < tol].mean() < 0.9  # the gate is exercised both ways
    pairs = pairs[keep]
    A = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(len(x), len(x)))
    nc, lab = connected_components(A, directed=False)
    sizes = np.bincount(lab)
    for mn in (0, 5, 50):
        cl = O.euclidean_clustering(x, tol, mn, len(x), residual=res)
        want = [np.flatnonzero(lab == c) for c in np.unique(lab) if sizes[c] >= max(mn, 1)]
        want.sort(key=lambda a: a[0])
        assert len(cl) == len(want)
        for a, b in zip(cl, want):
            np.testing.assert_array_equal(np.sort(a), b)
    # residuals all equal: the gate never fires, so the segments are the Euclidean clusters
    flat = O.euclidean_clustering(x, tol, 10, len(x), residual=np.zeros(len(x), np.float32))
    plain = O.euclidean_clustering(x, tol, 10, len(x))
    assert len(flat) == len(plain) and all(np.array_equal(a, b) for a, b in zip(flat, plain))


def test_residual_gate_is_checked_against_the_point_being_expanded():
    """Three collinear points 0 - 1 - 2 (spacing 0.6 tol, so 0 and 2 are not neighbours) with
    residuals 0, 0.002, 0.004: 0-1 and 1-2 pass the gate (0.002 < 0.003), so all three are one
    segment, although |r0 - r2| = 0.004; with r1 = 0.0035 neither edge passes."""
    tol = 0.1
    x = np.zeros((3, 4), np.float32)
    x[:, 0] = np.arange(3) * 0.06
    seg = O.euclidean_clustering(x, tol, 1, 10, residual=np.float32([0, 0.002, 0.004]))
    assert [s.tolist() for s in seg] == [[0, 1, 2]]
    seg = O.euclidean_clustering(x, tol, 1, 10, residual=np.float32([0, 0.0035, 0.0]))
    assert [s.tolist() for s in seg] == [[0], [1], [2]]


def test_clustering_truncates_at_max_size_like_the_reference():
    """A chain with one neighbour each way, searched from index 0 with max 3: the queue stops at
    3 points, so the chain of 10 comes out as 0-2, 3-5, 6-8, 9 (cuda_clustering.cu:75-96)."""
    tol = 0.1
    x = np.zeros((10, 4), np.float32)
    x[:, 0] = np.arange(10) * 0.06
    cl = O.euclidean_clustering(x, tol, 1, 3)
    assert [c.tolist() for c in cl] == [[0, 1, 2], [3, 4, 5], [6, 7, 8], [9]]
    assert [c.tolist() for c in O.euclidean_clustering(x, tol, 2, 3)] == [[0, 1, 2], [3, 4, 5], [6, 7, 8]]
    with pytest.raises(ValueError):
        O.euclidean_clustering(x, tol, 4, 3)


def test_nn_and_radius_match_brute_force():
    ref, qry, r = f4(GOLD["nb_ref"]), f4(GOLD["nb_qry"]), float(GOLD["nb_radius"])
    np.testing.assert_array_equal(O.nn_search(qry, ref, r), GOLD["nb_nn"])
    off, idx = O.radius_search(qry, ref, r, 0)
    np.testing.assert_array_equal(off, GOLD["nb_off"])
    np.testing.assert_array_equal(idx, GOLD["nb_idx"])
    off, idx = O.radius_search(qry, ref, r, 3)
    np.testing.assert_array_equal(off, GOLD["nb_offk"])
    np.testing.assert_array_equal(idx, GOLD["nb_idxk"])


def test_radius_sets_agree_with_kdtree():
    rng = np.random.default_rng(3)
    p = rng.uniform(0, 1, size=(3000, 3)).astype(np.float32)
    off, idx = O.radius_search(f4(p), f4(p), 0.05, 0)
    tree = cKDTree(p.astype(np.float64))
    ball = tree.query_ball_point(p.astype(np.float64), 0.05 * (1 - 1e-6))
    mism = 0
    for i, b in enumerate(ball):
        mine = set(idx[off[i]:off[i + 1]].tolist())
        mism += not set(b) <= mine
    assert mism == 0


def _cell_box_violations(p, cells, origin, cell, dims):
    """Exact-arithmetic check (fractions) that every point lies in the half-open box of its cell,
    origin + c * cell <= p < origin + (c + 1) * cell, except that a point below the first cell
    sits in cell 0 and one beyond the last in cell dims - 1."""
    from fractions import Fraction as F

    bad = 0
    for i in range(len(p)):
        for a in range(3):
            x, o, s, c = F(float(p[i, a])), F(float(origin[a])), F(float(cell)), int(cells[i, a])
            lo_ok = c == 0 or o + c * s <= x
            hi_ok = c == dims[a] - 1 or x < o + (c + 1) * s
            below = x < o + c * s
            above = x >= o + (c + 1) * s
            bad += not (lo_ok and hi_ok) or (below and c != 0) or (above and c != dims[a] - 1)
            bad += not (0 <= c < dims[a])
    return bad


def test_grid_cells_box_property_and_reference_extent():
    """Cell binning pinned by properties, not by its own formula: each point lies inside its
    cell's box (exact arithmetic), points outside the grid clamp to the border cells, the origin is
    the bound padded by 1.5 r (point_cloud.cu:528-537) and the cell counts follow the reference's
    FP32 extent formula int((max - min) / cell) + 1 (cuda_nn_search.cu:557-562)."""
    rng = np.random.default_rng(4)
    p = f4(rng.uniform(-1, 1, size=(500, 3)))
    r = np.float32(0.1)
    cells, origin, cell, dims = O.grid_cells(p, float(r))
    mn, mx = O.bounds(p)
    pad = np.float32(1.5) * r
    assert np.all(origin == mn - pad) and cell == r
    want_dims = ((mx + pad) - (mn - pad)) / r
    np.testing.assert_array_equal(dims, want_dims.astype(np.int32) + 1)
    assert _cell_box_violations(p[:, :3], cells, origin, cell, dims) == 0
    # a grid over a sub-box: points beyond it clamp to the border cells
    bmin, bmax = np.float32([-0.3, -0.2, -0.1]), np.float32([0.2, 0.3, 0.4])
    cells2, origin2, cell2, dims2 = O.grid_cells(p, float(r), bmin, bmax)
    np.testing.assert_array_equal(dims2, (((bmax + pad) - (bmin - pad)) / r).astype(np.int32) + 1)
    assert _cell_box_violations(p[:, :3], cells2, origin2, cell2, dims2) == 0
    outside = (p[:, :3] < origin2) | (p[:, :3] >= origin2 + dims2 * cell2)
    assert outside.any() and np.all((cells2 == 0) | (cells2 == dims2 - 1) | ~outside)
    # tiny radius: the cell edge doubles until the dense grid fits the cap
    _, _, cell2, dims2 = O.grid_cells(p, 1e-4)
    assert cell2 > np.float32(1e-4) and np.prod(dims2.astype(np.float64)) <= 2 ** 25


# ---- normals / eigen ----------------------------------------------------------------------

def test_eigen3_matches_numpy():
    rng = np.random.default_rng(5)
    for k in range(200):
        A = rng.normal(size=(3, 3))
        A = A + A.T
        if k % 10 == 0:
            A = np.diag(rng.normal(size=3))
        if k % 17 == 0:
            A = np.zeros((3, 3))
        ev, V = O.eigen3(A)
        np.testing.assert_allclose(np.sort(ev), np.linalg.eigvalsh(A), atol=1e-12 * (1 + np.abs(A).max()))
        np.testing.assert_allclose(A @ V, V * ev, atol=1e-11 * (1 + np.abs(A).max()))
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)


@pytest.mark.parametrize("K", [0, 15])
def test_normals_match_numpy_eigh(K):
    p, r = GOLD["nrm_pts"], float(GOLD["nrm_radius"])
    n = O.estimate_normals(f4(p), r, K)[:, :3].astype(np.float64)
    ref = GOLD[f"nrm_K{K}"]
    ok = np.abs(ref).sum(1) > 0
    assert angle(n[ok], ref[ok], signed=True).max() < 1e-6
    assert np.all(n[:, 2] <= 0)  # oriented toward (0,0,-1)
    assert np.all(n[~ok] == 0)


def test_normals_closed_forms():
    rng = np.random.default_rng(6)
    a = rng.uniform(-0.5, 0.5, size=(400, 2))
    nrm = np.array([0.2, -0.3, 0.9])
    nrm /= np.linalg.norm(nrm)
    e1 = np.cross(nrm, [1, 0, 0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(nrm, e1)
    plane = np.array([0, 0, 1.5]) + a[:, :1] * e1 + a[:, 1:] * e2
    n = O.estimate_normals(f4(plane), 0.1, 30)[:, :3]
    n = n[np.abs(n).sum(1) > 0]  # isolated border points (< 3 neighbours) have no normal
    assert len(n) > 390
    # float32 input quantisation bounds the fit: ~eps32 * |p| / r ~ 1e-6 rad
    assert angle(n, np.tile(nrm, (len(n), 1))).max() < 5e-6 and np.all(n[:, 2] < 0)
    u = rng.normal(size=(2000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = u[u[:, 2] < -0.3]
    sph = np.array([0, 0, 2.0]) + 0.3 * u
    n = O.estimate_normals(f4(sph), 0.06, 30)[:, :3]
    radial = (sph - [0, 0, 2.0]) / 0.3
    err = np.degrees(angle(n, radial))  # random sampling; rim points see one-sided patches
    assert np.median(err) < 2.0 and np.percentile(err, 90) < 5.0
    two = O.estimate_normals(f4([[0, 0, 1], [0.001, 0, 1]]), 0.1, 0)
    assert np.all(two == 0)  # count < 3


def test_gradient_matches_numpy_solve_and_plane_slope():
    p, r = GOLD["nrm_pts"], float(GOLD["grad_radius"])
    g = O.color_gradient(f4(p, GOLD["grad_int"]), f4(GOLD["grad_nrm"]), r, 0)[:, :3]
    ref = GOLD["grad"]
    np.testing.assert_allclose(g, ref, rtol=2e-6, atol=2e-6 * np.abs(ref).max())
    # linear intensity on a plane: gradient = tangential slope (up to the 1e-6 ridge)
    rng = np.random.default_rng(7)
    a = rng.uniform(-0.3, 0.3, size=(600, 2))
    pts = np.stack([a[:, 0], a[:, 1], np.full(600, 1.0)], 1)
    inten = 0.5 + 0.8 * a[:, 0] - 0.4 * a[:, 1]
    n = np.tile([[0, 0, -1.0]], (600, 1))
    g = O.color_gradient(f4(pts, inten), f4(n), 0.06, 0)[:, :3]
    inner = (np.abs(a) < 0.2).all(1)
    np.testing.assert_allclose(g[inner], np.tile([0.8, -0.4, 0.0], (inner.sum(), 1)), atol=2e-3)


# ---- Gauss-Newton, solve, ICP, adjoint ----------------------------------------------------------

def test_gauss_newton_matches_finite_difference_jacobian():
    G = GOLD
    out = O.build_gn(G["gn_src"], G["gn_tgt"], f4(G["gn_tn"]), f4(G["gn_tg"]), G["gn_nn"],
                     float(G["gn_lam"]), G["gn_T"])
    JTJ = np.zeros((6, 6))
    k = 0
    for a in range(6):
        for b in range(a, 6):
            JTJ[a, b] = JTJ[b, a] = out[k]
            k += 1
    scale = np.abs(G["gn_JTJ"]).max()
    np.testing.assert_allclose(JTJ, G["gn_JTJ"], rtol=1e-6, atol=1e-7 * scale)
    np.testing.assert_allclose(out[21:27], G["gn_JTr"], rtol=1e-6, atol=1e-7 * np.abs(G["gn_JTr"]).max())
    assert abs(out[27] - G["gn_r2"]) < 1e-9 * max(1, G["gn_r2"])
    assert out[28] == G["gn_ncorr"]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_residual_kinds_match_golden(kind):
    """Per-point residuals (cuda_color_icp_build_least_square.cu:13-132) against the numpy vectors
    of tests/golden/make_golden.py, and the identities that tie the kinds together: the objective
    kind sums to the weighted squares of the pair kind, i.e. to the Gauss-Newton r^2 of the same
    correspondences, and the absolute kind is |pair0| + |pair1|."""
    G = GOLD
    args = (G["gn_src"], G["gn_tgt"], f4(G["gn_tn"]), f4(G["gn_tg"]), G["gn_nn"], float(G["gn_lam"]))
    per, total = O.residuals(*args, kind, G["gn_T"])
    want = {0: G["res_pair"], 1: G["res_abs"], 2: G["res_rmse"]}[kind]
    np.testing.assert_allclose(per, want, rtol=1e-9, atol=1e-12)
    assert abs(total - want.sum()) <= 1e-9 * max(1.0, np.abs(want).sum())
    if kind == 2:
        assert abs(total - G["gn_r2"]) <= 1e-9 * G["gn_r2"]
        assert np.all(per[G["gn_nn"] < 0] == 0)


def test_solve6_and_euler():
    st, xi, T = O.solve6([G for G in [GOLD["sv_A"][a, b] for a in range(6) for b in range(a, 6)]], GOLD["sv_b"])
    assert st == 0
    np.testing.assert_allclose(xi, GOLD["sv_xi"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(T, GOLD["sv_T"], atol=1e-13)
    ang = np.array([0.1, -0.2, 0.3, 0, 0, 0])
    np.testing.assert_allclose(O.euler_to_T(ang)[:3, :3], Rotation.from_euler("xyz", ang[:3]).as_matrix(), atol=1e-15)
    # not positive definite -> identity step and status 4 (eigen_solver.hpp:60-62)
    st, xi, T = O.solve6(np.zeros(21), np.ones(6))
    assert st == 4 and np.all(xi == 0) and np.allclose(T, np.eye(4))


def noiseless_pair(seed=8, n=4000):
    rng = np.random.default_rng(seed)
    pts = []
    for nrm, d in (([0, 0, 1], 2.0), ([0, 1, 0], 0.8), ([1, 0, 0], -0.9)):
        a = rng.uniform(-0.8, 0.8, size=(n, 2))
        nrm = np.array(nrm, float)
        e1 = np.roll(nrm, 1)
        e2 = np.cross(nrm, e1)
        pts.append(nrm * d + a[:, :1] * e1 + a[:, 1:] * e2 * 0.999)
    p = np.concatenate(pts)
    p = p[(p[:, 2] > 0.5) & (np.abs(p[:, 0]) < 1.0) & (np.abs(p[:, 1]) < 0.9)]
    inten = 0.5 + 0.3 * np.sin(8 * p[:, 0]) * np.cos(6 * p[:, 1]) + 0.2 * np.sin(5 * p[:, 2])
    return p, inten


def test_icp_recovers_known_transform():
    p, inten = noiseless_pair()
    T_true = np.eye(4)
    T_true[:3, :3] = Rotation.from_euler("xyz", [0.03, -0.02, 0.04]).as_matrix()
    T_true[:3, 3] = [0.02, -0.015, 0.01]
    src = (p - T_true[:3, 3]) @ T_true[:3, :3]  # T_true maps src -> tgt
    tgt = f4(p, inten)
    sx = f4(src, inten)
    tn = O.estimate_normals(tgt, 0.05, 30)
    tg = O.color_gradient(tgt, tn, 0.04, 0)
    T, rep = O.icp_align(sx, tgt, tn, tg, O.icp_params(60, 0.1, 0.968, 0.0, 0.0))
    ang = np.arccos(np.clip((np.trace(T[:3, :3] @ T_true[:3, :3].T) - 1) / 2, -1, 1))
    assert ang < 1e-6 and np.linalg.norm(T[:3, 3] - T_true[:3, 3]) < 1e-6
    assert rep.iterations == 60 and rep.fitness > 0.95


def test_objective_gradient_matches_finite_differences():
    p, inten = noiseless_pair(9, 800)
    tgt = f4(p, inten)
    tn = O.estimate_normals(tgt, 0.1, 30)
    tg = O.color_gradient(tgt, tn, 0.08, 0)
    rng = np.random.default_rng(10)
    src = f4(p + rng.normal(0, 0.01, p.shape), inten + rng.normal(0, 0.02, len(p)))
    T = np.eye(4)
    T[:3, :3] = Rotation.from_euler("xyz", [0.01, 0.02, -0.01]).as_matrix()
    nn = O.nn_search(f4((src[:, :3].astype(np.float64) @ T[:3, :3].T) + T[:3, 3]), tgt, 0.1)
    xi = np.array([0.004, -0.003, 0.002, 0.001, -0.002, 0.003])
    E, dxi, dp = O.objective_grad(src, tgt, tn, tg, nn, 0.968, T, xi)
    h = 1e-6
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        fd[k] = (O.objective_grad(src, tgt, tn, tg, nn, 0.968, T, xi + e, False)[0] -
                 O.objective_grad(src, tgt, tn, tg, nn, 0.968, T, xi - e, False)[0]) / (2 * h)
    np.testing.assert_allclose(dxi, fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max())
    # at xi = 0 the pose gradient equals 2 J^T r of the Gauss-Newton build
    _, dxi0, _ = O.objective_grad(src, tgt, tn, tg, nn, 0.968, T, np.zeros(6))
    sysv = O.build_gn(src, tgt, tn, tg, nn, 0.968, T)
    np.testing.assert_allclose(dxi0, 2 * sysv[21:27], rtol=1e-10, atol=1e-12)
    # per-point gradients vs finite differences on a few points
    for i in np.nonzero(nn >= 0)[0][:5]:
        for a in range(4):
            s2 = src.copy()
            hp = 1e-4 if a < 3 else 1e-3
            s2[i, a] += np.float32(hp)
            ep = O.objective_grad(s2, tgt, tn, tg, nn, 0.968, T, xi, False)[0]
            s2[i, a] -= np.float32(2 * hp)
            em = O.objective_grad(s2, tgt, tn, tg, nn, 0.968, T, xi, False)[0]
            step = float(np.float32(src[i, a] + np.float32(hp)) - np.float32(src[i, a] - np.float32(hp)))
            assert abs((ep - em) / step - dp[i, a]) < 1e-4 * (1 + abs(dp[i, a]))


def test_outlier_removal_counts():
    rng = np.random.default_rng(11)
    p = np.concatenate([rng.uniform(0, 0.2, size=(400, 3)), rng.uniform(0.5, 3, size=(30, 3))]).astype(np.float32)
    ox, _ = O.radius_outlier_removal(f4(p), f4(np.zeros_like(p)), 0.03, 4)
    d2 = ((p[:, None].astype(np.float64) - p[None].astype(np.float64)) ** 2).sum(-1)
    keep = (d2 < 0.03 * 0.03).sum(1) > 4
    np.testing.assert_array_equal(ox[:, :3], p[keep])


PLANTS = {
    1: ("transposed cross product in J_g", lambda: test_gauss_newton_matches_finite_difference_jacobian()),
    2: ("dropped (k-1)^2 n n^T term", lambda: test_gradient_matches_numpy_solve_and_plane_slope()),
    3: ("wrong Euler composition order", lambda: test_solve6_and_euler()),
    4: ("largest instead of smallest eigenvector", lambda: test_normals_match_numpy_eigh(15)),
    5: ("nearest-neighbour tie toward the higher index", lambda: test_nn_and_radius_match_brute_force()),
    6: ("Gaussian weight with the 1/2 factor", lambda: test_bilateral_matches_direct_fp64_loop()),
    7: ("objective residual squares the weighted residuals", lambda: test_residual_kinds_match_golden(2)),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistakes_are_caught(plant, tmp_path):
    """Each plausible mistake, planted in the oracle (-DORC_PLANT), fails its pin."""
    path = O.build_variant(plant, str(tmp_path))
    with O.using(path):
        with pytest.raises(AssertionError):
            PLANTS[plant][1]()
