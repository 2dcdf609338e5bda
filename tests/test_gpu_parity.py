"""GPU parity: libgca.so (through the C ABI) against the FP64 oracle on the same seeded inputs.

Bars: bit-exact for pixel validity, point order and coordinates, voxel assignment and means,
grid cells, neighbour index sets and nearest neighbours; normals within 1e-5 rad (they come
out bit-exact); poses within 1e-5 rad / 1e-5 m of the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402

LAM = 0.968
CONFIGS = {"cfg0_320x240": (320, 240, 11), "cfg1_640x480": (640, 480, 12)}


def angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.abs((a * b).sum(-1)))


def pose_err(T, U):
    dR = T[:3, :3] @ U[:3, :3].T
    return (float(np.arctan2(np.linalg.norm([dR[2, 1] - dR[1, 2], dR[0, 2] - dR[2, 0], dR[1, 0] - dR[0, 1]]),
                             np.trace(dR) - 1)), float(np.linalg.norm(T[:3, 3] - U[:3, 3])))


class Frame:
    def __init__(self, g, ctx, W, H, seed):
        from cuda_3d_amd import scene

        (self.td, self.tc), (self.sd, self.sc), self.G, cam = scene.frame_pair(seed, W, H)
        self.camg = g.make_camera(W, H, cam.fx, cam.fy, cam.cx, cam.cy)
        self.camo = O.make_camera(W, H, cam.fx, cam.fy, cam.cx, cam.cy)
        self.bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
        self.bilo = O.Bilateral(1, 2, 2.0, 30.0)
        self.tgt = g.PointCloud.create_from_rgbd(self.td, self.tc, self.camg, self.bil, 0.2, 3.5, ctx=ctx)
        self.src = g.PointCloud.create_from_rgbd(self.sd, self.sc, self.camg, self.bil, 0.2, 3.5, ctx=ctx)
        self.otx, self.otr, _, _ = O.cloud_from_rgbd(self.td, self.tc, self.camo, self.bilo, 0.2, 3.5)
        self.osx, self.osr, _, _ = O.cloud_from_rgbd(self.sd, self.sc, self.camo, self.bilo, 0.2, 3.5)
        self.bmin, self.bmax = O.bounds(self.otx)


@pytest.fixture(scope="module", params=sorted(CONFIGS))
def frame(request, ctx):
    import cuda_3d_amd as g

    W, H, seed = CONFIGS[request.param]
    print(f"[gpu-parity] frame {request.param}", flush=True)
    return Frame(g, ctx, W, H, seed)


@pytest.fixture(scope="module")
def level(frame):
    """Target at voxel 0.02 with normals (r 0.05, K 30) and gradient (r 0.04), GPU and oracle."""
    v = 0.02
    vc = frame.tgt.voxel_grid_down_sample(v)
    ox, orgb, _ = O.voxel_downsample(frame.otx, frame.otr, v)
    vc.estimate_normals(2.5 * v, 30)
    vc.compute_color_gradient(2 * v, 0)
    on = O.estimate_normals(ox, 2.5 * v, 30, frame.bmin, frame.bmax)
    og = O.color_gradient(ox, on, 2 * v, 0, frame.bmin, frame.bmax)
    return dict(v=v, gpu=vc, ox=ox, orgb=orgb, on=on, og=og)


# ---- front end ---------------------------------------------------------------------------------

@pytest.mark.parametrize("kernels", ["onepass", "two_pass"])
@pytest.mark.parametrize("mode", ["fixed", "fixed_wide", "adaptive", "none"])
def test_frontend_bitexact(frame, ctx, mode, kernels):
    """fixed: 5x5 with its weight tables in LDS; fixed_wide: 15x15 with a range table too long for
    LDS (sigma_r = 200 mm: ~5500 entries), read from global memory.  onepass: pixels, look-back
    offset and compaction in one kernel; two_pass: pixels to staging, then compaction (default)."""
    import cuda_3d_amd as g

    print(f"[gpu-parity] front end {kernels} {mode}", flush=True)
    if kernels == "onepass":
        with tuned(ctx, frontend=1):
            _frontend_bitexact(g, frame, ctx, mode)
    else:
        _frontend_bitexact(g, frame, ctx, mode)


def _frontend_bitexact(g, frame, ctx, mode):
    if mode == "fixed":
        bg, bo = frame.bil, frame.bilo
    elif mode == "fixed_wide":
        bg, bo = g.bilateral(g.BILATERAL_FIXED, 7, 4.0, 200.0), O.Bilateral(1, 7, 4.0, 200.0)
    elif mode == "adaptive":
        bg, bo = g.bilateral(g.BILATERAL_REFERENCE_ADAPTIVE, 0, 0, 0), O.Bilateral(2, 0, 0, 0)
    else:
        bg, bo = g.bilateral(g.BILATERAL_NONE, 0, 1, 1), O.Bilateral(0, 0, 1, 1)
    pc = g.PointCloud.create_from_rgbd(frame.td, frame.tc, frame.camg, bg, 0.2, 3.5, ctx=ctx)
    xyz, rgb, inten = pc.download()
    ox, orgb, _, _ = O.cloud_from_rgbd(frame.td, frame.tc, frame.camo, bo, 0.2, 3.5)
    assert len(xyz) == len(ox) > 1000
    np.testing.assert_array_equal(xyz.view(np.uint32), np.ascontiguousarray(ox[:, :3]).view(np.uint32))
    np.testing.assert_array_equal(inten.view(np.uint32), np.ascontiguousarray(ox[:, 3]).view(np.uint32))
    np.testing.assert_array_equal(rgb, orgb[:, :3])
    mn, mx = O.bounds(ox)
    np.testing.assert_array_equal(pc.get_min_bound(), mn)
    np.testing.assert_array_equal(pc.get_max_bound(), mx)


def test_cloud_append(ctx):
    """point_cloud::operator+= (point_cloud.cu:55-73): the concatenation bit for bit, growth past
    the capacity, normals only when both clouds have them, bounds over every point (the max with
    the reference's FLT_MIN start), and a cloud appended to itself."""
    import cuda_3d_amd as g

    rng = np.random.default_rng(11)

    def cloud(n, lo, hi, normals):
        pc = g.PointCloud.from_numpy(rng.uniform(lo, hi, (n, 3)), rng.uniform(0, 1, (n, 3)), ctx=ctx)
        if normals:
            v = rng.normal(size=(n, 3))
            pc.set_normals(v / np.linalg.norm(v, axis=1, keepdims=True))
        return pc

    a, b = cloud(3000, -1.0, 2.0, True), cloud(5000, -3.0, 1.0, True)
    (ax, ar, ai), (bx, br, bi) = a.download(), b.download()
    an, bn = a.download_normals(), b.download_normals()
    a += b
    x, r, i = a.download()
    assert len(a) == 8000 and a.has_normals()
    np.testing.assert_array_equal(x, np.concatenate([ax, bx]))
    np.testing.assert_array_equal(r, np.concatenate([ar, br]))
    np.testing.assert_array_equal(i, np.concatenate([ai, bi]))
    np.testing.assert_array_equal(a.download_normals(), np.concatenate([an, bn]))
    mn, mx = O.bounds(np.concatenate([x, i[:, None]], 1))
    np.testing.assert_array_equal(a.get_min_bound(), mn)
    np.testing.assert_array_equal(a.get_max_bound(), mx)

    c = cloud(700, 0.5, 0.6, False)  # no normals: the sum has none (point_cloud.cu:64-70)
    cx, _, _ = c.download()
    a += c
    assert len(a) == 8700 and not a.has_normals()
    np.testing.assert_array_equal(a.download()[0], np.concatenate([ax, bx, cx]))

    d = cloud(1000, 2.0, 3.0, True)  # all coordinates positive: max bound from the points
    dx, _, _ = d.download()
    d += d
    np.testing.assert_array_equal(d.download()[0], np.concatenate([dx, dx]))
    assert d.has_normals()
    np.testing.assert_array_equal(d.get_max_bound(), dx.max(0))


def test_accumulation_step(frame):
    """One reconstruction step of apps/app.cpp:336-348 through the library: the model (the target
    frame voxelised at 1 cm) is moved by the frame pose, the source frame is appended and the sum is
    voxelised at 8 mm.  Bit-exact against the oracle at every stage."""
    T = frame.G
    model = frame.tgt.voxel_grid_down_sample(0.01)
    mx, mrgb, _ = O.voxel_downsample(frame.otx, frame.otr, 0.01)
    model.transform(T)
    tx, _ = O.transform(mx, T)
    x, _, i = model.download()
    np.testing.assert_array_equal(x, tx[:, :3])
    np.testing.assert_array_equal(i, tx[:, 3])
    model += frame.src
    ax, argb = O.accumulate(mx, mrgb, frame.osx, frame.osr, T)
    x, rgb, i = model.download()
    np.testing.assert_array_equal(x, ax[:, :3])
    np.testing.assert_array_equal(rgb, argb[:, :3])
    mn, mxb = O.bounds(ax)
    np.testing.assert_array_equal(model.get_min_bound(), mn)
    np.testing.assert_array_equal(model.get_max_bound(), mxb)
    out = model.voxel_grid_down_sample(0.008)
    vx, vrgb = O.accumulate(mx, mrgb, frame.osx, frame.osr, T, 0.008)
    x, rgb, i = out.download()
    assert len(x) == len(vx) > 1000
    np.testing.assert_array_equal(x, vx[:, :3])
    np.testing.assert_array_equal(i, vx[:, 3])
    np.testing.assert_array_equal(rgb, vrgb[:, :3])


def test_ply_ingest_and_self_registration(ctx):
    """data/frag_115.ply (every fifth vertex) through create_from_ply, then the reference's own
    example of apps/test_color_icp.cpp:399-410 on it: voxel 0.02, target normals at 0.05,
    color_icp(100, 0.05, 0.04), with the source the same fragment moved by a known pose.  Upload and
    transform bit-exact; the pose within 1e-5 of the oracle's."""
    import os

    import cuda_3d_amd as g
    from cuda_3d_amd.ply import read_ply

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frag_115_every5.ply")
    v = read_ply(path)
    tgt = g.PointCloud.create_from_ply(path, ctx=ctx)
    x, rgb, inten = tgt.download()
    assert not tgt.has_normals() and len(x) == len(v.xyz)
    np.testing.assert_array_equal(x, v.xyz)
    np.testing.assert_array_equal(rgb, v.rgb())
    with_n = g.PointCloud.create_from_ply(path, ctx=ctx, with_normals=True)
    np.testing.assert_array_equal(with_n.download_normals(), v.normals)

    txyzi = np.concatenate([x, inten[:, None]], 1)
    trgb = np.concatenate([rgb, np.zeros((len(x), 1), np.float32)], 1)
    K = np.eye(4)
    c, s_ = np.cos(0.03), np.sin(0.03)
    K[:3, :3] = [[c, -s_, 0], [s_, c, 0], [0, 0, 1]]
    K[:3, 3] = [0.02, -0.01, 0.015]
    src = tgt.copy()
    src.transform(K)
    sxyzi, _ = O.transform(txyzi, K)
    np.testing.assert_array_equal(src.download()[0], sxyzi[:, :3])

    ds, dt = src.voxel_grid_down_sample(0.02), tgt.voxel_grid_down_sample(0.02)
    dt.estimate_normals(0.05)
    icp = g.ColorICP(100, 0.05, 0.04, conv_rotation=0.0, conv_translation_sq=0.0)
    icp.set_source_point_cloud(ds)
    icp.set_target_point_cloud(dt)
    assert icp.align()
    T = icp.get_final_transformation_matrix()

    bmin, bmax = O.bounds(txyzi)
    osx, _, _ = O.voxel_downsample(sxyzi, trgb, 0.02)
    otx, _, _ = O.voxel_downsample(txyzi, trgb, 0.02)
    np.testing.assert_array_equal(dt.download()[0], otx[:, :3])
    on = O.estimate_normals(otx, 0.05, 0, bmin, bmax)
    og = O.color_gradient(otx, on, 0.04, 0, bmin, bmax)
    To, rep = O.icp_align(osx, otx, on, og, O.icp_params(100, 0.05, LAM, 0.0, 0.0), None, bmin, bmax)
    er, et = pose_err(T, To)
    kr, kt = pose_err(T, np.linalg.inv(K))
    print(f"[gpu-parity] frag_115 pose err vs oracle {er:.3e} rad {et:.3e} m; vs truth {kr:.3e} rad {kt:.3e} m",
          flush=True)
    assert er <= 1e-5 and et <= 1e-5
    assert kr < 1e-2 and kt < 1e-2


def _clusters_equal(got, want):
    assert len(got) == len(want)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, np.sort(b))


def test_euclidean_clustering_synthetic(ctx):
    """Connected components of the radius graph against the oracle's seeded search, on blobs,
    a 300-point chain (deep union-find trees), two bridged blobs and isolated points."""
    import cuda_3d_amd as g
    from test_oracle import cluster_scene

    tol = 0.02
    x = cluster_scene(3, tol)
    pc = g.PointCloud.from_numpy(x[:, :3], np.full((len(x), 3), 0.5), ctx=ctx)
    assert pc.euclidean_clustering(tol, 10, len(x)) == []  # no normals: empty (point_cloud.cu:290-295)
    pc.set_normals(np.tile([0.0, 0.0, 1.0], (len(x), 1)))
    xi = pc.download_xyzi()
    for mn in (0, 10, 100):
        print(f"[gpu-parity] clustering synthetic min {mn}", flush=True)
        _clusters_equal(pc.euclidean_clustering(tol, mn, len(x)), O.euclidean_clustering(xi, tol, mn, len(x)))
    labels, nc, no = pc.euclidean_cluster_labels(tol, 10, 500)
    whole = O.euclidean_clustering(xi, tol, 10, len(x))
    cut = O.euclidean_clustering(xi, tol, 10, 500)
    assert nc == len(cut) and no == sum(len(c) > 500 for c in whole) > 0
    lab = labels.cpu().numpy()
    for k, c in enumerate(cut):
        assert (lab[c] == k).all()
    assert (lab == -1).sum() == len(x) - sum(len(c) for c in cut)
    import ctypes as C

    host = np.full(len(x), -7, np.int32)
    hc, ho = C.c_int32(), C.c_int32()
    assert g._lib.lib.gca_euclidean_clustering_host(ctx.handle, pc.handle, tol, 10, 500, host.ctypes.data_as(C.c_void_p),
                                                    len(x), C.byref(hc), C.byref(ho)) == 0
    np.testing.assert_array_equal(host, lab)
    assert (hc.value, ho.value) == (nc, no)
    assert g._lib.lib.gca_euclidean_clustering_host(ctx.handle, pc.handle, tol, 10, 500, host.ctypes.data_as(C.c_void_p),
                                                    len(x) - 1, C.byref(hc), C.byref(ho)) != 0  # too small
    assert pc.euclidean_clustering(tol, 5, 4) == [] and pc.euclidean_clustering(0.0, 1, 10) == []


@pytest.mark.parametrize("mx", [500, 100, 37, 3, 1])
def test_euclidean_clustering_truncation(ctx, mx):
    """max_cluster_size below the largest components: the reference's searches stop at that
    size and leave the rest to later seeds (cuda_clustering.cu:75-96); the device replays them."""
    import cuda_3d_amd as g
    from test_oracle import cluster_scene

    tol = 0.02
    x = cluster_scene(3, tol)
    pc = g.PointCloud.from_numpy(x[:, :3], np.full((len(x), 3), 0.5), ctx=ctx)
    pc.set_normals(np.tile([0.0, 0.0, 1.0], (len(x), 1)))
    xi = pc.download_xyzi()
    for mn in (0, min(10, mx)):
        print(f"[gpu-parity] clustering truncation max {mx} min {mn}", flush=True)
        want = O.euclidean_clustering(xi, tol, mn, mx)
        got = pc.euclidean_clustering(tol, mn, mx)
        _clusters_equal(got, want)
        assert max(len(c) for c in want) >= mx


def test_residual_segmentation_synthetic(ctx):
    """Residual-based segmentation (cuda_clustering.cu:251-341) against the oracle, with residuals
    that cut many edges, whole and with max_cluster_size cutting components."""
    import cuda_3d_amd as g
    from test_oracle import cluster_scene

    tol = 0.02
    x = cluster_scene(5, tol)
    rng = np.random.default_rng(9)
    res = rng.choice(np.float32([0.0, 0.002, 0.004, 0.0055, 0.01]), len(x)).astype(np.float32)
    res += rng.uniform(0, 1e-4, len(x)).astype(np.float32)
    pc = g.PointCloud.from_numpy(x[:, :3], np.full((len(x), 3), 0.5), ctx=ctx)  # no normals needed
    xi = pc.download_xyzi()
    for mn, mx in ((0, len(x)), (5, len(x)), (5, 60), (1, 7)):
        print(f"[gpu-parity] residual segmentation synthetic min {mn} max {mx}", flush=True)
        _clusters_equal(pc.icp_residual_segmentation(tol, res, mn, mx),
                        O.euclidean_clustering(xi, tol, mn, mx, residual=res))
    assert pc.icp_residual_segmentation(0.0, res, 1, 10) == [] and pc.icp_residual_segmentation(tol, res, 5, 4) == []


def test_residual_segmentation_frame(level, gn_case):
    """Both frames: the 2 cm source moved by the test pose, its absolute residuals against the
    2 cm target (gca_compute_residuals, ABS kind), segmented at 3 cm, whole and cut at 300."""
    import cuda_3d_amd as g

    moved = gn_case["src"].copy()
    moved.transform(gn_case["T"])
    res, _ = g.compute_residuals(gn_case["src"], level["gpu"], gn_case["nn"], 1, LAM, gn_case["T"])
    res = np.asarray(res, np.float32)
    xi = moved.download_xyzi()
    for mn, mx in ((5, 10 ** 6), (5, 300)):
        print(f"[gpu-parity] residual segmentation frame n={len(xi)} min {mn} max {mx}", flush=True)
        got = moved.icp_residual_segmentation(0.03, res, mn, mx)
        _clusters_equal(got, O.euclidean_clustering(xi, 0.03, mn, mx, residual=res))
        assert len(got) >= 2


def test_euclidean_clustering_frame(level):
    """The 2 cm target of each frame with its normals (a copy: the shared cloud keeps its grid),
    tolerance 3 cm."""
    vc, ox = level["gpu"].copy(), level["ox"]
    assert vc.has_normals()
    print(f"[gpu-parity] clustering frame n={len(ox)}", flush=True)
    got = vc.euclidean_clustering(0.03, 20, 10 ** 6)
    _clusters_equal(got, O.euclidean_clustering(ox, 0.03, 20, 10 ** 6))
    assert len(got) >= 1
    # cut at a size far below the frame's surfaces
    got = vc.euclidean_clustering(0.03, 20, 1000)
    want = O.euclidean_clustering(ox, 0.03, 20, 1000)
    _clusters_equal(got, want)
    assert len(want) > len(O.euclidean_clustering(ox, 0.03, 20, 10 ** 6))


def test_frontend_edge_frames(ctx):
    import cuda_3d_amd as g

    cam = g.make_camera(64, 48, 60.0, 60.0, 32.0, 24.0)
    camo = O.make_camera(64, 48, 60.0, 60.0, 32.0, 24.0)
    bgr = np.full((48, 64, 3), 77, np.uint8)
    empty = np.zeros((48, 64), np.uint16)
    pc = g.PointCloud.create_from_rgbd(empty, bgr, cam, g.bilateral(), 0.2, 3.5, ctx=ctx)
    assert pc.points_number() == 0
    far = np.full((48, 64), 9000, np.uint16)  # beyond tmax
    assert g.PointCloud.create_from_rgbd(far, bgr, cam, g.bilateral(), 0.2, 3.5, ctx=ctx).points_number() == 0
    rng = np.random.default_rng(1)
    sparse = np.where(rng.random((48, 64)) < 0.3, rng.integers(300, 3400, (48, 64)), 0).astype(np.uint16)
    pc = g.PointCloud.create_from_rgbd(sparse, bgr, cam, g.bilateral(), 0.2, 3.5, ctx=ctx)
    ox, _, _, _ = O.cloud_from_rgbd(sparse, bgr, camo, O.Bilateral(1, 2, 2.0, 30.0), 0.2, 3.5)
    np.testing.assert_array_equal(pc.download()[0], ox[:, :3])
    bad = g.make_camera(64, 48, 60.0, 60.0, 32.0, 24.0, depth_scale=0.0)
    with pytest.raises(g.GcaError):
        g.PointCloud.create_from_rgbd(sparse, bgr, bad, g.bilateral(), 0.2, 3.5, ctx=ctx)


# ---- voxel ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("v", [0.04, 0.02, 0.01])
def test_voxel_bitexact(frame, v):
    vc = frame.tgt.voxel_grid_down_sample(v)
    xyz, rgb, inten = vc.download()
    ox, orgb, _ = O.voxel_downsample(frame.otx, frame.otr, v)
    assert len(xyz) == len(ox)
    np.testing.assert_array_equal(xyz, ox[:, :3])
    np.testing.assert_array_equal(rgb, orgb[:, :3])
    np.testing.assert_array_equal(inten, ox[:, 3])
    np.testing.assert_array_equal(vc.get_min_bound(), frame.bmin)  # output keeps input bounds


def test_voxel_edge_cases(ctx):
    import cuda_3d_amd as g

    one = g.PointCloud.from_numpy([[0.1, 0.2, 0.3]], [[0.5, 0.25, 0.125]], ctx=ctx)
    out = one.voxel_grid_down_sample(0.01)
    np.testing.assert_array_equal(out.download()[0], np.float32([[0.1, 0.2, 0.3]]))
    assert one.voxel_grid_down_sample(0.0) is None
    assert one.voxel_grid_down_sample(-1.0) is None
    empty = g.PointCloud.from_numpy(np.zeros((0, 3)), np.zeros((0, 3)), ctx=ctx)
    assert empty.voxel_grid_down_sample(0.02).points_number() == 0
    same = g.PointCloud.from_numpy(np.tile([[0.5, 0.5, 1.0]], (100, 1)), np.full((100, 3), 0.5), ctx=ctx)
    assert same.voxel_grid_down_sample(0.02).points_number() == 1
    huge = g.PointCloud.from_numpy([[0, 0, 0], [1e4, 1e4, 1e4]], np.zeros((2, 3)), ctx=ctx)
    assert huge.voxel_grid_down_sample(1e-6) is None  # over INT_MAX voxels per axis (point_cloud.cu:213-218)
    # repeated calls leave no state behind (bitmap and accumulators self-clear)
    a = same.voxel_grid_down_sample(0.02).download()[0]
    b = same.voxel_grid_down_sample(0.02).download()[0]
    np.testing.assert_array_equal(a, b)


def _voxel_vs_oracle(cloud, xyzi, rgb, v):
    vc = cloud.voxel_grid_down_sample(v)
    xyz, col, inten = vc.download()
    ox, orgb, _ = O.voxel_downsample(xyzi, rgb, v)
    assert len(xyz) == len(ox)
    np.testing.assert_array_equal(xyz, ox[:, :3])
    np.testing.assert_array_equal(col, orgb[:, :3])
    np.testing.assert_array_equal(inten, ox[:, 3])
    return len(ox)


@pytest.mark.parametrize("v", [0.04, 0.02, 0.01])
def test_voxel_sorted_path_bitexact(frame, ctx, v):
    """The radix-sort binning (tuning voxel=2) against the oracle on the frame; the default
    (bitmap) path on the same cloud is covered by test_voxel_bitexact."""
    with tuned(ctx, voxel=2):
        _voxel_vs_oracle(frame.tgt, frame.otx, frame.otr, v)


def _clustered_points(rng, n_vox, per, extent, v):
    """n_vox voxel centres in [0, extent)^3, `per` jittered points inside each voxel."""
    c = (np.floor(rng.uniform(0, extent / v, (n_vox, 3))) + 0.5) * v
    p = np.repeat(c, per, 0) + rng.uniform(-0.3 * v, 0.3 * v, (n_vox * per, 3))
    p = np.concatenate([p, [[0.0, 0.0, 0.0], [extent, extent, extent]]])  # pin the bounds
    rgb = rng.uniform(0, 1, (len(p), 3))
    order = rng.permutation(len(p))
    return p[order], rgb[order]


@pytest.mark.parametrize("case", ["48bit", "63bit", "90bit_split"])
def test_voxel_lattice_beyond_bitmap(ctx, case):
    """Lattices the occupancy bitmap cannot hold (over 2^34 cells) go through the sorted path with
    no refusal, as the reference's int3 sort takes any lattice under INT_MAX voxels per axis
    (point_cloud.cu:213-218): 2^48 cells (6 radix passes, 200k points over several sort tiles),
    2^63 (8 passes), and 2^90 (keys wider than 64 bits: sorted by (ky, kz), then kx).
    Bit-exact against the oracle, in either tuning."""
    import cuda_3d_amd as g

    rng = np.random.default_rng({"48bit": 1, "63bit": 2, "90bit_split": 3}[case])
    n_vox, per, extent, v = {"48bit": (50000, 4, 50.0, 0.001), "63bit": (3000, 3, 2000.0, 0.001),
                             "90bit_split": (3000, 3, 1000.0, 1e-6)}[case]
    p, rgb = _clustered_points(rng, n_vox, per, extent, v)
    pc = g.PointCloud.from_numpy(p, rgb, ctx=ctx)
    xyzi, col, inten = pc.download()
    ox = np.concatenate([xyzi, inten[:, None]], 1).astype(np.float32)
    orgb = np.concatenate([col, np.zeros((len(col), 1))], 1).astype(np.float32)
    print(f"[voxel-sort] {case}: {len(p)} points", flush=True)
    m = _voxel_vs_oracle(pc, ox, orgb, v)
    assert m >= n_vox * 0.99
    with tuned(ctx, voxel=1):  # bitmap asked for: the lattice still falls back to the sort
        _voxel_vs_oracle(pc, ox, orgb, v)


# ---- neighbour search -------------------------------------------------------------------------------

def test_grid_cells_bitexact(level, frame):
    c, o, s, d = level["gpu"].grid_cells(0.05)
    oc, oo, os_, od = O.grid_cells(level["ox"], 0.05, frame.bmin, frame.bmax)
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(o, oo)
    assert s == os_
    np.testing.assert_array_equal(d, od)


def test_nn_bitexact(level, frame):
    import cuda_3d_amd as g

    q = frame.src.voxel_grid_down_sample(0.04)
    oq, _, _ = O.voxel_downsample(frame.osx, frame.osr, 0.04)
    for r in (0.05, 0.02):
        idx = g.PointCloud.nn_search(q, level["gpu"], r)
        np.testing.assert_array_equal(idx, O.nn_search(oq, level["ox"], r, frame.bmin, frame.bmax))
    assert g.PointCloud.nn_search(q, level["gpu"], 0.0).size == 0  # point_cloud.cu:512-516


@pytest.mark.parametrize("K", [0, 30, 5])
def test_radius_search_bitexact(level, frame, K):
    import cuda_3d_amd as g

    offs, idx = g.PointCloud.radius_search(level["gpu"], level["gpu"], 0.05, K)
    ooffs, oidx = O.radius_search(level["ox"], level["ox"], 0.05, K, frame.bmin, frame.bmax)
    np.testing.assert_array_equal(offs, ooffs)
    np.testing.assert_array_equal(idx, oidx)


def test_outlier_removal_bitexact(frame):
    vc = frame.tgt.voxel_grid_down_sample(0.01)
    ox, orgb, _ = O.voxel_downsample(frame.otx, frame.otr, 0.01)
    out = vc.radius_outlier_removal(0.02, 4)
    oo, _ = O.radius_outlier_removal(ox, orgb, 0.02, 4, frame.bmin, frame.bmax)
    np.testing.assert_array_equal(out.download()[0], oo[:, :3])
    assert vc.radius_outlier_removal(0.0, 4).points_number() == 0


# ---- normals and colour gradient -----------------------------------------------------------------------

@pytest.mark.parametrize("K", [30, 0, 3])
def test_normals_parity(level, frame, K):
    c = level["gpu"].copy()
    assert c.estimate_normals(2.5 * level["v"], K)
    n = c.download_normals()
    on = O.estimate_normals(level["ox"], 2.5 * level["v"], K, frame.bmin, frame.bmax)[:, :3]
    nz = np.abs(on).sum(1) > 0
    assert np.array_equal(nz, np.abs(n).sum(1) > 0)
    assert angle(n[nz], on[nz]).max() <= 1e-5
    assert np.all((n[nz] * on[nz]).sum(1) > 0)  # same orientation
    assert np.array_equal(n, on), "normals are expected to be bit-exact"


def test_gradient_parity(level):
    gr = level["gpu"].download_gradient()
    og = level["og"][:, :3]
    scale = np.abs(og).max()
    np.testing.assert_allclose(gr, og, rtol=0, atol=1e-6 * scale)
    assert np.array_equal(gr, og), "colour gradients are expected to be bit-exact"


def test_gradient_requires_normals(ctx):
    import cuda_3d_amd as g

    pc = g.PointCloud.from_numpy(np.random.default_rng(0).uniform(0, 1, (50, 3)), np.full((50, 3), 0.5), ctx=ctx)
    assert not pc.compute_color_gradient(0.1)


# ---- Gauss-Newton system, residuals, solve -------------------------------------------------------------

@pytest.fixture(scope="module")
def gn_case(level, frame):
    import cuda_3d_amd as g

    src = frame.src.voxel_grid_down_sample(0.02)
    osx, _, _ = O.voxel_downsample(frame.osx, frame.osr, 0.02)
    T = frame.G.copy()
    T[:3, 3] += 0.003
    moved = src.copy()
    moved.transform(T)
    nn = g.PointCloud.nn_search(moved, level["gpu"], 0.05)
    return dict(src=src, osx=osx, T=T, nn=nn)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_build_gauss_newton(level, gn_case, mode):
    import cuda_3d_amd as g

    JTJ, JTr, r2, nc = g.build_gauss_newton(gn_case["src"], level["gpu"], gn_case["nn"], LAM, gn_case["T"], mode)
    o = O.build_gn(gn_case["osx"], level["ox"], level["on"], level["og"], gn_case["nn"], LAM, gn_case["T"])
    k = 0
    oJ = np.zeros((6, 6))
    for a in range(6):
        for b in range(a, 6):
            oJ[a, b] = oJ[b, a] = o[k]
            k += 1
    assert nc == int(o[28])
    tol = {0: 2e-6, 1: 1e-3, 2: 2e-6}[mode]  # naive FP32 accumulation is the ablation
    assert np.abs(JTJ - oJ).max() <= tol * np.abs(oJ).max()
    assert np.abs(JTr - o[21:27]).max() <= tol * np.abs(o[21:27]).max() + 1e-9
    assert abs(r2 - o[27]) <= tol * o[27]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_residuals(level, gn_case, kind):
    import cuda_3d_amd as g

    vals, s = g.compute_residuals(gn_case["src"], level["gpu"], gn_case["nn"], kind, LAM, gn_case["T"])
    ov, os_ = O.residuals(gn_case["osx"], level["ox"], level["on"], level["og"], gn_case["nn"], LAM, kind, gn_case["T"])
    np.testing.assert_allclose(vals, ov, rtol=1e-4, atol=1e-6)
    assert abs(s - os_) <= 1e-5 * max(1.0, abs(os_)) + 1e-6 * len(gn_case["nn"])


def test_solver_matches_oracle():
    import cuda_3d_amd as g

    rng = np.random.default_rng(2)
    A = rng.normal(size=(6, 6))
    A = A @ A.T + np.eye(6)
    b = rng.normal(size=6) * 1e-2
    st, xi, T = g.solve_gauss_newton(A, b)
    ost, oxi, oT = O.solve6([A[a, c] for a in range(6) for c in range(a, 6)], b)
    assert st == ost == 0
    np.testing.assert_allclose(xi, oxi, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(T, oT, atol=1e-15)


# ---- registration -------------------------------------------------------------------------------------

def icp_p(g, iters, corr, mode=0, conv=(0.0, 0.0)):
    return g.icp_params(iters, corr, LAM, mode, conv[0], conv[1])


def test_align_single_level(level, gn_case, frame):
    import cuda_3d_amd as g

    icp = g.ColorICP(30, 0.05, 0.04, conv_rotation=0.0, conv_translation_sq=0.0)
    icp.set_source_point_cloud(gn_case["src"])
    icp.set_target_point_cloud(level["gpu"])
    assert icp.align()
    T = icp.get_final_transformation_matrix()
    To, rep = O.icp_align(gn_case["osx"], level["ox"], level["on"], level["og"],
                          O.icp_params(30, 0.05, LAM, 0.0, 0.0), None, frame.bmin, frame.bmax)
    er, et = pose_err(T, To)
    print(f"[gpu-parity] align pose err vs oracle: {er:.3e} rad {et:.3e} m", flush=True)
    assert er <= 1e-5 and et <= 1e-5
    assert icp.report.iterations == rep.iterations == 30
    assert abs(icp.get_RSME() - rep.rmse) <= 1e-4 * rep.rmse
    res, mean = icp.get_abs_residual()
    assert len(res) == gn_case["src"].points_number() and mean > 0


def test_align_reference_stopping_rule(level, gn_case, frame):
    import cuda_3d_amd as g

    icp = g.ColorICP(100, 0.05, 0.04)  # reference thresholds (color_icp.hpp:47-48)
    icp.set_source_point_cloud(gn_case["src"])
    icp.set_target_point_cloud(level["gpu"])
    assert icp.align()
    To, rep = O.icp_align(gn_case["osx"], level["ox"], level["on"], level["og"],
                          O.icp_params(100, 0.05, LAM), None, frame.bmin, frame.bmax)
    assert icp.report.converged == rep.converged == 1
    assert icp.report.iterations == rep.iterations
    er, et = pose_err(icp.get_final_transformation_matrix(), To)
    assert er <= 1e-5 and et <= 1e-5


def test_align_preconditions(level, ctx):
    import cuda_3d_amd as g

    icp = g.ColorICP(10, 0.05, 0.04)
    assert not icp.align()  # no clouds
    icp.set_source_point_cloud(level["gpu"])
    bare = g.PointCloud.from_numpy(np.random.default_rng(0).uniform(0, 1, (50, 3)), np.full((50, 3), 0.5), ctx=ctx)
    icp.set_target_point_cloud(bare)
    assert not icp.align()  # target without normals (color_icp.cu:168-174)
    bad = g.ColorICP(10, 0.0, 0.04)
    bad.set_source_point_cloud(level["gpu"])
    bad.set_target_point_cloud(level["gpu"])
    assert not bad.align()


def test_multiscale_pose_parity(frame):
    import cuda_3d_amd as g

    lv = [g.level_params(v, icp=icp_p(g, 30, 2.5 * v)) for v in (0.04, 0.02, 0.01)]
    lo = [O.LevelParams(v, 2.5 * v, 30, 2 * v, 0, O.icp_params(30, 2.5 * v, LAM, 0.0, 0.0)) for v in (0.04, 0.02, 0.01)]
    T, reps = g.multiscale_icp(frame.src, frame.tgt, lv)
    To, oreps = O.multiscale_icp(frame.osx, frame.osr, frame.otx, frame.otr, lo)
    er, et = pose_err(T, To)
    gr, gt = pose_err(T, frame.G)
    print(f"[gpu-parity] multiscale vs oracle {er:.3e} rad {et:.3e} m; vs truth {gr:.3e} rad {gt:.3e} m", flush=True)
    assert er <= 1e-5 and et <= 1e-5
    assert [r.iterations for r in reps] == [30, 30, 30]
    assert gr < 5e-3 and gt < 5e-3  # registration itself succeeded


def test_accumulation_modes_pose_error(frame):
    """Precision variant: compensated FP32 partials vs naive FP32 ablation vs FP64."""
    import cuda_3d_amd as g

    lo = [O.LevelParams(v, 2.5 * v, 30, 2 * v, 0, O.icp_params(30, 2.5 * v, LAM, 0.0, 0.0)) for v in (0.04, 0.02, 0.01)]
    To, _ = O.multiscale_icp(frame.osx, frame.osr, frame.otx, frame.otr, lo)
    errs = {}
    for mode in (0, 1, 2):
        lv = [g.level_params(v, icp=icp_p(g, 30, 2.5 * v, mode)) for v in (0.04, 0.02, 0.01)]
        T, _ = g.multiscale_icp(frame.src, frame.tgt, lv)
        errs[mode] = pose_err(T, To)
    print(f"[gpu-parity] pose err by accumulation mode {errs}", flush=True)
    assert max(errs[0]) <= 1e-5 and max(errs[2]) <= 1e-5


@pytest.fixture(scope="module")
def uneven_pairs(ctx):
    """Four surface-sample pairs of uneven sizes (one degenerate), targets prepared."""
    import cuda_3d_amd as g
    from cuda_3d_amd import scene

    srcs, tgts = [], []
    for k, n in enumerate((3000, 20000, 7000, 500)):
        sx, sr, tx, tr, G = scene.sample_surface_pair(100 + k, n)
        s = g.PointCloud.from_numpy(sx, sr, ctx=ctx)
        t = g.PointCloud.from_numpy(tx, tr, ctx=ctx)
        t.estimate_normals(0.05, 30)
        t.compute_color_gradient(0.04, 0)
        srcs.append(s)
        tgts.append(t)
    return srcs, tgts


def rep_key(r):
    """Report fields that are results (latency_us and active_us are timing)."""
    d = r.as_dict()
    d.pop("latency_us")
    d.pop("active_us")
    return d


def single_align(src, tgt, iters=30, T0=None):
    import cuda_3d_amd as g

    icp = g.ColorICP(iters, 0.05, 0.04, conv_rotation=0.0, conv_translation_sq=0.0)
    icp.set_source_point_cloud(src)
    icp.set_target_point_cloud(tgt)
    assert icp.align(T0)
    return icp.get_final_transformation_matrix(), icp.report


class tuned:
    """Context manager: ctx tuning knobs set for the block, back to the heuristics after."""

    def __init__(self, ctx, **kw):
        self.ctx, self.kw = ctx, kw

    def __enter__(self):
        self.ctx.set_tuning(**self.kw)

    def __exit__(self, *a):
        self.ctx.set_tuning(**{k: 0 for k in self.kw})


# One launch shape for one-pair and batched calls (their defaults differ: lanes per query and the
# cap on work units, which fixes the partition and so the FP64 summation order).
SAME_SHAPE = dict(icp_lanes_small=4, icp_lanes_medium=4, icp_lanes_large=4, icp_lanes_batch=4, icp_max_units=512)


def test_batch_matches_single_pairs(ctx, uneven_pairs):
    import cuda_3d_amd as g

    srcs, tgts = uneven_pairs
    n = len(srcs)
    # one-pair and batched calls with the same lanes per query: each pair's work partition
    # depends on its own size only, so the poses are bit-identical
    with tuned(ctx, **SAME_SHAPE):
        Tb, reps = g.batch_icp(srcs, tgts, icp_p(g, 30, 0.05), [np.eye(4)] * n)
        singles = [single_align(srcs[k], tgts[k]) for k in range(n)]
    for k in range(n):
        T, rep = singles[k]
        np.testing.assert_array_equal(Tb[k], T)
        # a degenerate pair may stop early (failed solve -> identity step -> converged, as in
        # the reference); batched and single runs must agree on it
        assert (reps[k].iterations, reps[k].solver_failures, reps[k].n_correspondences) == \
               (rep.iterations, rep.solver_failures, rep.n_correspondences)
        assert reps[k].rmse == rep.rmse
    # default shapes (one lane per query when batched, wide groups for one pair): only the
    # order of the FP64 partial sums differs
    Tb, reps = g.batch_icp(srcs, tgts, icp_p(g, 30, 0.05), [np.eye(4)] * n)
    for k in range(n):
        T, rep = single_align(srcs[k], tgts[k])
        np.testing.assert_allclose(Tb[k], T, rtol=0, atol=1e-11)
        assert (reps[k].iterations, reps[k].n_correspondences) == (rep.iterations, rep.n_correspondences)


def test_launch_strategies_identical(ctx, uneven_pairs):
    """One persistent launch, one launch per iteration and one resident launch per pair (every
    workgroup finishing each iteration itself) run the same units and the same finish."""
    import cuda_3d_amd as g

    srcs, tgts = uneven_pairs
    p = icp_p(g, 12, 0.05)
    out = {}
    for launch in (1, 2, 3):
        print(f"[gpu-parity] launch strategy {launch}", flush=True)
        with tuned(ctx, icp_launch=launch):
            out[launch] = g.batch_icp(srcs, tgts, p, [np.eye(4)] * len(srcs))
    for launch in (2, 3):
        np.testing.assert_array_equal(out[1][0], out[launch][0])
        assert [rep_key(r) for r in out[1][1]] == [rep_key(r) for r in out[launch][1]]
    # one pair, three levels: resident launches level by level (the default) against one
    # persistent launch over the chain
    lv = [g.level_params(v, icp=icp_p(g, 10, 2.5 * v, conv=(1e-6, 1e-6))) for v in (0.04, 0.02)]
    res = {}
    for launch in (0, 1):
        with tuned(ctx, icp_launch=launch):
            res[launch] = g.multiscale_icp(srcs[1], tgts[1], lv)
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert [rep_key(r) for r in res[0][1]] == [rep_key(r) for r in res[1][1]]


def test_batch_plan_matches_batch_calls(ctx, uneven_pairs):
    """A batch plan's runs (eager, and a graph captured once then replayed) are bit-identical to
    one-shot batch calls: repeated runs, new initial poses, sources moved in place between runs.
    A plan whose target lost its gradient refuses to run."""
    import cuda_3d_amd as g
    from cuda_3d_amd import scene

    srcs0, tgts = uneven_pairs
    n = len(srcs0)
    p = icp_p(g, 30, 0.05)
    T0b = [scene.random_rigid(31 + k, 3.0, 0.02) for k in range(n)]
    M = scene.random_rigid(77, 2.0, 0.01)
    for graph in (False, True):
        srcs = [s.copy() for s in srcs0]
        rows_s, rows_t = [[s] for s in srcs], [[t] for t in tgts]
        want_a = g.batch_multiscale_icp(rows_s, rows_t, [p], [np.eye(4)] * n, [0.0])
        want_b = g.batch_multiscale_icp(rows_s, rows_t, [p], T0b, [0.0])
        plan = g.BatchPlan(rows_s, rows_t, [p], [np.eye(4)] * n, [0.0], graph=graph)
        with pytest.raises(g.GcaError):
            plan.results()  # no run yet
        for rep in range(3):
            print(f"[gpu-parity] batch plan graph={graph} run {rep}", flush=True)
            plan.enqueue()
            T, reps = plan.results()
            np.testing.assert_array_equal(T, want_a[0])
            assert [rep_key(r[0]) for r in reps] == [rep_key(r[0]) for r in want_a[1]]
        plan.set_initial(T0b)
        plan.enqueue()
        plan.enqueue()  # back to back on one stream: the later run's results
        T, reps = plan.results()
        np.testing.assert_array_equal(T, want_b[0])
        assert [rep_key(r[0]) for r in reps] == [rep_key(r[0]) for r in want_b[1]]
        for s in srcs:  # rewritten in place: the plan registers the moved points
            s.transform(M)
        want_c = g.batch_multiscale_icp(rows_s, rows_t, [p], T0b, [0.0])
        plan.enqueue()
        T, _ = plan.results()
        np.testing.assert_array_equal(T, want_c[0])
        assert not np.array_equal(want_c[0], want_b[0])
        assert plan.captures == (1 if graph else 0)
        plan.close()
    # a target whose gradient was recomputed: its gathered copy is stale, the plan refuses
    t = tgts[1].copy()
    plan = g.BatchPlan([[srcs0[1]]], [[t]], [p], [np.eye(4)], [0.0])
    plan.enqueue()
    plan.results()
    t.compute_color_gradient(0.04, 0)
    with pytest.raises(g.GcaError):
        plan.enqueue()
    plan.close()


def test_batch_plan_levels_with_skipped_level(ctx, uneven_pairs):
    """A two-level plan whose second level has max_iterations 0 (skipped: the chain ends at the
    first level) and a three-level plan, each against the one-shot call, bit for bit."""
    import cuda_3d_amd as g

    srcs0, tgts0 = uneven_pairs
    vs = (0.04, 0.02, 0.01)
    S, Tg = [], []
    for s, t in zip(srcs0, tgts0):
        S.append([s.voxel_grid_down_sample(v) for v in vs])
        row = []
        for v in vs:
            c = t.voxel_grid_down_sample(v)
            c.estimate_normals(2.5 * v, 30)
            c.compute_color_gradient(2.0 * v, 0)
            row.append(c)
        Tg.append(row)
    n = len(S)
    T0 = [np.eye(4)] * n
    for levels, iters in (((0, 1), (10, 0)), ((0, 1, 2), (10, 10, 10))):
        src = [[S[k][l] for l in levels] for k in range(n)]
        tgt = [[Tg[k][l] for l in levels] for k in range(n)]
        prm = [icp_p(g, it, 2.5 * vs[l], conv=(1e-6, 1e-6)) for l, it in zip(levels, iters)]
        gr = [2.0 * vs[l] for l in levels]
        want = g.batch_multiscale_icp(src, tgt, prm, T0, gr)
        for graph in (False, True):
            print(f"[gpu-parity] batch plan levels {levels} iterations {iters} graph={graph}", flush=True)
            plan = g.BatchPlan(src, tgt, prm, T0, gr, graph=graph)
            for _ in range(2):
                plan.enqueue()
                T, reps = plan.results()
                np.testing.assert_array_equal(T, want[0])
                assert [[rep_key(r) for r in row] for row in reps] == [[rep_key(r) for r in row] for row in want[1]]
            plan.close()
        if iters[-1] == 0:
            assert all(row[-1].iterations == 0 for row in want[1])
            assert any(row[0].iterations > 0 for row in want[1])


def test_batch_plan_in_caller_graph(ctx, uneven_pairs):
    """A plan's enqueue captured into the caller's own graph (torch's, on the ctx stream): each
    replay, with eager work and synchronisations between them, gives the one-shot call's poses."""
    import torch

    import cuda_3d_amd as g

    srcs, tgts = uneven_pairs
    n = len(srcs)
    p = icp_p(g, 30, 0.05)
    rows_s, rows_t = [[s] for s in srcs], [[t] for t in tgts]
    want = g.batch_multiscale_icp(rows_s, rows_t, [p], [np.eye(4)] * n, [0.0])
    plan = g.BatchPlan(rows_s, rows_t, [p], [np.eye(4)] * n, [0.0], graph=True)
    plan.enqueue()
    plan.results()  # warm: the plan's own graph
    assert plan.captures == 1
    graph = torch.cuda.CUDAGraph()
    stream = torch.cuda.ExternalStream(ctx.stream)
    with torch.cuda.graph(graph, stream=stream):
        plan.enqueue()  # the stream is capturing: plain launches into the caller's graph
    assert plan.captures == 1
    for rep in range(3):
        print(f"[gpu-parity] batch plan in a caller graph: replay {rep}", flush=True)
        x = torch.arange(1 << 16, device="cuda", dtype=torch.float32)
        assert float(x.sum()) == float((1 << 16) * ((1 << 16) - 1) // 2)
        graph.replay()
        torch.cuda.synchronize()
        T, reps = plan.results()
        np.testing.assert_array_equal(T, want[0])
        assert [rep_key(r[0]) for r in reps] == [rep_key(r[0]) for r in want[1]]
    del graph
    plan.close()


def test_batch_multiscale_matches_sequential(ctx, frame):
    """Chained levels in one launch == level-by-level calls carrying the pose on the host."""
    import cuda_3d_amd as g

    vs = (0.04, 0.02, 0.01)
    S, Tg = [], []
    for v in vs:
        s = frame.src.voxel_grid_down_sample(v)
        t = frame.tgt.voxel_grid_down_sample(v)
        t.estimate_normals(2.5 * v, 30)
        t.compute_color_gradient(2 * v, 0)
        S.append(s)
        Tg.append(t)
    params = [icp_p(g, 30, 2.5 * v, conv=(1e-6, 1e-6)) for v in vs]
    for launch in (1, 2):
        with tuned(ctx, icp_launch=launch, **SAME_SHAPE):
            Tb, reps = g.batch_multiscale_icp([S, S], [Tg, Tg], params, [np.eye(4), frame.G], [2 * v for v in vs])
            for k, T0 in enumerate((np.eye(4), frame.G)):
                T = T0
                for l, v in enumerate(vs):
                    icp = g.ColorICP(30, 2.5 * v, 2 * v)  # reference stopping rule
                    icp.set_source_point_cloud(S[l])
                    icp.set_target_point_cloud(Tg[l])
                    assert icp.align(T)
                    T = icp.get_final_transformation_matrix()
                    assert rep_key(reps[k][l]) == rep_key(icp.report)
                np.testing.assert_array_equal(Tb[k], T)
        print(f"[gpu-parity] chained levels (launch {launch}) iterations "
              f"{[[r.iterations for r in row] for row in reps]}", flush=True)


def test_warm_start_matches_full_search(level, gn_case):
    """Iteration k+1 of the device loop (nearest neighbours searched in a box shrunk by the
    previous match) against an independent full search + Gauss-Newton build + host solve at the
    pose iteration k ends with: same correspondences, hence the same system and step."""
    import cuda_3d_amd as g

    src, tgt = gn_case["src"], level["gpu"]
    ctx = src.ctx
    P = src.download()[0].astype(np.float64)
    for k in (0, 1, 2, 5):  # k = 0: the first iteration, searched without a previous match
        Tk, _ = single_align(src, tgt, k)
        Tk1, rep1 = single_align(src, tgt, k + 1)
        # the source mapped exactly as the device maps it: FP64 affine map, rounded once to FP32
        m = np.stack([((Tk[a, 0] * P[:, 0] + Tk[a, 1] * P[:, 1]) + Tk[a, 2] * P[:, 2]) + Tk[a, 3]
                      for a in range(3)], 1).astype(np.float32)
        moved = g.PointCloud.from_numpy(m, np.zeros_like(m), ctx=ctx)
        nn = g.PointCloud.nn_search(moved, tgt, 0.05)
        JTJ, JTr, r2, nc = g.build_gauss_newton(src, tgt, nn, LAM, Tk, 0)
        assert nc == rep1.n_correspondences
        assert abs(r2 / src.points_number() - rep1.rmse) <= 1e-12 * rep1.rmse
        _, _, dT = g.solve_gauss_newton(JTJ, JTr)
        np.testing.assert_allclose(dT @ Tk, Tk1, rtol=0, atol=1e-11)


@pytest.mark.parametrize("shape", [dict(icp_lanes_batch=16), dict(icp_lanes_batch=64),
                                   dict(icp_max_units=1), dict(icp_max_units=7, icp_persist_blocks=3)])
def test_shape_tuning_invariance(ctx, uneven_pairs, shape):
    """Launch shapes change only the FP64 summation order of the partials."""
    import cuda_3d_amd as g

    srcs, tgts = uneven_pairs
    p = icp_p(g, 10, 0.05)
    Tref, rref = g.batch_icp(srcs, tgts, p, [np.eye(4)] * len(srcs))
    with tuned(ctx, **shape):
        T, r = g.batch_icp(srcs, tgts, p, [np.eye(4)] * len(srcs))
    np.testing.assert_allclose(T, Tref, rtol=0, atol=1e-10)
    assert [x.n_correspondences for x in r] == [x.n_correspondences for x in rref]
    assert [x.iterations for x in r] == [x.iterations for x in rref]


@pytest.mark.parametrize("lanes", [1, 64])
def test_single_pair_lanes_invariance(ctx, uneven_pairs, lanes):
    srcs, tgts = uneven_pairs
    Tref, rref = single_align(srcs[1], tgts[1], 10)
    with tuned(ctx, icp_lanes_small=lanes, icp_lanes_medium=lanes, icp_lanes_large=lanes):
        T, r = single_align(srcs[1], tgts[1], 10)
    np.testing.assert_allclose(T, Tref, rtol=0, atol=1e-10)
    assert (r.iterations, r.n_correspondences) == (rref.iterations, rref.n_correspondences)


def test_determinism(frame):
    import cuda_3d_amd as g

    outs = []
    for _ in range(2):
        vc = frame.tgt.voxel_grid_down_sample(0.01)
        vc.estimate_normals(0.025, 30)
        vc.compute_color_gradient(0.02, 0)
        lv = [g.level_params(v, icp=icp_p(g, 10, 2.5 * v)) for v in (0.04, 0.02)]
        T, _ = g.multiscale_icp(frame.src, frame.tgt, lv)
        outs.append((vc.download()[0], vc.download_normals(), vc.download_gradient(), T))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


# ---- differentiation ---------------------------------------------------------------------------------

def test_objective_adjoint(level, gn_case):
    import cuda_3d_amd as g

    xi = np.array([0.002, -0.001, 0.0015, 0.001, 0.0005, -0.001])
    E, dxi, dp = g.objective_grad(gn_case["src"], level["gpu"], gn_case["nn"], gn_case["T"], xi, LAM)
    oE, odxi, odp = O.objective_grad(gn_case["osx"], level["ox"], level["on"], level["og"], gn_case["nn"], LAM,
                                     gn_case["T"], xi)
    assert abs(E - oE) <= 1e-9 * abs(oE)
    np.testing.assert_allclose(dxi, odxi, rtol=1e-9, atol=1e-12 * np.abs(odxi).max())
    np.testing.assert_allclose(dp.cpu().numpy(), odp, rtol=1e-5, atol=1e-6 * np.abs(odp).max())


def test_autograd_function_fd(level, gn_case):
    import torch

    import cuda_3d_amd as g
    from cuda_3d_amd.autograd import colored_icp_objective

    src_xyzi = torch.as_tensor(gn_case["src"].download_xyzi(), device="cuda").requires_grad_(True)
    nn = g.device_tensor(gn_case["nn"])
    rng = np.random.default_rng(3)
    worst = 0.0
    for k in range(8):
        xi0 = np.r_[rng.uniform(-0.01, 0.01, 3), rng.uniform(-0.01, 0.01, 3)]
        xi = torch.tensor(xi0, dtype=torch.float64, device="cuda", requires_grad=True)
        E = colored_icp_objective(xi, src_xyzi, level["gpu"], nn, gn_case["T"], LAM)
        E.backward()
        h = 1e-4
        fd = np.zeros(6)
        for a in range(6):
            e = np.zeros(6)
            e[a] = h
            Ep = colored_icp_objective(torch.tensor(xi0 + e, device="cuda"), src_xyzi.detach(), level["gpu"], nn, gn_case["T"], LAM)
            Em = colored_icp_objective(torch.tensor(xi0 - e, device="cuda"), src_xyzi.detach(), level["gpu"], nn, gn_case["T"], LAM)
            fd[a] = (Ep.item() - Em.item()) / (2 * h)
        rel = np.abs(xi.grad.cpu().numpy() - fd).max() / np.abs(fd).max()
        worst = max(worst, rel)
        src_xyzi.grad = None
    print(f"[gpu-parity] autograd vs central FD (h=1e-4) worst relative error {worst:.2e}", flush=True)
    assert worst < 1e-5


def test_cpp_facade_matches_python(ctx, tmp_path):
    """gca::color_icp through the C++ facade == ColorICP through the Python mirror (same ABI
    calls, same defaults, so the same bits), run as a separate process."""
    import json
    import os
    import subprocess

    import cuda_3d_amd as g
    from cuda_3d_amd import scene

    sx, sr, tx, tr, G = scene.sample_surface_pair(7, 6000)
    path = tmp_path / "clouds.bin"
    with open(path, "wb") as fh:
        np.array([len(sx), len(tx)], np.int64).tofile(fh)
        for a in (sx, sr, tx, tr):
            np.ascontiguousarray(a, np.float32).tofile(fh)
    exe = os.path.join(os.path.dirname(__file__), "cpp", "facade_icp")
    r = subprocess.run([exe, str(path), "0.05", "30", "0.05", "0.04", "30"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    s = g.PointCloud.from_numpy(sx, sr, ctx=ctx)
    t = g.PointCloud.from_numpy(tx, tr, ctx=ctx)
    t.estimate_normals(0.05, 30)
    icp = g.ColorICP(30, 0.05, 0.04)
    icp.set_source_point_cloud(s)
    icp.set_target_point_cloud(t)
    assert icp.align()
    np.testing.assert_array_equal(np.array(out["T"]).reshape(4, 4), icp.get_final_transformation_matrix())
    assert out["iterations"] == icp.report.iterations and out["n_src"] == len(sx) == out["moved_src"]
    # get_abs_residual_host, euclidean_clustering and icp_residual_segmentation through the facade
    res, mean = icp.get_abs_residual()
    np.testing.assert_array_equal(np.array(out["abs"], np.float32), res)
    assert abs(out["abs_mean"] - mean) <= 1e-6 * max(abs(mean), 1e-30)
    clusters = t.euclidean_clustering(0.02, 20, 2000)
    assert out["clusters"] == [[len(c), int(c[0])] for c in clusters] and len(clusters) > 0
    segments = icp.get_transformed_source_point_cloud().icp_residual_segmentation(0.02, res, 10, 1500)
    assert out["segments"] == [[len(c), int(c[0])] for c in segments] and len(segments) > 0


# ---- precision variant and differentiation at their stated sizes -----------------------------------

def test_precision_variant_1280x720(ctx):
    """1280x720 (~9e5 points), voxel pyramid down to 0.01 m (~2.5e5 points), 3-level colored
    ICP with 30 iterations per level: pose error against the FP64 oracle for compensated FP32
    partials (default), the naive-FP32 ablation, and FP64 partials."""
    import cuda_3d_amd as g
    from cuda_3d_amd import scene

    W, H = 1280, 720
    (td, tc), (sd, sc), G, cam = scene.frame_pair(13, W, H)
    camg = g.make_camera(W, H, cam.fx, cam.fy, cam.cx, cam.cy)
    camo = O.make_camera(W, H, cam.fx, cam.fy, cam.cx, cam.cy)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    tgt = g.PointCloud.create_from_rgbd(td, tc, camg, bil, 0.2, 3.5, ctx=ctx)
    src = g.PointCloud.create_from_rgbd(sd, sc, camg, bil, 0.2, 3.5, ctx=ctx)
    vs = (0.04, 0.02, 0.01)
    lo = [O.LevelParams(v, 2.5 * v, 30, 2 * v, 0, O.icp_params(30, 2.5 * v, LAM, 0.0, 0.0)) for v in vs]
    otx, otr, _, _ = O.cloud_from_rgbd(td, tc, camo, O.Bilateral(1, 2, 2.0, 30.0), 0.2, 3.5)
    osx, osr, _, _ = O.cloud_from_rgbd(sd, sc, camo, O.Bilateral(1, 2, 2.0, 30.0), 0.2, 3.5)
    print(f"[gpu-parity] 1280x720: {src.points_number()} / {tgt.points_number()} points; oracle running", flush=True)
    To, _ = O.multiscale_icp(osx, osr, otx, otr, lo)
    errs = {}
    for mode, name in ((0, "compensated"), (1, "naive_fp32"), (2, "fp64")):
        lv = [g.level_params(v, icp=icp_p(g, 30, 2.5 * v, mode)) for v in vs]
        T, reps = g.multiscale_icp(src, tgt, lv)
        errs[name] = pose_err(T, To)
    fine = tgt.voxel_grid_down_sample(0.01).points_number()
    print(f"[gpu-parity] 1280x720 finest level {fine} points; pose error vs FP64 oracle (rad, m): {errs}", flush=True)
    assert max(errs["compensated"]) <= 1e-5 and max(errs["fp64"]) <= 1e-5
    assert all(np.isfinite(errs["naive_fp32"]))
    # The accumulation error itself: the Gauss-Newton system of the finest level at the oracle's
    # final pose, where J^T r nearly cancels, built in each mode and compared with the oracle's
    # FP64 build over the same correspondences.
    v = 0.01
    sf, tf = src.voxel_grid_down_sample(v), tgt.voxel_grid_down_sample(v)
    tf.estimate_normals(2.5 * v, 30)
    tf.compute_color_gradient(2.0 * v, 0)
    osf, _, _ = O.voxel_downsample(osx, osr, v)
    otf, _, _ = O.voxel_downsample(otx, otr, v)
    on = O.estimate_normals(otf, 2.5 * v, 30)
    og = O.color_gradient(otf, on, 2.0 * v, 0)
    moved = sf.copy()
    moved.transform(To)
    nn = g.PointCloud.nn_search(moved, tf, 2.5 * v)
    o = O.build_gn(osf, otf, on, og, nn, LAM, To)
    oJ = np.zeros((6, 6))
    k = 0
    for a in range(6):
        for b in range(a, 6):
            oJ[a, b] = oJ[b, a] = o[k]
            k += 1
    sys_err = {}
    for mode, name in ((0, "compensated"), (1, "naive_fp32"), (2, "fp64")):
        JTJ, JTr, r2, nc = g.build_gauss_newton(sf, tf, nn, LAM, To, mode)
        assert nc == int(o[28])
        sys_err[name] = (float(np.abs(JTJ - oJ).max() / np.abs(oJ).max()),
                         float(np.abs(JTr - o[21:27]).max() / np.abs(o[21:27]).max()),
                         float(abs(r2 - o[27]) / o[27]))
    print(f"[gpu-parity] 1280x720 finest-level system at the final pose, {int(o[28])} correspondences; "
          f"relative error vs FP64 (JTJ, JTr, r^2): {sys_err}", flush=True)
    # J^T r at the solution is a sum that cancels: its error is dominated by the FP32 rounding of
    # the per-point terms in every mode, so it is reported, and only J^T J and r^2 are bounded
    for name in ("compensated", "fp64"):
        assert sys_err[name][0] <= 1e-5 and sys_err[name][2] <= 1e-5, (name, sys_err[name])


def test_differentiation_100k_32_poses(ctx):
    """Adjoint of the colored-ICP objective on a 1e5-point pair at 32 seeded poses (rotation
    <= 10 deg, translation <= 0.1 m), through the torch.autograd.Function: dE/dxi against central
    differences (h = 1e-4) and the FP64 analytic gradient of the oracle; dE/d(x,y,z,I) against
    central differences on sampled points."""
    import torch

    import cuda_3d_amd as g
    from cuda_3d_amd import scene
    from cuda_3d_amd.autograd import colored_icp_objective

    sx, sr, tx, tr, G = scene.sample_surface_pair(31, 100000)
    src = g.PointCloud.from_numpy(sx, sr, ctx=ctx)
    tgt = g.PointCloud.from_numpy(tx, tr, ctx=ctx)
    tgt.estimate_normals(0.03, 30)
    tgt.compute_color_gradient(0.03, 0)
    pad = lambda a: np.concatenate([a, np.zeros((len(a), 1), np.float32)], 1)  # noqa: E731  (float4 rows)
    ox, on, og = tgt.download_xyzi(), pad(tgt.download_normals()), pad(tgt.download_gradient())
    osx = src.download_xyzi()
    pts = torch.as_tensor(osx, device="cuda")
    work = g.PointCloud(ctx, len(osx))
    rng = np.random.default_rng(2031)
    worst = {"xi_fd": 0.0, "xi_fp64": 0.0, "pts_fd": 0.0}
    h = 1e-4
    for k in range(32):
        T = scene.random_rigid(5000 + k, 10.0, 0.1) @ G
        moved = src.copy()
        moved.transform(T)
        nn = g.PointCloud.nn_search(moved, tgt, 0.1)
        nn_dev = g.device_tensor(nn)
        xi = torch.zeros(6, dtype=torch.float64, device="cuda", requires_grad=True)
        p = pts.clone().requires_grad_(True)
        E = colored_icp_objective(xi, p, tgt, nn_dev, T, LAM, work)
        E.backward()
        gxi, gp = xi.grad.cpu().numpy(), p.grad.cpu().numpy()
        # FP64 analytic gradient of the oracle at the same pose and correspondences
        oE, odxi, _ = O.objective_grad(osx, ox, on, og, nn, LAM, T, np.zeros(6))
        worst["xi_fp64"] = max(worst["xi_fp64"], np.abs(gxi - odxi).max() / np.abs(odxi).max())
        fd = np.zeros(6)
        for a in range(6):
            e = np.zeros(6)
            e[a] = h
            Ep = colored_icp_objective(torch.tensor(e, device="cuda"), pts, tgt, nn_dev, T, LAM, work).item()
            Em = colored_icp_objective(torch.tensor(-e, device="cuda"), pts, tgt, nn_dev, T, LAM, work).item()
            fd[a] = (Ep - Em) / (2 * h)
        worst["xi_fd"] = max(worst["xi_fd"], np.abs(gxi - fd).max() / np.abs(fd).max())
        # per-point gradients: central differences on 4 matched points, each coordinate
        matched = np.nonzero(np.asarray(nn) >= 0)[0]
        for i in rng.choice(matched, 4, replace=False):
            for c in range(4):
                q = pts.clone()
                base = float(osx[i, c])
                up, dn = np.float32(base + h), np.float32(base - h)
                q[i, c] = float(up)
                Ep = colored_icp_objective(torch.zeros(6, dtype=torch.float64, device="cuda"), q, tgt, nn_dev, T,
                                           LAM, work).item()
                q[i, c] = float(dn)
                Em = colored_icp_objective(torch.zeros(6, dtype=torch.float64, device="cuda"), q, tgt, nn_dev, T,
                                           LAM, work).item()
                fdp = (Ep - Em) / (float(up) - float(dn))
                scale = max(abs(fdp), 1e-3 * np.abs(gp[i]).max(), 1e-12)
                worst["pts_fd"] = max(worst["pts_fd"], abs(gp[i, c] - fdp) / scale)
    print(f"[gpu-parity] adjoint on 1e5 points, 32 poses: worst relative error {worst}", flush=True)
    assert worst["xi_fp64"] < 1e-8
    assert worst["xi_fd"] < 1e-5
    assert worst["pts_fd"] < 1e-3


# ---- native frame pipeline and its graph replay ---------------------------------------------------

def _pipeline_frames(g, W, H, seed):
    from cuda_3d_amd import scene

    (td, tc), (sd, sc), G, cm = scene.frame_pair(seed, W, H)
    cam = g.make_camera(W, H, cm.fx, cm.fy, cm.cx, cm.cy)
    dev = [(g.device_tensor(td.view(np.int16)), g.device_tensor(tc)),
           (g.device_tensor(sd.view(np.int16)), g.device_tensor(sc))]
    return cam, dev, ((td, tc), (sd, sc)), G


def test_pipeline_matches_composed_calls(ctx, frame):
    """gca_pipeline (device-side voxel lattice, frustum grids, one chained launch) gives the
    poses of the same frames registered through gca_multiscale_icp, bit for bit, and matches the
    oracle."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    W, H = (320, 240) if frame.camg.width == 320 else (640, 480)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    pipe = FramePipeline(ctx, frame.camg, bil, graph=False)
    dt = g.device_tensor(frame.td.view(np.int16))
    ct = g.device_tensor(frame.tc)
    ds = g.device_tensor(frame.sd.view(np.int16))
    cs = g.device_tensor(frame.sc)
    pipe.step(dt.data_ptr(), ct.data_ptr())
    assert pipe.result()[0] is None  # first frame: nothing to register against
    pipe.step(ds.data_ptr(), cs.data_ptr())  # source onto target
    T, reps = pipe.result()
    lv = [g.level_params(v, icp=g.icp_params(30, 2.5 * v)) for v in (0.04, 0.02, 0.01)]
    Tm, mreps = g.multiscale_icp(frame.src, frame.tgt, lv)
    np.testing.assert_array_equal(T, Tm)
    assert [rep_key(r) for r in reps] == [rep_key(r) for r in mreps]
    lo = [O.LevelParams(v, 2.5 * v, 30, 2 * v, 0, O.icp_params(30, 2.5 * v, LAM, 1e-6, 1e-6)) for v in (0.04, 0.02, 0.01)]
    To, oreps = O.multiscale_icp(frame.osx, frame.osr, frame.otx, frame.otr, lo)
    er, et = pose_err(T, To)
    print(f"[gpu-parity] pipeline {W}x{H} vs oracle {er:.3e} rad {et:.3e} m; iterations "
          f"{[r.iterations for r in reps]} (oracle {[r.iterations for r in oreps]})", flush=True)
    assert er <= 1e-5 and et <= 1e-5
    assert [r.iterations for r in reps] == [r.iterations for r in oreps]
    full_n, lv_n = pipe.sizes(1)
    assert full_n == frame.src.points_number()
    pipe.close()


def test_pipeline_graph_replays_match_eager(ctx):
    """Each frame parity's step captured once and replayed: every replay's pose equals the
    eager pipeline's, with unrelated library work and synchronisations between replays."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    cam, dev, _, _ = _pipeline_frames(g, 320, 240, 21)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    eager = FramePipeline(ctx, cam, bil, graph=False)
    graph = FramePipeline(ctx, cam, bil, graph=True)
    other = g.PointCloud.from_numpy(np.random.default_rng(5).uniform(0, 1, (2000, 3)), np.full((2000, 3), 0.5), ctx=ctx)
    for k in range(12):
        d, c = dev[k % 2]
        eager.step(d.data_ptr(), c.data_ptr())
        Te, re = eager.result()
        graph.step(d.data_ptr(), c.data_ptr())
        Tg, rg = graph.result()
        if k == 0:
            assert Te is None and Tg is None
            continue
        np.testing.assert_array_equal(Tg, Te)
        assert [rep_key(r) for r in rg] == [rep_key(r) for r in re]
        # unrelated work on another ctx and stream between replays
        other.voxel_grid_down_sample(0.05)
        ctx.synchronize()
    eager.close()
    graph.close()


@pytest.mark.parametrize("frontend", [0, 1])
def test_pipeline_graph_follows_new_frame_buffers(ctx, frontend):
    """Frames arriving in four different device buffers (each frame of the pair in two copies,
    used in turn): the replayed graphs follow the new buffers through a node update, so exactly
    two captures happen (one per frame parity), and every pose equals the eager pipeline's.
    frontend 0: the two-kernel front end; 1: the one-pass kernel (its node carries other
    arguments)."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    cam, dev, _, _ = _pipeline_frames(g, 320, 240, 23)
    copies = [(d.clone(), c.clone()) for d, c in dev]
    bufs = [dev[0], dev[1], copies[0], copies[1]]  # frame A, B, A, B at four addresses
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    eager = FramePipeline(ctx, cam, bil, graph=False)
    graph = FramePipeline(ctx, cam, bil, graph=True)
    if frontend:
        eager.ctx.set_tuning(frontend=frontend)
        graph.ctx.set_tuning(frontend=frontend)
    for k in range(14):
        d, c = bufs[k % 4] if k < 8 else bufs[(k * 3) % 4]
        eager.step(d.data_ptr(), c.data_ptr())
        Te, re = eager.result()
        graph.step(d.data_ptr(), c.data_ptr())
        Tg, rg = graph.result()
        if k == 0:
            assert Te is None and Tg is None
            continue
        np.testing.assert_array_equal(Tg, Te)
        assert [rep_key(r) for r in rg] == [rep_key(r) for r in re]
    captures, updates = graph.graph_stats()
    print(f"[gpu-parity] rotating frame buffers: {captures} captures, {updates} node updates", flush=True)
    assert captures == 2 and updates >= 6
    eager.close()
    graph.close()


def test_pipeline_tuning_change_rebuilds(ctx):
    """Changing the pipeline ctx's tuning drops the plans and graphs; the following steps run
    eagerly, capture again, and keep agreeing with an untouched pipeline to the last bits."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    cam, dev, _, _ = _pipeline_frames(g, 320, 240, 22)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    ref = FramePipeline(ctx, cam, bil, graph=False)
    pipe = FramePipeline(ctx, cam, bil, graph=True)
    for k in range(10):
        if k == 5:
            pipe.ctx.set_tuning(icp_lanes_small=4, icp_lanes_medium=4, icp_lanes_large=4)
        d, c = dev[k % 2]
        ref.step(d.data_ptr(), c.data_ptr())
        Tr, rr = ref.result()
        pipe.step(d.data_ptr(), c.data_ptr())
        Tp, rp = pipe.result()
        if k == 0:
            continue
        np.testing.assert_allclose(Tp, Tr, rtol=0, atol=1e-10)
        assert [r.iterations for r in rp] == [r.iterations for r in rr]
    ref.close()
    pipe.close()


def test_pipeline_grid_cell_change_mid_sequence(ctx):
    """Regression: a registration plan reads its target's grid through the geometry and arrays
    it saw when it was made.  Changing the grid's cell edge between steps (as the autotuner does)
    once left the plan made at the change with the old geometry while later steps rebuilt its
    target with the new one, an inconsistent cell table.  Now the previous frame's grids are
    re-binned into the new geometry before the plan is made (and a plan whose target grid has
    changed under it is refused, never launched).  The neighbour results, and so the poses, do
    not depend on the cell edge: every step equals an untouched pipeline's, bit for bit."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    cam, dev, _, _ = _pipeline_frames(g, 320, 240, 24)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    ref = FramePipeline(ctx, cam, bil, graph=False)
    pipe = FramePipeline(ctx, cam, bil, graph=True)
    cells = {3: 5, 6: 8, 9: 4, 12: 6, 15: 0}
    for k in range(18):
        if k in cells:
            print(f"[gpu-parity] step {k}: grid_cell {cells[k]}", flush=True)
            pipe.ctx.set_tuning(grid_cell=cells[k])
        d, c = dev[k % 2]
        ref.step(d.data_ptr(), c.data_ptr())
        Tr, rr = ref.result()
        pipe.step(d.data_ptr(), c.data_ptr())
        Tp, rp = pipe.result()
        if k == 0:
            continue
        np.testing.assert_array_equal(Tp, Tr)
        assert [rep_key(r) for r in rp] == [rep_key(r) for r in rr]
    ref.close()
    pipe.close()


def test_pipeline_overlap_tuning_changes(ctx):
    """Two-lane overlapped steps across tuning changes (grid cell edge, lanes per query): a change
    synchronises both lanes, drops the plans and graphs, and the following steps run eagerly and
    capture again.  Every pose still equals the sequential pipeline's one step later: to the last
    bits for the cell edge and the normals' lanes (exact neighbour sets and fixed-point moments),
    within 1e-10 for the ICP's lanes (they only reorder its FP64 partial sums), with the same
    iteration counts.  Poses are read every third step, so steps are in flight across a change."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    cam, dev, _, _ = _pipeline_frames(g, 320, 240, 24)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    ref = FramePipeline(ctx, cam, bil, graph=False)
    ovl = FramePipeline(ctx, cam, bil, graph=True, overlap=True)
    assert ovl.slots == 4
    changes = {4: dict(grid_cell=5), 9: dict(normals_lanes=4, icp_lanes_small=2, icp_lanes_medium=2),
               14: dict(grid_cell=4, normals_lanes=8)}
    want = []
    for k in range(20):
        if k in changes:
            print(f"[gpu-parity] step {k}: {changes[k]}", flush=True)
            ovl.ctx.set_tuning(**changes[k])
        d, c = dev[k % 2]
        ref.step(d.data_ptr(), c.data_ptr())
        Tr, rr = ref.result()
        want.append((Tr, [r.iterations for r in rr]))  # the reports are views of the pipeline's buffer
        ovl.step(d.data_ptr(), c.data_ptr())
        if k % 3 != 2:
            continue
        To, ro = ovl.result()
        np.testing.assert_allclose(To, want[k - 1][0], rtol=0, atol=1e-10)
        assert [r.iterations for r in ro] == want[k - 1][1]
    ovl.flush()
    To, ro = ovl.result()
    np.testing.assert_allclose(To, want[-1][0], rtol=0, atol=1e-10)
    assert [r.iterations for r in ro] == want[-1][1]
    ref.close()
    ovl.close()


def test_pipeline_overlap_graph_toggled(ctx):
    """Two-lane overlapped steps with graph replay switched off (steps in flight: the graphs are
    dropped after the lanes drain) and on again: every pose still equals the sequential
    pipeline's one step later, to the last bits, with the same iteration counts."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    cam, dev, _, _ = _pipeline_frames(g, 320, 240, 31)
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    ref = FramePipeline(ctx, cam, bil, graph=False)
    ovl = FramePipeline(ctx, cam, bil, graph=True, overlap=True)
    want = []
    for k in range(18):
        if k in (7, 12):
            print(f"[gpu-parity] step {k}: graph {'off' if k == 7 else 'on'}", flush=True)
            ovl.set_graph(k == 12)
        d, c = dev[k % 2]
        ref.step(d.data_ptr(), c.data_ptr())
        Tr, rr = ref.result()
        want.append((Tr, [r.iterations for r in rr]))
        ovl.step(d.data_ptr(), c.data_ptr())
        if k % 4 != 3:
            continue
        To, ro = ovl.result()
        np.testing.assert_array_equal(To, want[k - 1][0])
        assert [r.iterations for r in ro] == want[k - 1][1]
    assert ovl.graph_stats()[0] > 0
    ovl.flush()
    To, ro = ovl.result()
    np.testing.assert_array_equal(To, want[-1][0])
    ref.close()
    ovl.close()


def test_pipeline_rejects_bad_arguments(ctx):
    import ctypes as C

    import cuda_3d_amd as g
    from cuda_3d_amd import _lib as L

    cam = g.make_camera(64, 48, 60.0, 60.0, 32.0, 24.0)
    bil = g.bilateral()
    lv = (L.LevelParams * 1)(g.level_params(0.02))
    h = C.c_void_p()
    assert L.lib.gca_pipeline_create(ctx.handle, C.byref(cam), C.byref(bil), 0.2, 3.5, 0, lv, C.byref(h)) == L.GCA_ERR_INVALID_VALUE
    assert L.lib.gca_pipeline_create(ctx.handle, C.byref(cam), C.byref(bil), 3.5, 0.2, 1, lv, C.byref(h)) == L.GCA_ERR_INVALID_VALUE
    bad = (L.LevelParams * 1)(g.level_params(0.0))
    assert L.lib.gca_pipeline_create(ctx.handle, C.byref(cam), C.byref(bil), 0.2, 3.5, 1, bad, C.byref(h)) == L.GCA_ERR_INVALID_VALUE
    assert L.lib.gca_pipeline_create(ctx.handle, C.byref(cam), C.byref(bil), 0.2, 3.5, 1, lv, C.byref(h)) == 0
    assert L.lib.gca_pipeline_step(h, None, None) == L.GCA_ERR_INVALID_VALUE
    T = np.zeros(16)
    valid = C.c_int32(7)
    assert L.lib.gca_pipeline_result(h, T.ctypes.data_as(C.c_void_p), None, C.byref(valid)) == 0
    assert valid.value == 0 and np.array_equal(T.reshape(4, 4), np.eye(4))
    assert L.lib.gca_pipeline_destroy(h) == 0



@pytest.mark.parametrize("graph,size,levels", [(False, (320, 240), 3), (True, (320, 240), 3), (True, (640, 480), 3),
                                               (False, (320, 240), 5), (True, (320, 240), 5)])
def test_pipeline_overlap_matches_sequential(ctx, graph, size, levels):
    """Overlapped steps (the previous frame registered beside the new frame's build) give every
    pose of the sequential pipeline, bit for bit, one step later, and every frame's level sizes;
    flush() registers the last frame.  Three levels: lockstep builds in two lanes, four frame
    slots; five levels (beyond one multi-level launch): one lane, per-level streams, three slots.
    Poses are read only every few steps, so that several steps (both lanes) are in flight
    between reads.  Frames from six buffers, so the replayed graphs also follow new frame
    buffers."""
    import cuda_3d_amd as g
    from cuda_3d_amd.pipeline import FramePipeline

    vox = (0.04, 0.02, 0.01) if levels == 3 else (0.08, 0.04, 0.03, 0.02, 0.01)
    cam, dev, _, _ = _pipeline_frames(g, size[0], size[1], 26)
    bufs = list(dev) + [(d.clone(), c.clone()) for _ in range(2) for d, c in dev]  # frame k % 2 in bufs[k % 6]
    bil = g.bilateral(g.BILATERAL_FIXED, 2, 2.0, 30.0)
    seq = FramePipeline(ctx, cam, bil, vox, graph=False)
    ovl = FramePipeline(ctx, cam, bil, vox, graph=graph, overlap=True)
    slots = 4 if levels == 3 else 3
    assert ovl.slots == slots
    want, sizes = [], []
    for k in range(16):
        d, c = bufs[k % 6]
        seq.step(d.data_ptr(), c.data_ptr())
        Ts, rs = seq.result()
        want.append((Ts, [rep_key(r) for r in rs]) if Ts is not None else None)
        sizes.append(seq.sizes(k % 2))
    checked = 0
    for k in range(16):
        d, c = bufs[k % 6]
        ovl.step(d.data_ptr(), c.data_ptr())
        if k not in (0, 1, 2, 5, 6, 7, 11, 15):
            continue
        To, ro = ovl.result()
        for j in range(max(0, k - slots + 1), k + 1):  # the slots still holding frames j
            assert ovl.sizes(j % slots) == sizes[j]
        if k < 2:
            assert To is None  # nothing registered yet: frame k - 1 has no predecessor for k < 2
            continue
        np.testing.assert_array_equal(To, want[k - 1][0])
        assert [rep_key(r) for r in ro] == want[k - 1][1]
        checked += 1
    ovl.flush()
    To, ro = ovl.result()
    np.testing.assert_array_equal(To, want[-1][0])
    assert [rep_key(r) for r in ro] == want[-1][1]
    captures, updates = ovl.graph_stats()
    print(f"[gpu-parity] overlapped steps {size[0]}x{size[1]}, {levels} levels (graph {graph}): {checked + 1} poses "
          f"checked, {captures} captures, {updates} node updates", flush=True)
    if graph:  # two lanes: a build graph and a registration graph per slot
        assert captures == (2 * slots if slots == 4 else slots) and updates >= 3
    seq.close()
    ovl.close()
