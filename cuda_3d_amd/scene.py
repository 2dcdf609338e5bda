"""Seeded synthetic RGB-D scenes: 3 planes + 4 spheres at 0.5-3 m.

Stands in for the reference's RealSense source (src/camera/realsense_device.cpp:87-120): it
produces Z16 depth (millimetres, scale 0.001) and BGR8 colour frames of the same layout,
rendered through the reference's pixel convention (ray through (x - 0.5, y - 0.5),
src/geometry/cuda_point_cloud_factory.cu:232).

Depth noise is Gaussian (sigma 2 mm) then rounded to millimetres, 5 % of pixels are zero
holes, and colour is a per-object base colour plus uniform integer noise in [-10, 10].
A source frame is rendered from a camera displaced by a seeded rigid transform G (rotation
axis uniform on the sphere, angle uniform in [0, max_angle], translation uniform in a ball);
G maps source-camera coordinates to target-camera coordinates, so G is the pose that
registration should recover.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class CameraModel:
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    depth_scale: float = 0.001


def camera_for(width: int, height: int) -> CameraModel:
    """Intrinsics of the configurations: 320x240 f=307.5, 640x480 f=615, 1280x720 f=922.5."""
    f = {320: 307.5, 640: 615.0, 1280: 922.5}.get(width, 0.9609375 * width)
    return CameraModel(width, height, f, f, width / 2.0, height / 2.0)


@dataclass
class Scene:
    planes: list      # (unit normal n, offset d, base BGR) with n . X = d
    spheres: list     # (centre, radius, base BGR)
    background: tuple


def make_scene(seed: int, scale: float = 1.0) -> Scene:
    """Seeded scene; `scale` multiplies every distance (a camera at the origin sees the same
    image, `scale` times farther away, over `scale`^2 times the surface area)."""
    rng = np.random.default_rng(np.random.PCG64(seed))

    def colour():
        return tuple(int(v) for v in rng.integers(30, 226, size=3))

    def tilt(n, amount=0.15):
        v = np.asarray(n, np.float64) + rng.uniform(-amount, amount, 3)
        return v / np.linalg.norm(v)

    planes = [
        (tilt([0.0, 0.0, 1.0]), 2.8 + rng.uniform(-0.1, 0.1), colour()),   # back wall
        (tilt([0.0, 1.0, 0.0]), 0.9 + rng.uniform(-0.1, 0.1), colour()),   # floor (y down)
        (tilt([1.0, 0.0, 0.0]), -1.1 + rng.uniform(-0.1, 0.1), colour()),  # left wall
    ]
    spheres = []
    for _ in range(4):
        z = rng.uniform(1.0, 2.2)
        c = np.array([rng.uniform(-0.45, 0.45) * z, rng.uniform(-0.3, 0.3) * z, z])
        spheres.append((c, rng.uniform(0.12, 0.3), colour()))
    if scale != 1.0:
        planes = [(n, d * scale, col) for n, d, col in planes]
        spheres = [(c * scale, r * scale, col) for c, r, col in spheres]
    return Scene(planes, spheres, colour())


def random_rigid(seed: int, max_angle_deg: float, max_trans: float) -> np.ndarray:
    """Seeded rigid transform: axis uniform on S^2, angle U[0, max], translation uniform in a ball."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = math.radians(rng.uniform(0.0, max_angle_deg))
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    t = d * max_trans * rng.uniform(0.0, 1.0) ** (1.0 / 3.0)
    G = np.eye(4)
    G[:3, :3] = R
    G[:3, 3] = t
    return G


def _raycast(scene: Scene, cam: CameraModel, G: np.ndarray):
    """Depth (metres along camera z) and object id per pixel for a camera with pose G."""
    H, W = cam.height, cam.width
    xs = (np.arange(W, dtype=np.float64) - 0.5 - cam.cx) / cam.fx
    ys = (np.arange(H, dtype=np.float64) - 0.5 - cam.cy) / cam.fy
    dx, dy = np.meshgrid(xs, ys)
    dirs = np.stack([dx, dy, np.ones_like(dx)], -1)             # camera frame, z = 1
    R, o = G[:3, :3], G[:3, 3]
    wd = dirs @ R.T                                              # world directions
    best = np.full((H, W), np.inf)
    obj = np.full((H, W), -1, np.int32)
    k = 0
    for n, d, _ in scene.planes:
        den = wd @ n
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (d - o @ n) / den
        t = np.where((t > 1e-6) & np.isfinite(t), t, np.inf)
        m = t < best
        best[m], obj[m] = t[m], k
        k += 1
    for c, r, _ in scene.spheres:
        oc = o - c
        b = wd @ oc
        a = np.einsum("ijk,ijk->ij", wd, wd)
        cc = oc @ oc - r * r
        disc = b * b - a * cc
        with np.errstate(invalid="ignore"):
            t = (-b - np.sqrt(disc)) / a
        t = np.where((disc >= 0) & (t > 1e-6), t, np.inf)
        m = t < best
        best[m], obj[m] = t[m], k
        k += 1
    return best, obj    # best is the ray parameter, equal to camera-frame z since dir_z = 1


def render(scene: Scene, cam: CameraModel, G: np.ndarray | None, seed: int,
           z_range=(0.5, 3.0), noise_m=0.002, hole_frac=0.05, colour_noise=10):
    """Render (depth uint16 [H,W] in millimetres, bgr uint8 [H,W,3])."""
    G = np.eye(4) if G is None else G
    rng = np.random.default_rng(np.random.PCG64(seed))
    z, obj = _raycast(scene, cam, G)
    valid = np.isfinite(z) & (z >= z_range[0]) & (z <= z_range[1])
    zn = z + rng.normal(0.0, noise_m, size=z.shape)
    mm = np.where(valid, np.rint(zn / cam.depth_scale), 0.0)
    mm = np.clip(mm, 0, 65535).astype(np.uint16)
    holes = rng.random(size=z.shape) < hole_frac
    mm[holes] = 0
    base = np.array([p[2] for p in scene.planes] + [s[2] for s in scene.spheres] +
                    [scene.background], np.int32)
    col = base[np.where(obj >= 0, obj, len(base) - 1)]
    col = col + rng.integers(-colour_noise, colour_noise + 1, size=col.shape)
    bgr = np.clip(col, 0, 255).astype(np.uint8)
    return mm, bgr


def frame_pair(seed: int, width: int, height: int, max_angle_deg=10.0, max_trans=0.1):
    """Target frame (pose I), source frame (pose G), G and the camera model."""
    cam = camera_for(width, height)
    scene = make_scene(seed)
    G = random_rigid(seed + 7919, max_angle_deg, max_trans)
    tgt = render(scene, cam, None, seed + 1)
    src = render(scene, cam, G, seed + 2)
    return tgt, src, G, cam


def sample_surface_pair(seed: int, n_points: int, max_angle_deg=15.0, max_trans=0.2,
                        noise_m=0.002, colour_noise=10, scale=1.0):
    """Two independent samplings (n_points each) of one scene's visible surfaces.

    Returns (src_xyz, src_rgb, tgt_xyz, tgt_rgb, G) with G mapping source to target.  Used
    for batched-registration workloads where each pair has its own point count.  Points are
    pixels of a virtual camera whose resolution gives about n_points valid pixels; `scale`
    scales the scene (make_scene).
    """
    cam_w = int(max(64, math.sqrt(n_points * 4.0 / 3.0 / 0.9)))
    cam_h = int(cam_w * 3 // 4)
    cam = CameraModel(cam_w, cam_h, 0.9609375 * cam_w, 0.9609375 * cam_w, cam_w / 2.0, cam_h / 2.0)
    scene = make_scene(seed, scale)
    G = random_rigid(seed + 104729, max_angle_deg, max_trans)
    out = []
    for k, pose in enumerate((G, np.eye(4))):
        rng = np.random.default_rng(np.random.PCG64(seed * 31 + k))
        z, obj = _raycast(scene, cam, pose)
        valid = np.isfinite(z) & (z >= 0.5 * scale) & (z <= 3.0 * scale)
        ys, xs = np.nonzero(valid)
        sel = rng.permutation(len(xs))[:n_points]
        ys, xs = ys[sel], xs[sel]
        zz = z[ys, xs] + rng.normal(0.0, noise_m, size=len(xs))
        X = (xs - 0.5 - cam.cx) * zz / cam.fx
        Y = (ys - 0.5 - cam.cy) * zz / cam.fy
        xyz = np.stack([X, Y, zz], -1).astype(np.float32)
        base = np.array([p[2] for p in scene.planes] + [s[2] for s in scene.spheres], np.int32)
        col = base[obj[ys, xs]] + rng.integers(-colour_noise, colour_noise + 1, size=(len(xs), 3))
        bgr = np.clip(col, 0, 255).astype(np.float32)
        rgb = (bgr[:, ::-1] * np.float32(1.0 / 255.0)).astype(np.float32)
        out.append((xyz, np.ascontiguousarray(rgb)))
    (sx, sr), (tx, tr) = out
    return sx, sr, tx, tr, G


def frame_sequence(seed: int, n_frames: int, width: int, height: int, max_angle_deg=10.0, max_trans=0.1):
    """n_frames distinct frames of one seeded scene, each from its own camera pose.  The poses
    scatter around the identity by at most half the motion bounds each, so any two frames (in
    particular consecutive ones, and the last and the first) differ by at most max_angle_deg /
    max_trans.  Returns ([(depth, bgr)], [pose], camera model); pose k maps frame k's camera
    coordinates into the first frame's world."""
    cam = camera_for(width, height)
    scene = make_scene(seed)
    frames, poses = [], []
    for k in range(n_frames):
        G = np.eye(4) if k == 0 else random_rigid(seed + 104729 * (k + 1), 0.5 * max_angle_deg, 0.5 * max_trans)
        frames.append(render(scene, cam, G, seed + 1 + k))
        poses.append(G)
    return frames, poses, cam
