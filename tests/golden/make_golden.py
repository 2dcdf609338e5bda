"""Writes tests/golden/golden_small.npz: small seeded inputs with expected outputs computed
independently of the oracle, by brute force in numpy/scipy and from closed forms.

The reference repository has no tests, golden vectors or known-answer fixtures for this path,
so these vectors are the external pins of the FP64 oracle:
  bilateral   direct FP64 double loop of the filter definition (exp(-x^2/sigma^2) weights,
              zero taps skipped) on a 16x12 frame with holes
  backproj    pinhole closed form in float32 with the (x-0.5, y-0.5) pixel centre
  voxel       np.floor keys, np.unique(axis=0) order, np.add.at FP64 means
  neighbours  O(N*M) FP64 distances: nearest (d2, index), radius lists, K-nearest
  normals     np.linalg.eigh of the centred covariance of brute-force neighbour sets
  gradient    np.linalg.solve of the regularised 3x3 normal equations
  jacobian    central finite differences of the colored-ICP residuals over the Euler pose
  residuals   the per-point residual kinds of cuda_color_icp_build_least_square.cu:13-132:
              weighted pair (sqrt(l) r_g, sqrt(1-l) r_c), |.|+|.| of the pair, and the
              unweighted objective l r_g^2 + (1-l) r_c^2; zero where there is no correspondence
  solve6      np.linalg.solve + scipy Rotation.from_euler('xyz') for Rz Ry Rx
Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np
from scipy.spatial.transform import Rotation

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")


def bilateral_case(rng):
    H, W = 12, 16
    depth = rng.integers(900, 1300, size=(H, W)).astype(np.uint16)
    depth[rng.random((H, W)) < 0.1] = 0
    depth[5, 7] = 0
    R, ss, sd = 2, 2.0, 30.0
    filt = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            d0 = int(depth[y, x])
            if d0 == 0:
                continue
            sw = swd = 0.0
            for dy in range(-R, R + 1):
                for dx in range(-R, R + 1):
                    yy, xx = y + dy, x + dx
                    if not (0 <= yy < H and 0 <= xx < W):
                        continue
                    dn = int(depth[yy, xx])
                    if dn == 0:
                        continue
                    w = np.exp(-(dx * dx + dy * dy) / (ss * ss)) * np.exp(-((d0 - dn) ** 2) / (sd * sd))
                    sw += w
                    swd += w * dn
            filt[y, x] = swd / sw
    bgr = rng.integers(0, 256, size=(H, W, 3)).astype(np.uint8)
    fx = fy = np.float32(20.0)
    cx, cy = np.float32(8.0), np.float32(6.0)
    z = (filt.astype(np.float32) * np.float32(0.001)).astype(np.float32)
    valid = (depth > 0) & (z >= np.float32(0.2)) & (z <= np.float32(3.5))
    ys, xs = np.nonzero(valid)
    zz = z[ys, xs]
    X = (((xs.astype(np.float32) - np.float32(0.5)) - cx) * zz) / fx
    Y = (((ys.astype(np.float32) - np.float32(0.5)) - cy) * zz) / fy
    # colour pixel of the reference's float32 re-projection, int(u + 0.5) truncation
    # (cuda_point_cloud_factory.cu:182-187,242-243): may land one pixel left / up of the
    # depth pixel, because u + 0.5 is an integer only in exact arithmetic
    tx = ((X * fx) / zz + cx + np.float32(0.5)).astype(np.int32)
    ty = ((Y * fy) / zz + cy + np.float32(0.5)).astype(np.int32)
    col = bgr[ty, tx].astype(np.float32) * np.float32(1.0 / 255.0)
    return dict(bil_depth=depth, bil_bgr=bgr, bil_filtered=filt, bil_valid=valid,
                bil_xyz=np.stack([X, Y, zz], 1).astype(np.float32), bil_rgb=col[:, ::-1].copy())


def voxel_case(rng):
    n = 300
    xyz = rng.uniform(-0.3, 0.3, size=(n, 3)).astype(np.float32)
    rgb = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    v = np.float32(0.05)
    mn = xyz.min(0)
    keys = np.floor((xyz - mn) / v).astype(np.int64)           # float32 sub/div as the reference
    uk, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    s = np.zeros((len(uk), 6))
    np.add.at(s, inv, np.concatenate([xyz, rgb], 1).astype(np.float64))
    cnt = np.bincount(inv).astype(np.float64)
    mean = s / cnt[:, None]
    return dict(vox_xyz=xyz, vox_rgb=rgb, vox_size=v, vox_keys=uk, vox_mean=mean, vox_count=cnt)


def neighbour_case(rng):
    ref = rng.uniform(0, 0.6, size=(400, 3)).astype(np.float32)
    qry = rng.uniform(-0.05, 0.65, size=(150, 3)).astype(np.float32)
    qry[:10] = ref[:10]                      # exact hits
    qry[10:12] = ref[20] + np.float32(0.0)   # duplicate query
    ref[30] = ref[31]                        # duplicate reference points: tie by index
    r = np.float32(0.08)
    d2 = ((qry[:, None, :].astype(np.float64) - ref[None, :, :].astype(np.float64)) ** 2).sum(-1)
    r2 = float(r) * float(r)
    nn = np.full(len(qry), -1, np.int32)
    lists, lists_k = [], []
    for i in range(len(qry)):
        inr = np.nonzero(d2[i] < r2)[0]
        order = inr[np.lexsort((inr, d2[i, inr]))]
        if len(order):
            nn[i] = order[0]
        lists.append(order)
        lists_k.append(order[:3])
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int32)
    offk = np.concatenate([[0], np.cumsum([len(l) for l in lists_k])]).astype(np.int32)
    return dict(nb_ref=ref, nb_qry=qry, nb_radius=r, nb_nn=nn, nb_off=off,
                nb_idx=np.concatenate(lists).astype(np.int32), nb_offk=offk,
                nb_idxk=np.concatenate(lists_k).astype(np.int32))


def surface_points(rng, n):
    a = rng.uniform(-0.3, 0.3, size=(n // 2, 2))
    plane = np.stack([a[:, 0], a[:, 1], 1.0 + 0.2 * a[:, 0] - 0.1 * a[:, 1]], 1)
    u = rng.normal(size=(n - n // 2, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u[:, 2] = -np.abs(u[:, 2])
    sphere = np.array([0.1, 0.0, 1.2]) + 0.15 * u
    p = np.concatenate([plane, sphere]) + rng.normal(0, 0.002, size=(n, 3))
    return p.astype(np.float32)


def brute_lists(p, r, K):
    d2 = ((p[:, None, :].astype(np.float64) - p[None, :, :].astype(np.float64)) ** 2).sum(-1)
    out = []
    for i in range(len(p)):
        inr = np.nonzero(d2[i] < float(r) * float(r))[0]
        order = inr[np.lexsort((inr, d2[i, inr]))]
        out.append(order[:K] if K > 0 else order)
    return out


def normals_case(rng):
    p = surface_points(rng, 250)
    inten = rng.uniform(0, 1, len(p)).astype(np.float32)
    r = np.float32(0.1)
    res = {}
    for K in (0, 15):
        nrm = np.zeros((len(p), 3))
        for i, nb in enumerate(brute_lists(p, r, K)):
            if len(nb) < 3:
                continue
            X = p[nb].astype(np.float64)
            C = np.cov(X.T, bias=True)
            w, V = np.linalg.eigh(C)
            v = V[:, 0]
            if v[2] > 0:
                v = -v
            nrm[i] = v
        res[f"nrm_K{K}"] = nrm
    # colour gradient on the K=15 normals, uncapped radius 0.08 neighbourhoods
    n = res["nrm_K15"].astype(np.float32)
    rg = np.float32(0.08)
    grad = np.zeros((len(p), 3))
    for i, nb in enumerate(brute_lists(p, rg, 0)):
        k = len(nb)
        if k < 5:
            continue
        nv = n[i].astype(np.float64)
        A = np.zeros((3, 3))
        b = np.zeros(3)
        for j in nb:
            if j == i:
                continue
            e = p[j].astype(np.float64) - p[i].astype(np.float64)
            v = e - (e @ nv) * nv
            A += np.outer(v, v)
            b += v * (float(inten[j]) - float(inten[i]))
        A += (k - 1) ** 2 * np.outer(nv, nv) + 1e-6 * np.eye(3)
        grad[i] = np.linalg.solve(A, b)
    return dict(nrm_pts=p, nrm_radius=r, nrm_K0=res["nrm_K0"], nrm_K15=res["nrm_K15"],
                grad_int=inten, grad_nrm=n, grad_radius=rg, grad=grad)


def euler(xi):
    return Rotation.from_euler("xyz", xi[:3]).as_matrix()  # extrinsic x, y, z == Rz Ry Rx


def residuals(xi, T, src, tgt, tn, tg, nn, lam):
    R = euler(xi)
    out = []
    for i, j in enumerate(nn):
        if j < 0:
            continue
        y = T[:3, :3] @ src[i, :3].astype(np.float64) + T[:3, 3]
        p = R @ y + xi[3:]
        q, n, d = tgt[j, :3].astype(np.float64), tn[j].astype(np.float64), tg[j].astype(np.float64)
        rg = (p - q) @ n
        rc = tgt[j, 3] + d @ (p - rg * n - q) - src[i, 3]
        out += [np.sqrt(lam) * rg, np.sqrt(1 - lam) * rc]
    return np.array(out)


def gn_case(rng):
    ns, nt = 40, 50
    tgt = np.concatenate([rng.uniform(-0.5, 0.5, (nt, 2)), rng.uniform(0.8, 1.2, (nt, 1)),
                          rng.uniform(0, 1, (nt, 1))], 1).astype(np.float32)
    src = np.concatenate([rng.uniform(-0.5, 0.5, (ns, 2)), rng.uniform(0.8, 1.2, (ns, 1)),
                          rng.uniform(0, 1, (ns, 1))], 1).astype(np.float32)
    tn = rng.normal(size=(nt, 3))
    tn = (tn / np.linalg.norm(tn, axis=1, keepdims=True)).astype(np.float32)
    tg = rng.normal(0, 2, size=(nt, 3)).astype(np.float32)
    nn = rng.integers(-1, nt, size=ns).astype(np.int32)
    T = np.eye(4)
    T[:3, :3] = euler(np.array([0.05, -0.03, 0.02]))
    T[:3, 3] = [0.02, -0.01, 0.03]
    lam = float(np.float32(0.968))  # the C ABI takes lambda as float
    r0 = residuals(np.zeros(6), T, src, tgt, tn, tg, nn, lam)
    h = 1e-6
    J = np.zeros((len(r0), 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        J[:, k] = (residuals(e, T, src, tgt, tn, tg, nn, lam) - residuals(-e, T, src, tgt, tn, tg, nn, lam)) / (2 * h)
    JTJ = J.T @ J
    JTr = J.T @ r0
    # per-point residual kinds, scattered back to every source point (0 without a match)
    pair = np.zeros((ns, 2))
    pair[nn >= 0] = r0.reshape(-1, 2)
    raw = pair / np.array([np.sqrt(lam), np.sqrt(1 - lam)])
    return dict(gn_src=src, gn_tgt=tgt, gn_tn=tn, gn_tg=tg, gn_nn=nn, gn_T=T, gn_lam=lam,
                gn_JTJ=JTJ, gn_JTr=JTr, gn_r2=float(r0 @ r0), gn_ncorr=int((nn >= 0).sum()),
                res_pair=pair, res_abs=np.abs(pair).sum(1),
                res_rmse=lam * raw[:, 0] ** 2 + (1 - lam) * raw[:, 1] ** 2)


def solve_case(rng):
    A = rng.normal(size=(6, 6))
    A = A @ A.T + 6 * np.eye(6)
    b = rng.normal(size=6)
    xi = np.linalg.solve(A, -b)
    T = np.eye(4)
    T[:3, :3] = euler(xi)
    T[:3, 3] = xi[3:]
    return dict(sv_A=A, sv_b=b, sv_xi=xi, sv_T=T)


def main():
    rng = np.random.default_rng(np.random.PCG64(20261202))
    d = {}
    for f in (bilateral_case, voxel_case, neighbour_case, normals_case, gn_case, solve_case):
        d.update(f(rng))
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, {k: getattr(v, "shape", ()) for k, v in d.items()})


if __name__ == "__main__":
    main()
