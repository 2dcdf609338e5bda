"""gca_build_gauss_newton_noalloc (the allocation-free entry point) against gca_build_gauss_newton:
the same kernels over the same work partition, so the same bits, with the sums left in device
memory.  Two calls in a row on one device, the second on a smaller input, so the second must not
see the first's partials (its ticket and partial slots are the shared static scratch)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAM = 0.968


def _case(g, ctx, n, seed):
    from cuda_3d_amd import scene

    sx, sr, tx, tr, G = scene.sample_surface_pair(seed, n)
    src = g.PointCloud.from_numpy(sx, sr, ctx=ctx)
    tgt = g.PointCloud.from_numpy(tx, tr, ctx=ctx)
    tgt.estimate_normals(0.05, 30)
    tgt.compute_color_gradient(0.04, 0)
    T = G.copy()
    T[:3, 3] += 0.002
    moved = src.copy()
    moved.transform(T)
    nn = g.PointCloud.nn_search(moved, tgt, 0.05, as_tensor=True)
    return src, tgt, nn, T


def test_noalloc_gauss_newton_matches_allocating_entry_point(ctx):
    import torch

    import cuda_3d_amd as g
    from cuda_3d_amd import _lib as L

    cases = [_case(g, ctx, 60000, 41), _case(g, ctx, 3000, 42)]  # the second call on a smaller input
    for mode in (0, 1, 2):
        outs = []
        for src, tgt, nn, T in cases:
            out = torch.full((29,), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            Tc = np.ascontiguousarray(T.reshape(16), np.float64)
            st = L.lib.gca_build_gauss_newton_noalloc(ctx.handle, src.handle, tgt.handle, C.c_void_p(nn.data_ptr()),
                                                      LAM, Tc.ctypes.data_as(C.c_void_p), mode,
                                                      C.c_void_p(out.data_ptr()))
            assert st == L.GCA_OK
            outs.append(out)
        ctx.synchronize()
        for (src, tgt, nn, T), out in zip(cases, outs):
            want = np.zeros(29)
            Tc = np.ascontiguousarray(T.reshape(16), np.float64)
            assert L.lib.gca_build_gauss_newton(ctx.handle, src.handle, tgt.handle, C.c_void_p(nn.data_ptr()), LAM,
                                                Tc.ctypes.data_as(C.c_void_p), mode,
                                                want.ctypes.data_as(C.c_void_p)) == L.GCA_OK
            got = out.cpu().numpy()
            print(f"[no-alloc] mode {mode}, {src.points_number()} points: {int(got[28])} correspondences", flush=True)
            np.testing.assert_array_equal(got, want)
            assert got[28] > 0.5 * src.points_number()
    # argument errors are refused before anything is enqueued
    src, tgt, nn, T = cases[1]
    Tc = np.ascontiguousarray(T.reshape(16), np.float64)
    assert L.lib.gca_build_gauss_newton_noalloc(ctx.handle, src.handle, tgt.handle, C.c_void_p(nn.data_ptr()), 1.5,
                                                Tc.ctypes.data_as(C.c_void_p), 0, C.c_void_p(1)) == L.GCA_ERR_INVALID_VALUE
    assert L.lib.gca_build_gauss_newton_noalloc(ctx.handle, src.handle, tgt.handle, C.c_void_p(nn.data_ptr()), LAM,
                                                Tc.ctypes.data_as(C.c_void_p), 0, None) == L.GCA_ERR_INVALID_VALUE
    bare = g.PointCloud.from_numpy(np.zeros((10, 3)), np.zeros((10, 3)), ctx=ctx)  # no normals
    assert L.lib.gca_build_gauss_newton_noalloc(ctx.handle, src.handle, bare.handle, C.c_void_p(nn.data_ptr()), LAM,
                                                Tc.ctypes.data_as(C.c_void_p), 0, C.c_void_p(1)) == L.GCA_ERR_NOT_READY
