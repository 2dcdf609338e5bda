"""C ABI checks that need no GPU: the library loads, exports every declared entry point, the
ctypes struct layouts match the C header, and argument validation fails cleanly."""
import ctypes as C
import os
import subprocess

import pytest

import cuda_3d_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    declared = L.declared_symbols()
    assert len(declared) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(L.lib, s)


def test_status_strings_and_version():
    assert L.lib.gca_abi_version() == 1
    for s in range(6):
        assert L.lib.gca_status_string(s).decode()
    assert L.lib.gca_status_string(99) == b"unknown status"


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "gca/gca.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(gca_camera),"
        " sizeof(gca_bilateral), sizeof(gca_icp_params), sizeof(gca_icp_report),"
        " sizeof(gca_level_params), offsetof(gca_camera, depth_to_color),"
        " offsetof(gca_icp_params, conv_rotation), offsetof(gca_level_params, icp));return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(L.Camera), C.sizeof(L.Bilateral), C.sizeof(L.IcpParams), C.sizeof(L.IcpReport),
            C.sizeof(L.LevelParams), L.Camera.depth_to_color.offset, L.IcpParams.conv_rotation.offset,
            L.LevelParams.icp.offset]
    assert got == want


def test_null_arguments_rejected_without_device():
    assert L.lib.gca_ctx_create(0, None) == L.GCA_ERR_INVALID_VALUE
    assert L.lib.gca_cloud_create(None, 10, None) == L.GCA_ERR_INVALID_VALUE
    assert L.lib.gca_voxel_downsample(None, None, 0.01, None) == L.GCA_ERR_INVALID_VALUE
    assert L.lib.gca_solve_gauss_newton(None, None, None, None) == L.GCA_ERR_INVALID_VALUE


def test_host_solver_matches_closed_form():
    import numpy as np

    A = np.diag([2.0, 3.0, 4.0, 5.0, 6.0, 7.0])
    up = np.ascontiguousarray([A[a, b] for a in range(6) for b in range(a, 6)])
    b = np.array([0.02, -0.03, 0.04, 0.1, -0.2, 0.3])
    xi = np.zeros(6)
    T = np.zeros(16)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    assert L.lib.gca_solve_gauss_newton(ptr(up), ptr(b), ptr(xi), ptr(T)) == 0
    np.testing.assert_allclose(xi, -b / np.diag(A), rtol=1e-15)
    zero = np.zeros(21)
    assert L.lib.gca_solve_gauss_newton(ptr(zero), ptr(b), ptr(xi), ptr(T)) == L.GCA_ERR_NOT_SPD
    np.testing.assert_array_equal(T.reshape(4, 4), np.eye(4))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="only meaningful without a device")
def test_context_creation_fails_cleanly_without_device():
    h = C.c_void_p()
    assert L.lib.gca_ctx_create(0, C.byref(h)) == L.GCA_ERR_DEVICE


def test_cpp_facade_builds_and_links():
    """The C++ facade (include/gca/gca.hpp) compiles warning-free against gca.h and links
    with libgca.so; the example prints its usage without touching a device."""
    cpp = os.path.join(ROOT, "tests", "cpp")
    r = subprocess.run(["make", "-s", "-B", "-C", cpp], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr
    r = subprocess.run([os.path.join(cpp, "facade_icp")], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
