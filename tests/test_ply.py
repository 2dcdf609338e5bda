"""PLY ingest (cuda_3d_amd.ply): the reference's own data/frag_115.ply, every fifth vertex, read back
against the figures a plain numpy structured read recorded; ASCII and binary files with elements
before the vertices; NaN removal; malformed files."""
import json
import os

import numpy as np
import pytest

from cuda_3d_amd.ply import read_ply

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "frag_115_every5.ply")


def test_fixture_reads_back():
    meta = json.load(open(os.path.join(HERE, "golden", "frag_115_every5.json")))
    v = read_ply(FIX)
    assert len(v.xyz) == meta["vertices"] == -(-meta["source_vertices"] // meta["step"])
    np.testing.assert_allclose(v.xyz.astype(np.float64).sum(0), meta["xyz_sum"], rtol=1e-12)
    assert v.rgb8.astype(np.int64).sum(0).tolist() == meta["rgb_sum"]
    for row, ref in ((0, meta["first"]), (-1, meta["last"])):
        np.testing.assert_array_equal(v.xyz[row], np.float32(ref[:3]))
        np.testing.assert_array_equal(v.normals[row], np.float32(ref[3:6]))
        assert v.rgb8[row].tolist() == ref[6:9]
    np.testing.assert_array_equal(v.xyz.min(0), np.float32(meta["bounds_min"]))
    np.testing.assert_array_equal(v.xyz.max(0), np.float32(meta["bounds_max"]))
    assert np.abs(np.linalg.norm(v.normals, axis=1) - 1).max() < 1e-6
    c = v.rgb()
    assert c.dtype == np.float32 and c.min() >= 0 and c.max() <= 1.0
    # create_from_pcl's c * (1/255) sits within half an FP32 ulp of c / 255
    assert np.abs(c.astype(np.float64) * 255 - v.rgb8).max() < 1e-4


def test_ascii_with_leading_element_and_nan(tmp_path):
    p = tmp_path / "a.ply"
    p.write_text("ply\nformat ascii 1.0\ncomment test\nelement camera 1\nproperty float view_px\n"
                 "property float view_py\nelement vertex 3\nproperty double x\nproperty double y\n"
                 "property double z\nproperty uchar red\nproperty uchar green\nproperty uchar blue\n"
                 "element face 1\nproperty list uchar int vertex_indices\nend_header\n"
                 "1.5 2.5\n0.25 -1 2 10 20 30\nnan 0 1 1 1 1\n3 4 5 255 0 7\n3 0 1 2\n")
    v = read_ply(str(p))
    np.testing.assert_array_equal(v.xyz, np.float32([[0.25, -1, 2], [3, 4, 5]]))
    assert v.rgb8.tolist() == [[10, 20, 30], [255, 0, 7]] and v.normals is None
    assert len(read_ply(str(p), remove_nan=False).xyz) == 3


def test_black_vertices_are_dropped_like_the_reference(tmp_path):
    """removeNaNRGBFromPointCloud (test_color_icp.cpp:262-290) drops points whose packed colour
    word is zero: a pure-black vertex without alpha, or with alpha 0; black with alpha 255 and any
    vertex of a file without colour stay."""
    p = tmp_path / "black.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
                 "property float z\nproperty uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n"
                 "0 0 1 0 0 0\n1 0 1 0 0 1\n2 0 1 1 0 0\n3 0 1 0 0 0\n")
    v = read_ply(str(p))
    np.testing.assert_array_equal(v.xyz[:, 0], np.float32([1, 2]))
    assert v.rgb8.tolist() == [[0, 0, 1], [1, 0, 0]]
    assert len(read_ply(str(p), remove_nan=False).xyz) == 4
    q = tmp_path / "alpha.ply"
    q.write_text("ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\n"
                 "property float z\nproperty uchar red\nproperty uchar green\nproperty uchar blue\n"
                 "property uchar alpha\nend_header\n0 0 1 0 0 0 255\n1 0 1 0 0 0 0\n")
    np.testing.assert_array_equal(read_ply(str(q)).xyz[:, 0], np.float32([0]))
    r = tmp_path / "plain.ply"
    r.write_text("ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\n"
                 "property float z\nend_header\n0 0 0\n1 1 1\n")
    assert len(read_ply(str(r)).xyz) == 2


def test_binary_with_leading_element_and_no_colour(tmp_path):
    p = tmp_path / "b.ply"
    xyz = np.float32([[1, 2, 3], [-4, 5.5, 6]])
    head = ("ply\nformat binary_little_endian 1.0\nelement extra 2\nproperty int a\nproperty short b\n"
            "element vertex 2\nproperty float x\nproperty float y\nproperty float z\nend_header\n")
    p.write_bytes(head.encode() + np.zeros(2, dtype=[("a", "<i4"), ("b", "<i2")]).tobytes() + xyz.tobytes())
    v = read_ply(str(p))
    np.testing.assert_array_equal(v.xyz, xyz)
    assert v.rgb8 is None and not v.rgb().any()


@pytest.mark.parametrize("text", [
    b"plx\nend_header\n",
    b"ply\nformat binary_big_endian 1.0\nelement vertex 1\nproperty float x\nend_header\n",
    b"ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
    b"property float z\nend_header\n\x00\x00",
    b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\nproperty float y\nend_header\n"
    b"\x00" * 8,
])
def test_malformed_files_raise(tmp_path, text):
    p = tmp_path / "bad.ply"
    p.write_bytes(text)
    with pytest.raises(ValueError):
        read_ply(str(p))
