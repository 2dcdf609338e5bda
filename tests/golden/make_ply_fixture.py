"""Writes tests/golden/frag_115_every5.ply and frag_115_every5.json from the reference's
data/frag_115.ply (64 365 vertices: x y z nx ny nz float, red green blue uchar, then 127 229
faces).  Every fifth vertex is kept, in order, in a binary little-endian vertex block followed by
an empty face element; the JSON holds what the file must read back as (count, sums, the first and
last vertices, colour histogram checksum), taken with a plain numpy structured read independent of
cuda_3d_amd.ply.

Run from the repository root with the reference checked out at /root/reference.
"""
import json
import os

import numpy as np

SRC = "/root/reference/data/frag_115.ply"
HERE = os.path.dirname(os.path.abspath(__file__))
STEP = 5

VDT = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4"),
                ("red", "u1"), ("green", "u1"), ("blue", "u1")])


def main():
    buf = open(SRC, "rb").read()
    off = buf.index(b"end_header\n") + len(b"end_header\n")
    n = int(buf[:off].split(b"element vertex ")[1].split(b"\n")[0])
    v = np.frombuffer(buf, VDT, count=n, offset=off)[::STEP].copy()
    head = ("ply\nformat binary_little_endian 1.0\ncomment every 5th vertex of frag_115.ply\n"
            f"element vertex {len(v)}\n"
            "property float x\nproperty float y\nproperty float z\n"
            "property float nx\nproperty float ny\nproperty float nz\n"
            "property uchar red\nproperty uchar green\nproperty uchar blue\n"
            "element face 0\nproperty list uchar int vertex_indices\nend_header\n")
    with open(os.path.join(HERE, "frag_115_every5.ply"), "wb") as f:
        f.write(head.encode("ascii"))
        f.write(v.tobytes())
    xyz = np.stack([v["x"], v["y"], v["z"]], 1).astype(np.float64)
    rgb = np.stack([v["red"], v["green"], v["blue"]], 1).astype(np.int64)
    meta = {
        "source": "data/frag_115.ply", "step": STEP, "source_vertices": n, "vertices": len(v),
        "xyz_sum": xyz.sum(0).tolist(), "rgb_sum": rgb.sum(0).tolist(),
        "first": [float(c) for c in v[0].tolist()], "last": [float(c) for c in v[-1].tolist()],
        "bounds_min": xyz.min(0).tolist(), "bounds_max": xyz.max(0).tolist(),
    }
    with open(os.path.join(HERE, "frag_115_every5.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {len(v)} vertices")


if __name__ == "__main__":
    main()
