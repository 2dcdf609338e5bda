"""PLY ingest: the host side of apps/test_color_icp.cpp:298-343 (loadPointCloud, removeNaNPoints)
feeding point_cloud::create_from_pcl (point_cloud.cu:481-504).

Reads binary little-endian and ASCII PLY files whose ``vertex`` element holds x, y, z and,
optionally, nx, ny, nz and red, green, blue (uchar).  Elements before ``vertex`` are skipped;
elements after it (faces) are never read.  Colours become r * (1.0f / 255.0f) in FP32, exactly as
create_from_pcl forms them.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


@dataclass
class PlyVertices:
    xyz: np.ndarray                 # float32 [n, 3]
    rgb8: Optional[np.ndarray]      # uint8 [n, 3] or None
    normals: Optional[np.ndarray]   # float32 [n, 3] or None

    def rgb(self) -> np.ndarray:
        """create_from_pcl's colour: (float)c * (1.0f / 255.0f); black when the file has none."""
        if self.rgb8 is None:
            return np.zeros_like(self.xyz)
        return self.rgb8.astype(np.float32) * (np.float32(1.0) / np.float32(255.0))


def _header(buf: bytes):
    end = buf.find(b"end_header")
    if not buf.startswith(b"ply") or end < 0:
        raise ValueError("not a PLY file")
    nl = buf.find(b"\n", end)
    body = (nl + 1) if nl >= 0 else len(buf)
    fmt, elements = None, []
    for raw in buf[:end].decode("ascii", "replace").splitlines()[1:]:
        t = raw.split()
        if not t or t[0] in ("comment", "obj_info"):
            continue
        if t[0] == "format":
            fmt = t[1]
        elif t[0] == "element":
            elements.append([t[1], int(t[2]), []])
        elif t[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if t[1] == "list":
                elements[-1][2].append((t[4], None, (t[2], t[3])))
            else:
                if t[1] not in _TYPES:
                    raise ValueError(f"unknown PLY type {t[1]!r}")
                elements[-1][2].append((t[2], _TYPES[t[1]], None))
    if fmt not in ("binary_little_endian", "ascii"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, elements, body


def read_ply(path: str, remove_nan: bool = True) -> PlyVertices:
    """Vertices of a PLY file.  ``remove_nan`` is the reference's removeNaNPoints
    (test_color_icp.cpp:335-342): it drops points with a non-finite coordinate
    (removeNaNFromPointCloudBruteForce, :223-259) and points whose packed colour word is zero
    (removeNaNRGBFromPointCloud, :262-290).  PCL's PLY reader packs r << 16 | g << 8 | b with a zero
    alpha byte unless the file has an alpha property, so a pure-black vertex of a coloured file is
    dropped; a file without colour keeps PCL's default alpha of 255 and loses nothing here."""
    with open(path, "rb") as f:
        buf = f.read()
    fmt, elements, off = _header(buf)
    names = [e[0] for e in elements]
    if "vertex" not in names:
        raise ValueError("PLY file has no vertex element")
    vi = names.index("vertex")
    _, n, props = elements[vi]
    if any(p[2] is not None for p in props):
        raise ValueError("list property in the vertex element")
    if fmt == "ascii":
        lines = buf[off:].decode("ascii").splitlines()
        row0 = sum(e[1] for e in elements[:vi])
        rows = [ln.split() for ln in lines[row0:row0 + n]]
        if len(rows) != n or any(len(r) != len(props) for r in rows):
            raise ValueError("truncated ASCII PLY vertex block")
        cols = {p[0]: np.array([r[k] for r in rows], dtype=np.float64).astype(p[1])
                for k, p in enumerate(props)}
    else:
        for name, count, eprops in elements[:vi]:
            if any(p[2] is not None for p in eprops):
                raise ValueError(f"list property in element {name!r} before the vertices")
            off += count * sum(np.dtype(p[1]).itemsize for p in eprops)
        dt = np.dtype([(p[0], "<" + p[1]) for p in props])
        if off + n * dt.itemsize > len(buf):
            raise ValueError("truncated binary PLY vertex block")
        v = np.frombuffer(buf, dt, count=n, offset=off)
        cols = {p[0]: v[p[0]] for p in props}
    for c in ("x", "y", "z"):
        if c not in cols:
            raise ValueError(f"PLY vertices have no {c!r}")
    xyz = np.stack([cols["x"], cols["y"], cols["z"]], 1).astype(np.float32)
    rgb8 = None
    if all(c in cols for c in ("red", "green", "blue")):
        rgb8 = np.stack([cols["red"], cols["green"], cols["blue"]], 1).astype(np.uint8)
    nrm = None
    if all(c in cols for c in ("nx", "ny", "nz")):
        nrm = np.stack([cols["nx"], cols["ny"], cols["nz"]], 1).astype(np.float32)
    if remove_nan:
        keep = np.isfinite(xyz).all(1)
        if rgb8 is not None:
            word_zero = (rgb8 == 0).all(1)
            alpha = cols.get("alpha")
            if alpha is not None:
                word_zero &= np.asarray(alpha).astype(np.uint8) == 0
            keep &= ~word_zero
        if not keep.all():
            xyz = xyz[keep]
            rgb8 = None if rgb8 is None else rgb8[keep]
            nrm = None if nrm is None else nrm[keep]
    return PlyVertices(np.ascontiguousarray(xyz), rgb8, nrm)
