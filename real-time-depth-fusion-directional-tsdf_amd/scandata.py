"""On-disk inputs of the reference's driver (VoxelSG/src/scanData.cpp), for hosts that feed
the depth-map generator from files.

* point cloud ``%06d.bin`` (ScanDataLoader::loadPointCloudFormat, scanData.cpp:778-834):
  a ``size_t`` point count, then per point 3 float32 position, 3 float32 normal and 3
  uint8 colour, packed (27 bytes), little-endian;
* pose text ``%06d.txt`` (ScanDataLoader::loadMatrixParams, scanData.cpp:223-459): sections
  headed by lines containing ``Transform_0``, ``Transform_45`` or ``CameraRT``; a data
  line (text after a ``:`` if there is one) with 4 numbers fills the next row of the
  current section; camera->world = Transform * CameraRT^-1 when CameraRT's rotation part
  is not the identity, else Transform.
"""
from __future__ import annotations

import re

import numpy as np

_REC = np.dtype([("p", "<f4", 3), ("n", "<f4", 3), ("c", "u1", 3)])


def load_point_cloud_bin(path: str):
    """(points float32 [n, 3], normals float32 [n, 3], colours uint8 [n, 3])."""
    with open(path, "rb") as fh:
        head = fh.read(8)
        if len(head) < 8:
            raise ValueError(f"{path}: no point count")
        n = int(np.frombuffer(head, "<u8")[0])
        rec = np.frombuffer(fh.read(n * _REC.itemsize), _REC)
    if len(rec) != n:
        raise ValueError(f"{path}: {n} points announced, {len(rec)} stored")
    return rec["p"].copy(), rec["n"].copy(), rec["c"].copy()


def save_point_cloud_bin(path: str, points, normals, colors) -> None:
    """Writes the format load_point_cloud_bin reads (the reference's reader order)."""
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    rec = np.zeros(len(pts), _REC)
    rec["p"] = pts
    rec["n"] = np.asarray(normals, np.float32).reshape(-1, 3)
    rec["c"] = np.asarray(colors, np.uint8).reshape(-1, 3)
    with open(path, "wb") as fh:
        fh.write(np.array([len(pts)], "<u8").tobytes())
        fh.write(rec.tobytes())


def _stof(tok: str) -> float:
    """std::stof: the longest leading float of the token (leading blanks skipped); raises
    when there is none.  Values go through float32."""
    m = re.match(r"\s*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", tok,
                 re.IGNORECASE)
    if not m:
        raise ValueError(tok)
    return float(np.float32(float(m.group(1))))


def parse_matrix_params(text: str) -> dict:
    """The matrices of one pose file: transform_0, transform_45, localToCamera (float32 4x4,
    identity where absent) and cameraToWorld0 / cameraToWorld45."""
    mats = {k: np.eye(4, dtype=np.float32) for k in ("transform_0", "transform_45", "localToCamera")}
    cur, row = None, 0
    for line in text.splitlines():
        if "Transform_0" in line:
            cur, row = "transform_0", 0
            continue
        if "Transform_45" in line:
            cur, row = "transform_45", 0
            continue
        if "CameraRT" in line:
            cur, row = "localToCamera", 0
            continue
        if cur is None or not line or not any(ch in "0123456789.-" for ch in line):
            continue
        data = line.split(":", 1)[1] if ":" in line else line
        col = 0
        for tok in data.split():
            if col >= 4:
                break
            try:
                val = _stof(tok)
            except ValueError:
                break
            if row < 4:
                mats[cur][row, col] = val
            col += 1
        if col == 4:
            row += 1
    L = mats["localToCamera"]
    loaded = any(abs(float(L[i, j]) - (1.0 if i == j else 0.0)) > 1e-6 for i in range(3) for j in range(3))
    if loaded:
        inv = np.linalg.pinv(L.astype(np.float64)).astype(np.float32)
        c0 = (mats["transform_0"].astype(np.float64) @ inv.astype(np.float64)).astype(np.float32)
        c45 = (mats["transform_45"].astype(np.float64) @ inv.astype(np.float64)).astype(np.float32)
    else:
        c0, c45 = mats["transform_0"].copy(), mats["transform_45"].copy()
    return {**mats, "cameraToWorld0": c0, "cameraToWorld45": c45}


def load_matrix_params(path: str) -> dict:
    with open(path) as fh:
        return parse_matrix_params(fh.read())
