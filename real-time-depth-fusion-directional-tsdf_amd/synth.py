"""Seeded synthetic depth frames of the fusion workloads (lib/libvsg_synth.so).

Workloads (scene sizes, sensor and volume parameters):

  0  "sphere-arc 320x240": 30 frames, f=262.5, sphere r=0.3 m + ground plane,
     camera on a 20 deg arc, noise N(0, 1 mm); voxel 4 mm, truncation 0.02 m,
     200000 buckets, FP32 store.
  1  "room 640x480": 300 frames, f=525, 4x4x3 m room with 20 primitives,
     random-walk camera, noise 0.0012+0.0019(z-0.4)^2; voxel 4 mm, truncation
     0.02 m, 2^20 buckets, FP32 store.
  4  "room 1024x768": 1000 frames, f=840, 8x8x3 m room with 60 primitives;
     voxel 2 mm, truncation 0.01 m, 2^22 buckets, FP16 store.

Depths outside [0.2, 4] m are invalid (point (0,0,0)).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libvsg_synth.so")

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with build_lib.py")
        L = C.CDLL(LIB_PATH)
        L.vsgs_create.restype = C.c_void_p
        L.vsgs_create.argtypes = [C.c_int, C.c_uint64, C.c_int]
        L.vsgs_destroy.argtypes = [C.c_void_p]
        L.vsgs_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.vsgs_pose.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.vsgs_render.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]
        L.vsgs_depth_range.argtypes = [C.c_void_p, C.c_long, C.POINTER(C.c_float)]
        _lib = L
    return _lib


@dataclass(frozen=True)
class Workload:
    name: str
    config: int
    voxel: float
    trunc: float
    buckets: int
    blocks: int
    store: str


WORKLOADS = {
    0: Workload("sphere-arc 320x240", 0, 0.004, 0.02, 200000, 1 << 17, "fp32"),
    1: Workload("room 640x480", 1, 0.004, 0.02, 1 << 20, 1 << 19, "fp32"),
    4: Workload("room 1024x768", 4, 0.002, 0.01, 1 << 22, 1 << 22, "fp16"),
}


class SynthScene:
    """Scene + trajectory of one workload; renders frames on demand."""

    def __init__(self, config: int, seed: int = 0, nframes: int = 0):
        self._h = C.c_void_p(lib().vsgs_create(int(config), int(seed), int(nframes)))
        whn = (C.c_int * 3)()
        intr = (C.c_double * 4)()
        lib().vsgs_info(self._h, whn, intr)
        self.config = config
        self.W, self.H, self.nframes = whn[0], whn[1], whn[2]
        self.fx, self.fy, self.cx, self.cy = (float(v) for v in intr)

    def __del__(self):
        try:
            if self._h and self._h.value:
                lib().vsgs_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    @property
    def intrinsics(self):
        return self.fx, self.fy, self.cx, self.cy

    def pose(self, frame: int) -> np.ndarray:
        """Trajectory pose of a frame: float64 4x4 row-major camera->world."""
        T = (C.c_double * 16)()
        lib().vsgs_pose(self._h, int(frame), T)
        return np.array(T[:], dtype=np.float64).reshape(4, 4)

    def render(self, frame: int, pose: np.ndarray | None = None, noise: bool = True) -> dict:
        """points float32 [H,W,3], colors uint8 [H,W,3], normals float32 [H,W,3], depth float32 [H,W],
        pose float32 [4,4]."""
        H, W = self.H, self.W
        pts = np.zeros((H, W, 3), np.float32)
        rgb = np.zeros((H, W, 3), np.uint8)
        nrm = np.zeros((H, W, 3), np.float32)
        depth = np.zeros((H, W), np.float32)
        T = self.pose(frame) if pose is None else np.asarray(pose, np.float64).reshape(4, 4)
        Tc = np.ascontiguousarray(T.reshape(16))
        lib().vsgs_render(self._h, int(frame), Tc.ctypes.data, 1 if noise else 0, pts.ctypes.data, rgb.ctypes.data,
                          nrm.ctypes.data, depth.ctypes.data)
        return {"points": pts, "colors": rgb, "normals": nrm, "depth": depth, "pose": T.astype(np.float32)}


def depth_range(depth: np.ndarray) -> tuple[float, float]:
    """The driver's per-frame ray range (VoxelSG/src/voxelSG.cpp:594-608)."""
    d = np.ascontiguousarray(depth, dtype=np.float32).reshape(-1)
    out = (C.c_float * 2)()
    lib().vsgs_depth_range(d.ctypes.data, d.size, out)
    return float(out[0]), float(out[1])
