"""ctypes binding of the VoxelSG oracle (oracle/libvsg_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker and the CPU baseline; the product path never
imports it.  The oracle restates the reference's kernels on the CPU
(oracle/vsg_oracle.hpp); float arithmetic for parity, double for precision.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "libvsg_oracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run make in {ORACLE_DIR}")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.vsgo_create.restype = vp
        L.vsgo_create.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_int, C.c_int]
        L.vsgo_destroy.argtypes = [vp]
        L.vsgo_reset.argtypes = [vp]
        L.vsgo_integrate.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_double, vp, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int, vp]
        L.vsgo_set_directional.argtypes = [vp, C.c_int]
        L.vsgo_stage_times.argtypes = [vp, vp]
        L.vsgo_set_threads.restype = C.c_int
        L.vsgo_set_threads.argtypes = [C.c_int]
        L.vsgo_export_directional.argtypes = [vp, C.c_uint64, vp, vp, vp]
        L.vsgo_alloc.argtypes = [vp, vp, C.c_int, C.c_int, C.c_double, vp]
        L.vsgo_render.restype = C.c_int
        L.vsgo_render.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, vp, vp, vp, vp, vp, vp, C.c_long]
        L.vsgo_num_allocated.restype = C.c_uint64
        L.vsgo_num_allocated.argtypes = [vp]
        L.vsgo_heap_counter.restype = C.c_uint32
        L.vsgo_heap_counter.argtypes = [vp]
        L.vsgo_export_blocks.restype = C.c_uint64
        L.vsgo_export_blocks.argtypes = [vp, vp, vp, vp, C.c_uint64]
        L.vsgo_export_voxels.argtypes = [vp, C.c_uint64, vp, vp, vp, vp]
        L.vsgo_extract_surface_points.restype = C.c_uint64
        L.vsgo_extract_surface_points.argtypes = [vp, C.c_double, C.c_double, C.c_uint32, C.c_uint64, vp, vp, vp]
        L.vsgo_marching_cubes.restype = C.c_uint64
        L.vsgo_marching_cubes.argtypes = [vp, C.c_double, C.c_double, C.c_int32, C.c_uint32, vp, C.c_uint64, vp, vp]
        L.vsgo_depthmap_from_points.restype = C.c_uint32
        L.vsgo_depthmap_from_points.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                                vp, C.c_uint64, vp, C.c_uint64, vp, C.c_uint64, vp, vp, vp, vp]
        L.vsgo_hash.restype = C.c_uint32
        L.vsgo_hash.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]
        L.vsgo_alloc_world_to_block.argtypes = [vp, C.c_float, vp, vp]
        L.vsgo_ray_world_to_voxel.argtypes = [vp, C.c_float, vp, vp, vp]
        L.vsgo_dda.restype = C.c_int
        L.vsgo_dda.argtypes = [vp, vp, C.c_float, vp, C.c_int]
        L.vsgo_f32_to_f16.restype = C.c_uint16
        L.vsgo_f32_to_f16.argtypes = [C.c_float]
        L.vsgo_f16_to_f32.restype = C.c_float
        L.vsgo_f16_to_f32.argtypes = [C.c_uint16]
        L.vsgo_trilinear.restype = C.c_int
        L.vsgo_trilinear.argtypes = [vp, vp, C.c_float, vp, vp]
        L.vsgo_depth_range.argtypes = [vp, C.c_long, C.c_long, vp]
        L.vsgo_march_brackets.argtypes = [vp, C.c_int, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_double, vp, C.c_long, vp]
        L.vsgo_render_adjoint.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_double, vp, C.c_long, vp, vp, vp,
                                          vp, vp]
        L.vsgo_get_sdf.restype = C.c_double
        L.vsgo_get_sdf.argtypes = [vp, C.c_int64]
        L.vsgo_set_sdf.argtypes = [vp, C.c_int64, C.c_double]
        L.vsgo_set_voxels.argtypes = [vp, C.c_uint64, vp, vp, vp, vp]
        L.vsgo_stream_out.restype = C.c_uint64
        L.vsgo_stream_out.argtypes = [vp, vp, C.c_double, C.c_uint64, vp, vp, vp, vp, vp, vp, vp]
        L.vsgo_stream_in.argtypes = [vp, C.c_uint64, vp, vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def set_threads(n: int = 0) -> int:
    """OpenMP threads of the oracle's parallel loops (n <= 0: unchanged); returns the
    number in effect."""
    return int(lib().vsgo_set_threads(int(n)))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _pose(T) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(T, dtype=np.float64).reshape(16))


COUNTER_NAMES = ["valid_pixels", "new_blocks", "heap_fail", "num_allocated", "visible", "updated",
                 "heap_counter", "frames"]


class OracleScene:
    """The reference's hash/voxel state on the CPU (float or double arithmetic)."""

    def __init__(self, num_buckets: int, num_blocks: int, voxel: float, store_fp16: bool = False,
                 fp64: bool = False, directional: bool = False):
        self.voxel = voxel
        self.fp64 = fp64
        self._h = C.c_void_p(lib().vsgo_create(num_buckets, num_blocks, voxel, int(store_fp16), int(fp64)))
        if directional:
            lib().vsgo_set_directional(self._h, 1)

    def __del__(self):
        try:
            if self._h.value:
                lib().vsgo_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def reset(self) -> None:
        lib().vsgo_reset(self._h)

    def integrate(self, points, colors, trunc: float, pose, intr, visible_only: bool = False,
                  normals=None) -> dict:
        """One integrateFromScanData; `normals` (camera frame, like points) feed the
        directional extension when it is on."""
        pts = _f32(points)
        H, W = pts.shape[0], pts.shape[1]
        rgb = np.ascontiguousarray(colors, dtype=np.uint8)
        nrm = None if normals is None else _f32(normals)
        ctr = np.zeros(8, np.uint32)
        fx, fy, cx, cy = intr
        lib().vsgo_integrate(self._h, pts.ctypes.data, rgb.ctypes.data, None if nrm is None else nrm.ctypes.data, W,
                             H, float(trunc), _pose(pose).ctypes.data, fx, fy, cx, cy, int(visible_only),
                             ctr.ctypes.data)
        return dict(zip(COUNTER_NAMES, (int(v) for v in ctr)))

    def stage_times(self) -> tuple[float, float, float]:
        """Seconds of the last integrate's allocation, compaction and integration."""
        out = np.zeros(3, np.float64)
        lib().vsgo_stage_times(self._h, out.ctypes.data)
        return float(out[0]), float(out[1]), float(out[2])

    def export_directional(self, nblocks: int) -> dict:
        """Direction slots of the first nblocks blocks: id, weight, dtsdf, each [nblocks*512, 2]."""
        n = int(nblocks) * 512 * 2
        ids, ws, dts = np.zeros(n, np.uint8), np.zeros(n, np.uint8), np.zeros(n, np.float64)
        lib().vsgo_export_directional(self._h, int(nblocks), ids.ctypes.data, ws.ctypes.data, dts.ctypes.data)
        return {"id": ids.reshape(-1, 2), "weight": ws.reshape(-1, 2), "dtsdf": dts.reshape(-1, 2)}

    def alloc(self, points, trunc: float, pose) -> None:
        pts = _f32(points)
        H, W = pts.shape[0], pts.shape[1]
        lib().vsgo_alloc(self._h, pts.ctypes.data, W, H, float(trunc), _pose(pose).ctypes.data)

    def render(self, W: int, H: int, cam, pose, intr, minD: float, maxD: float, voxel: float | None = None,
               init: float = np.nan, pixels=None) -> dict:
        """Ray casts every pixel, or only `pixels` (int indices; the rest keep `init`)."""
        depth = np.full(W * H, init, np.float32)
        pos = np.full((W * H, 4), init, np.float32)
        col = np.full((W * H, 4), init, np.float32)
        nrm = np.full((W * H, 4), init, np.float32)
        hm = np.zeros(2, np.uint32)
        camd = np.ascontiguousarray(np.asarray(cam, np.float64).reshape(3))
        fx, fy, cx, cy = intr
        plist = None if pixels is None else np.ascontiguousarray(pixels, dtype=np.int32)
        ran = lib().vsgo_render(self._h, W, H, camd.ctypes.data, _pose(pose).ctypes.data, fx, fy, cx, cy, minD, maxD,
                                self.voxel if voxel is None else voxel, depth.ctypes.data, pos.ctypes.data,
                                col.ctypes.data, nrm.ctypes.data, hm.ctypes.data,
                                None if plist is None else plist.ctypes.data, 0 if plist is None else plist.size)
        return {"depth": depth, "pos": pos, "col": col, "nrm": nrm, "hits": int(hm[0]), "misses": int(hm[1]),
                "ran": bool(ran)}

    def render_adjoint(self, W: int, H: int, cam, pose, intr, minD: float, maxD: float, pixels,
                       grad_depth=None, voxel: float | None = None):
        """(J [n,6] = d depth / d xi for pose * exp(xi^), hit [n] bool, grad_sdf or None).
        self.last_depth receives the depth of the listed pixels in the oracle's arithmetic."""
        pix = np.ascontiguousarray(pixels, dtype=np.int32)
        n = pix.size
        J = np.zeros((n, 6), np.float64)
        hit = np.zeros(n, np.uint8)
        depth = np.zeros(n, np.float64)
        camd = np.ascontiguousarray(np.asarray(cam, np.float64).reshape(3))
        gd = gs = None
        if grad_depth is not None:
            gd = np.ascontiguousarray(grad_depth, dtype=np.float64)
            gs = np.zeros(self.num_allocated() * 512, np.float64)
        fx, fy, cx, cy = intr
        lib().vsgo_render_adjoint(self._h, W, H, camd.ctypes.data, _pose(pose).ctypes.data, fx, fy, cx, cy, minD, maxD,
                                  self.voxel if voxel is None else voxel, pix.ctypes.data, n, J.ctypes.data,
                                  hit.ctypes.data, depth.ctypes.data, None if gd is None else gd.ctypes.data,
                                  None if gs is None else gs.ctypes.data)
        self.last_depth = depth
        return J, hit.astype(bool), gs

    def march_brackets(self, W: int, cam, pose, intr, minD: float, maxD: float, pixels,
                       voxel: float | None = None) -> np.ndarray:
        """[n, 3] float64: the accepted bracket's ray parameters (a, b) and the refined
        alpha of each listed pixel, NaN where the ray misses."""
        pix = np.ascontiguousarray(pixels, dtype=np.int32)
        out = np.zeros((pix.size, 3), np.float64)
        camd = np.ascontiguousarray(np.asarray(cam, np.float64).reshape(3))
        fx, fy, cx, cy = intr
        if pix.size:
            lib().vsgo_march_brackets(self._h, W, camd.ctypes.data, _pose(pose).ctypes.data, fx, fy, cx, cy, minD,
                                      maxD, self.voxel if voxel is None else voxel, pix.ctypes.data, pix.size,
                                      out.ctypes.data)
        return out

    def depth_at(self, W, H, pose, intr, minD, maxD, pixels) -> np.ndarray:
        """Rendered depth of listed pixels in the oracle's arithmetic (float64 output)."""
        T = np.asarray(pose, np.float64)
        self.render_adjoint(W, H, T[:3, 3], T, intr, minD, maxD, pixels)
        return self.last_depth.copy()

    def get_sdf(self, idx: int) -> float:
        return float(lib().vsgo_get_sdf(self._h, int(idx)))

    def set_sdf(self, idx: int, value: float) -> None:
        lib().vsgo_set_sdf(self._h, int(idx), float(value))

    def set_voxels(self, idx, sdf, weight, rgb=None) -> None:
        """Overwrites SDF, weight and (rgb [n, 3] given) colour of voxels (global indices
        ptr*512 + voxel) of allocated blocks."""
        i = np.ascontiguousarray(idx, np.int64).ravel()
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sdf, np.float64).ravel(), i.shape))
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(weight, np.uint8).ravel(), i.shape))
        c = None if rgb is None else np.ascontiguousarray(np.asarray(rgb, np.uint8).reshape(i.size, 3))
        if i.size:
            lib().vsgo_set_voxels(self._h, i.size, i.ctypes.data, s.ctypes.data, w.ctypes.data,
                                  None if c is None else c.ctypes.data)

    def num_allocated(self) -> int:
        return int(lib().vsgo_num_allocated(self._h))

    def heap_counter(self) -> int:
        return int(lib().vsgo_heap_counter(self._h))

    def stream_out(self, center, radius: float, capacity: int = 1 << 30, directional: bool = False) -> dict:
        """Stream-out (VoxelStreamingManager.cu:36-91,258-358): the live blocks (table
        order) at distance >= radius from center, at most capacity, removed from the
        scene.  Returns coords int32 [m, 3], sdf float64 [m, 512], weight uint16 [m, 512],
        rgbf uint8 [m, 512, 4] and, with directional, dir_id / dir_w uint8 and dtsdf
        float64 [m, 512, 2]."""
        c = np.ascontiguousarray(center, np.float64).reshape(3)
        cap = int(min(capacity, max(self.num_allocated(), 1)))
        coords = np.zeros((cap, 3), np.int32)
        sdf = np.zeros((cap, 512), np.float64)
        weight = np.zeros((cap, 512), np.uint16)
        rgbf = np.zeros((cap, 512, 4), np.uint8)
        dir_id = np.zeros((cap, 512, 2), np.uint8)
        dir_w = np.zeros((cap, 512, 2), np.uint8)
        dtsdf = np.zeros((cap, 512, 2), np.float64)
        dptr = (dir_id.ctypes.data, dir_w.ctypes.data, dtsdf.ctypes.data) if directional else (None, None, None)
        m = int(lib().vsgo_stream_out(self._h, c.ctypes.data, float(radius), cap, coords.ctypes.data, sdf.ctypes.data,
                                      weight.ctypes.data, rgbf.ctypes.data, *dptr))
        out = {"coords": coords[:m], "sdf": sdf[:m], "weight": weight[:m], "rgbf": rgbf[:m]}
        if directional:
            out.update(dir_id=dir_id[:m], dir_w=dir_w[:m], dtsdf=dtsdf[:m])
        return out

    def stream_in(self, blocks: dict) -> np.ndarray:
        """Stream-in of the blocks (stream_out's layout) in order; status uint8 per block:
        0 inserted, 1 already allocated, 2 bucket full, 3 heap empty, 4 out of range."""
        coords = np.ascontiguousarray(blocks["coords"], np.int32)
        n = len(coords)
        sdf = np.ascontiguousarray(blocks["sdf"], np.float64)
        weight = np.ascontiguousarray(blocks["weight"], np.uint16)
        rgbf = np.ascontiguousarray(blocks["rgbf"], np.uint8)
        status = np.zeros(n, np.uint8)
        if "dir_id" in blocks:
            keep = [np.ascontiguousarray(blocks[k], t) for k, t in (("dir_id", np.uint8), ("dir_w", np.uint8),
                                                                     ("dtsdf", np.float64))]
            dptr = tuple(a.ctypes.data for a in keep)
        else:
            dptr = (None, None, None)
        if n:
            lib().vsgo_stream_in(self._h, n, coords.ctypes.data, sdf.ctypes.data, weight.ctypes.data,
                                 rgbf.ctypes.data, *dptr, status.ctypes.data)
        return status

    def export_blocks(self):
        """(coords int32 [n,3], ptrs int32 [n], hash entry indices int32 [n]) in pointer order."""
        n = int(lib().vsgo_export_blocks(self._h, None, None, None, 0))
        coords = np.zeros((n, 3), np.int32)
        ptrs = np.zeros(n, np.int32)
        slots = np.zeros(n, np.int32)
        if n:
            lib().vsgo_export_blocks(self._h, coords.ctypes.data, ptrs.ctypes.data, slots.ctypes.data, n)
        return coords, ptrs, slots

    def extract_surface_points(self, min_sdf: float = -0.01, max_sdf: float = 0.01, min_weight: int = 1,
                               max_points: int = 10_000_000) -> dict:
        """Surface points (RayCastRender.cu:684-868) in hash-entry then voxel order: the
        first max_points of them (positions and normals float64 [n, 3], colours uint8
        [n, 3]) and the number found (total)."""
        total = int(lib().vsgo_extract_surface_points(self._h, float(min_sdf), float(max_sdf), int(min_weight), 0,
                                                      None, None, None))
        n = min(total, int(max_points))
        pos = np.zeros((n, 3), np.float64)
        nrm = np.zeros((n, 3), np.float64)
        col = np.zeros((n, 3), np.uint8)
        if n:
            lib().vsgo_extract_surface_points(self._h, float(min_sdf), float(max_sdf), int(min_weight), n,
                                              pos.ctypes.data, nrm.ctypes.data, col.ctypes.data)
        return {"pos": pos, "nrm": nrm, "col": col, "total": total}

    def marching_cubes(self, table, iso: float = 0.0, voxel: float | None = None, cross_block: bool = False,
                       min_weight: int = 0, max_triangles: int = 50_000_000) -> dict:
        """Marching-cubes triangles (MarchingCubes.cu:352-467) with the given table (int8
        [256, 32]) in hash-entry, cube and table order: vertices and per-vertex normals
        float64 [n, 3, 3] of the first max_triangles, and the number found (total)."""
        tab = np.ascontiguousarray(table, np.int8)
        assert tab.shape == (256, 32)
        v = self.voxel if voxel is None else voxel
        args = (float(iso), float(v), int(bool(cross_block)), int(min_weight), tab.ctypes.data)
        total = int(lib().vsgo_marching_cubes(self._h, *args, 0, None, None))
        n = min(total, int(max_triangles))
        verts = np.zeros((n, 3, 3), np.float64)
        nrms = np.zeros((n, 3, 3), np.float64)
        if n:
            lib().vsgo_marching_cubes(self._h, *args, n, verts.ctypes.data, nrms.ctypes.data)
        return {"verts": verts, "nrms": nrms, "total": total}

    def export_voxels(self, nblocks: int | None = None):
        """(sdf float64 [n*512], weight uint8, colour uint8 [n*512,3], flags uint8) for pointers [0, n)."""
        if nblocks is None:
            nblocks = self.num_allocated()
        nv = nblocks * 512
        sdf = np.zeros(nv, np.float64)
        w = np.zeros(nv, np.uint8)
        c = np.zeros((nv, 3), np.uint8)
        f = np.zeros(nv, np.uint8)
        if nblocks:
            lib().vsgo_export_voxels(self._h, nblocks, sdf.ctypes.data, w.ctypes.data, c.ctypes.data, f.ctypes.data)
        return sdf, w, c, f

    def trilinear(self, p, voxel: float | None = None):
        p3 = _f32(np.asarray(p).reshape(3))
        d = C.c_float()
        col = np.zeros(3, np.uint8)
        ok = lib().vsgo_trilinear(self._h, p3.ctypes.data, self.voxel if voxel is None else voxel, C.byref(d),
                                  col.ctypes.data)
        return bool(ok), float(d.value), col


def hash_bucket(x: int, y: int, z: int, num_buckets: int) -> int:
    return int(lib().vsgo_hash(x, y, z, num_buckets))


def alloc_world_to_block(p, v: float):
    p3 = _f32(np.asarray(p).reshape(3))
    vox = np.zeros(3, np.int32)
    blk = np.zeros(3, np.int32)
    lib().vsgo_alloc_world_to_block(p3.ctypes.data, v, vox.ctypes.data, blk.ctypes.data)
    return vox, blk


def ray_world_to_voxel(p, v: float):
    p3 = _f32(np.asarray(p).reshape(3))
    vox = np.zeros(3, np.int32)
    blk = np.zeros(3, np.int32)
    loc = np.zeros(3, np.int32)
    lib().vsgo_ray_world_to_voxel(p3.ctypes.data, v, vox.ctypes.data, blk.ctypes.data, loc.ctypes.data)
    return vox, blk, loc


def dda(a, b, v: float, cap: int = 2048) -> np.ndarray:
    a3, b3 = _f32(np.asarray(a).reshape(3)), _f32(np.asarray(b).reshape(3))
    out = np.zeros((cap, 3), np.int32)
    n = lib().vsgo_dda(a3.ctypes.data, b3.ctypes.data, v, out.ctypes.data, cap)
    return out[:min(n, cap)].copy()


def f32_to_f16(x: float) -> int:
    return int(lib().vsgo_f32_to_f16(x))


def f16_to_f32(h: int) -> float:
    return float(lib().vsgo_f16_to_f32(h))


def depth_range(depth) -> tuple[float, float]:
    d = _f32(np.asarray(depth).reshape(-1))
    out = np.zeros(2, np.float32)
    lib().vsgo_depth_range(d.ctypes.data, d.size, 1, out.ctypes.data)
    return float(out[0]), float(out[1])


def se3_exp(xi) -> np.ndarray:
    """4x4 exp of xi = (omega, upsilon) (rotation vector, translation), float64."""
    xi = np.asarray(xi, np.float64)
    w, u = xi[:3], xi[3:]
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        R = np.eye(3) + K
        V = np.eye(3) + 0.5 * K
    else:
        A = np.sin(th) / th
        B = (1 - np.cos(th)) / th ** 2
        Cc = (th - np.sin(th)) / th ** 3
        R = np.eye(3) + A * K + B * K @ K
        V = np.eye(3) + B * K + Cc * K @ K
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ u
    return T


def depthmap_from_points(points, normals, colors, W: int, H: int, intrinsics, orthographic: bool = False) -> dict:
    """generateFromPointCloud (CustomDepthMapGenerator.cpp:81-222) on host arrays: points
    float32 [n, 3] (camera frame), normals float32 [m, 3] (m <= n; m == 0: finite
    differences), colours uint8 [k, 3] (k <= n).  Returns pointmap / normalmap float32
    [H*W, 3], depth float32 [H*W], colormap uint8 [H*W, 3] and the valid pixel count."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    nrm = np.ascontiguousarray(normals if normals is not None else np.zeros((0, 3)), np.float32).reshape(-1, 3)
    col = np.ascontiguousarray(colors if colors is not None else np.zeros((0, 3)), np.uint8).reshape(-1, 3)
    fx, fy, cx, cy = (float(v) for v in intrinsics)
    npix = W * H
    out = {"pointmap": np.zeros((npix, 3), np.float32), "depth": np.zeros(npix, np.float32),
           "normalmap": np.zeros((npix, 3), np.float32), "colormap": np.zeros((npix, 3), np.uint8)}
    valid = lib().vsgo_depthmap_from_points(int(W), int(H), fx, fy, cx, cy, 1 if orthographic else 0,
                                            pts.ctypes.data, len(pts), nrm.ctypes.data, len(nrm), col.ctypes.data,
                                            len(col), out["pointmap"].ctypes.data, out["depth"].ctypes.data,
                                            out["normalmap"].ctypes.data, out["colormap"].ctypes.data)
    out["valid"] = int(valid)
    return out
