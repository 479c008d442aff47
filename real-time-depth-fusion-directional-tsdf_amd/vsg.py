"""Host mirror of the reference's VoxelScene / RayCastRender API over the C ABI.

Class and method names, argument order and meaning follow
VoxelSG/include/VoxelScene.h:52-209 and VoxelSG/include/RayCastRender.h:20-123;
the work runs in lib/libvsg.so (hand-written HIP kernels for gfx950) through the
C ABI of include/vsg.h.  Device buffers are passed as torch CUDA tensors (or raw
device addresses); PyTorch only provides device memory and streams here.

Error behaviour mirrors the reference where it has one: ``initialize`` returns
False on failure (RayCastRender.cpp:105-111); calls on an unallocated object
report and return.  Everything else raises :class:`VSGError` with the library's
message, and loading fails loudly when lib/libvsg.so is missing: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libvsg.so")

VSG_OK = 0
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "HIP_ERROR", 4: "NOT_READY"}
STORE_FP32 = 0
STORE_FP16 = 1
# timing ids of vsg_timing_read, in the C enum's order; "alloc_scan" is no longer
# launched (the insertion sums the workgroup counts itself) and reads zero launches
KERNELS = ["alloc_claim", "alloc_resolve", "alloc_scan", "alloc_insert", "compact", "integrate", "raycast",
           "reset", "adjoint", "gauss_newton", "extract", "depthgen", "mesh"]

SDF_BLOCK_SIZE = 8      # VoxelSG/include/globalPramasConfig.h:8
HASH_BUCKET_SIZE = 10   # VoxelSG/include/globalPramasConfig.h:9


class VSGError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"vsg status {STATUS.get(status, status)}: {msg}")
        self.status = status


class vsg_params(C.Structure):
    _fields_ = [("hash_slot_num", C.c_int32), ("sdf_block_num", C.c_int32), ("voxel_size", C.c_float),
                ("store", C.c_int32), ("flags", C.c_uint32), ("reserved", C.c_int32 * 3)]


SCENE_DIRECTIONAL = 1


class vsg_intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float)]


class vsg_frame(C.Structure):
    _fields_ = [("points", C.c_void_p), ("colors", C.c_void_p), ("normals", C.c_void_p),
                ("width", C.c_int32), ("height", C.c_int32)]


class vsg_counters(C.Structure):
    _fields_ = [("valid_pixels", C.c_uint32), ("new_blocks", C.c_uint32), ("heap_fail", C.c_uint32),
                ("num_allocated", C.c_uint32), ("visible_blocks", C.c_uint32), ("updated_voxels", C.c_uint32),
                ("claimed_buckets", C.c_uint32), ("out_of_range", C.c_uint32), ("frames", C.c_uint32),
                ("reserved", C.c_uint32 * 7)]


class vsg_hash_data(C.Structure):
    _fields_ = [("buckets", C.c_void_p), ("bucket_used", C.c_void_p), ("blocks", C.c_void_p), ("voxels", C.c_void_p),
                ("colors", C.c_void_p), ("block_index", C.c_void_p), ("num_allocated_device", C.c_void_p),
                ("num_buckets", C.c_uint32), ("num_blocks", C.c_uint32), ("voxel_bytes", C.c_uint32),
                ("index_entries", C.c_uint32), ("store", C.c_int32)]


class vsg_diag(C.Structure):
    _fields_ = [("alloc_buckets", C.c_uint32), ("ray_buckets", C.c_uint32), ("any_buckets", C.c_uint32),
                ("ray_blocks", C.c_uint32), ("march", C.c_uint32 * 8)]


# Every symbol include/vsg.h declares, with its ctypes signature.
_SIGNATURES = {
    "vsg_version": (C.c_int32, []),
    "vsg_last_error": (C.c_char_p, []),
    "vsg_scene_create": (C.c_int, [C.POINTER(vsg_params), C.POINTER(C.c_void_p)]),
    "vsg_scene_destroy": (C.c_int, [C.c_void_p]),
    "vsg_scene_reset": (C.c_int, [C.c_void_p]),
    "vsg_scene_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vsg_scene_get_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "vsg_integrate": (C.c_int, [C.c_void_p, C.POINTER(vsg_frame), C.c_float, C.POINTER(C.c_float),
                                C.POINTER(vsg_intrinsics)]),
    "vsg_integrate_noalloc": (C.c_int, [C.c_void_p, C.POINTER(vsg_frame), C.c_float, C.POINTER(C.c_float),
                                        C.POINTER(vsg_intrinsics)]),
    "vsg_scene_counters": (C.c_int, [C.c_void_p, C.POINTER(vsg_counters)]),
    "vsg_heap_free_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "vsg_num_integrated_frames": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "vsg_scene_params": (C.c_int, [C.c_void_p, C.POINTER(vsg_params)]),
    "vsg_scene_hash_data": (C.c_int, [C.c_void_p, C.POINTER(vsg_hash_data)]),
    "vsg_export_blocks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "vsg_export_directional": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsg_export_voxels": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsg_render_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "vsg_render_destroy": (C.c_int, [C.c_void_p]),
    "vsg_render_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vsg_render_set_sampling": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsg_render_set_step_budget": (C.c_int, [C.c_void_p, C.c_uint32]),
    "vsg_render_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(vsg_intrinsics), C.c_float, C.c_float, C.c_float]),
    "vsg_render_frame_device_pose": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_void_p,
                                               C.POINTER(vsg_intrinsics), C.c_float, C.c_float, C.c_float]),
    "vsg_render_buffers": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_void_p)] * 4),
    "vsg_render_download": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4),
    "vsg_render_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vsg_render_set_max_points": (C.c_int, [C.c_void_p, C.c_uint32]),
    "vsg_extract_surface_points": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_uint32,
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "vsg_extracted_points": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_void_p)] * 3 + [C.POINTER(C.c_uint32)]),
    "vsg_extracted_points_download": (C.c_int, [C.c_void_p] + [C.c_void_p] * 3 + [C.c_uint32,
                                                                                   C.POINTER(C.c_uint32)]),
    "vsg_depthgen_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "vsg_depthgen_destroy": (C.c_int, [C.c_void_p]),
    "vsg_depthgen_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vsg_depthgen_set_projection": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsg_depthgen_from_points": (C.c_int, [C.c_void_p, C.POINTER(vsg_intrinsics), C.c_void_p, C.c_uint32, C.c_void_p,
                                           C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "vsg_depthgen_valid_pixels": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "vsg_mc_table": (C.c_int, [C.c_void_p]),
    "vsg_mesh_create": (C.c_int, [C.c_uint32, C.c_void_p]),
    "vsg_mesh_destroy": (C.c_int, [C.c_void_p]),
    "vsg_marching_cubes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_int32, C.c_uint32, C.c_void_p,
                                     C.c_void_p]),
    "vsg_mesh_buffers": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsg_mesh_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "vsg_render_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsg_render_set_gn_jacobian": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsg_render_set_gn_noise": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "vsg_render_set_gn_window": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsg_pose_gn_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(vsg_intrinsics), C.c_float,
                                   C.c_float, C.c_float, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "vsg_pose_gn_refine": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(vsg_intrinsics), C.c_float,
                                     C.c_float, C.c_uint32, C.POINTER(C.c_float), C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "vsg_import_voxels": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsg_export_cell_masks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "vsg_block_image_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "vsg_stream_out": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_uint32)]),
    "vsg_stream_in": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_uint32)]),
    "vsg_diag_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsg_diag_convert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "vsg_diag_read": (C.c_int, [C.c_void_p, C.POINTER(vsg_diag), C.c_int32]),
    "vsg_timing_enable": (C.c_int, [C.c_int32]),
    "vsg_timing_reset": (C.c_int, []),
    "vsg_timing_read": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads lib/libvsg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with build_lib.py (no CPU fallback exists)")
        # PyTorch-ROCm ships its own HIP runtime and asks for it as "libamdhip64.so";
        # loaded after libvsg.so (which asks for "libamdhip64.so.7") it would be a second
        # runtime in the process, with its own streams.  Loading torch first makes
        # libvsg.so bind to torch's runtime, so stream handles and device pointers
        # cross between them.  Without torch the system runtime is used.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def _check(status: int) -> None:
    if status != VSG_OK:
        raise VSGError(status, lib().vsg_last_error().decode(errors="replace"))


def _dev_ptr(x) -> int:
    """Device address of a torch CUDA tensor (contiguous) or an int address."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_cuda:
            raise ValueError("device buffers must be CUDA tensors")
        if not x.is_contiguous():
            raise ValueError("device buffers must be contiguous")
        return int(x.data_ptr())
    raise TypeError(f"cannot take a device address of {type(x)!r}")


def _pose16(T) -> C.Array:
    a = np.ascontiguousarray(np.asarray(T, dtype=np.float32).reshape(16))
    return (C.c_float * 16)(*a.tolist())


class Params:
    """Params (VoxelSG/include/CUDAHashData.h:9-42).

    ``truncationScale``/``truncation`` keep the reference's names; the reference
    swaps them when filling from the config (CUDAHashData.h:27-28) and its
    integrate call takes the truncation as an argument instead (voxelSG.cpp:501).
    """

    def __init__(self, hashSlotNum: int, SDFBlockNum: int, voxelSize: float, truncation: float = 0.02,
                 truncationScale: float = 0.0):
        self.hashSlotNum = int(hashSlotNum)
        self.SDFBlockNum = int(SDFBlockNum)
        self.SDFBlockSize = SDF_BLOCK_SIZE
        self.slotSize = HASH_BUCKET_SIZE
        self.totalHashSize = self.hashSlotNum * HASH_BUCKET_SIZE
        self.totalBlockSize = self.SDFBlockNum * SDF_BLOCK_SIZE ** 3
        self.voxelSize = float(voxelSize)
        self.truncation = float(truncation)
        self.truncationScale = float(truncationScale)

    def __repr__(self) -> str:
        return (f"Params(hashSlotNum={self.hashSlotNum}, SDFBlockNum={self.SDFBlockNum}, "
                f"voxelSize={self.voxelSize}, truncation={self.truncation})")


class VoxelScene:
    """VoxelScene (VoxelSG/include/VoxelScene.h:52-209).

    store: "fp32" (f32 SDF + u16 weight) or "fp16" (f16 SDF + u16 weight).
    """

    def __init__(self, params: Params, store: str = "fp32", directional: bool = False):
        """directional: also keep the two direction slots per voxel of the
        directional TSDF extension (export_directional)."""
        self.m_params = params
        self.m_store = store
        p = vsg_params()
        p.hash_slot_num = params.hashSlotNum
        p.sdf_block_num = params.SDFBlockNum
        p.voxel_size = params.voxelSize
        p.store = {"fp32": STORE_FP32, "fp16": STORE_FP16}[store]
        p.flags = SCENE_DIRECTIONAL if directional else 0
        h = C.c_void_p()
        _check(lib().vsg_scene_create(C.byref(p), C.byref(h)))
        self._h = h
        self.m_isAllocated = True

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(lib().vsg_scene_destroy(self._h))
            self._h = C.c_void_p()
            self.m_isAllocated = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # -- reference API --------------------------------------------------------
    def reset(self) -> None:
        """VoxelScene::reset (VoxelSG/src/VoxelScene.cpp:276-302)."""
        if not self.m_isAllocated:
            print("VoxelScene: Not allocated yet!", file=sys.stderr)
            return
        _check(lib().vsg_scene_reset(self._h))

    def integrateFromScanData(self, depthmap, colormap, normalmap, width: int, height: int,
                              truncationDistance: float, cameraTransform, fx: float, fy: float, cx: float,
                              cy: float) -> None:
        """VoxelScene::integrateFromScanData (VoxelSG/src/VoxelScene.cpp:404-833).

        depthmap: device float3[H*W] camera-space points; colormap: device
        uchar3[H*W]; normalmap: device float3[H*W] (may be None);
        cameraTransform: host 4x4 row-major camera->world.
        """
        if depthmap is None or colormap is None:
            print("VoxelScene: Invalid pointers (nullptr)!", file=sys.stderr)
            return
        if not self.m_isAllocated:
            print("VoxelScene: Not allocated!", file=sys.stderr)
            return
        f = vsg_frame(_dev_ptr(depthmap), _dev_ptr(colormap), _dev_ptr(normalmap), int(width), int(height))
        k = vsg_intrinsics(fx, fy, cx, cy)
        _check(lib().vsg_integrate(self._h, C.byref(f), float(truncationDistance), _pose16(cameraTransform),
                                   C.byref(k)))

    def compactifyHashEntries(self) -> None:
        """Compaction runs inside every integrate call; nothing to do here."""

    def getHeapFreeCount(self) -> int:
        """VoxelScene::getHeapFreeCount (VoxelSG/src/VoxelScene.cpp:370-375)."""
        n = C.c_uint32()
        _check(lib().vsg_heap_free_count(self._h, C.byref(n)))
        return n.value

    def getNumIntegratedFrames(self) -> int:
        n = C.c_uint32()
        _check(lib().vsg_num_integrated_frames(self._h, C.byref(n)))
        return n.value

    def isAllocated(self) -> bool:
        return self.m_isAllocated

    def getParams(self) -> Params:
        return self.m_params

    def printStats(self) -> None:
        """VoxelScene::printStats (VoxelSG/src/VoxelScene.cpp:380-399)."""
        p = self.m_params
        print("\n=== VoxelScene Stats ===")
        print(f"Hash Slots: {p.hashSlotNum}")
        print(f"Bucket Size: {p.slotSize}")
        print(f"Total Hash Entries: {p.totalHashSize}")
        print(f"SDF Blocks: {p.SDFBlockNum}")
        print(f"Block Size: {p.SDFBlockSize}x{p.SDFBlockSize}x{p.SDFBlockSize}")
        print(f"Total Voxels: {p.totalBlockSize}")
        print(f"Integrated Frames: {self.getNumIntegratedFrames()}")
        if self.m_isAllocated:
            free = self.getHeapFreeCount()
            print(f"Free Blocks: {free} / {p.SDFBlockNum} ({free * 100.0 / p.SDFBlockNum}%)")
        print("========================\n")

    def saveToMesh(self, filename: str, iso: float = 0.0, min_weight: int = 1) -> bool:
        """VoxelScene::saveToMesh (a stub in the reference, VoxelSG/src/VoxelScene.cpp:838-845):
        marching cubes over the whole scene, cubes across block borders included and cubes
        with a corner of weight below min_weight left out, written as a binary PLY with the
        vertices shared by neighbouring triangles merged (equal float3) and one face per
        triangle.  False (and a message) when the scene has no triangles."""
        out = MarchingCubesOutput()
        allocateMarchingCubesOutput(out, 1 << 20)
        try:
            mc = MarchingCubesParams()
            mc.isoValue = iso
            mc.voxelSize = 0.0  # the scene's
            total = runMarchingCubesCUDA(self, self.m_params, mc, out, cross_block=True, min_weight=min_weight)
            if total > out.maxTriangles:
                allocateMarchingCubesOutput(out, total)
                total = runMarchingCubesCUDA(self, self.m_params, mc, out, cross_block=True, min_weight=min_weight)
            if total == 0:
                print("VoxelScene::saveToMesh: no triangles", file=sys.stderr)
                return False
            verts = downloadMarchingCubes(out)["verts"]
        finally:
            freeMarchingCubesOutput(out)
        write_mesh_ply(filename, verts)
        return True

    # -- extensions -----------------------------------------------------------
    def set_stream(self, hip_stream: int) -> None:
        _check(lib().vsg_scene_set_stream(self._h, C.c_void_p(hip_stream)))

    def stream(self) -> int:
        s = C.c_void_p()
        _check(lib().vsg_scene_get_stream(self._h, C.byref(s)))
        return s.value or 0

    def getHashData(self) -> dict:
        """VoxelScene::getHashData (VoxelScene.h:125-136): the scene's device arrays
        (addresses and sizes, this library's layout; vsg_hash_data in include/vsg.h)."""
        h = vsg_hash_data()
        _check(lib().vsg_scene_hash_data(self._h, C.byref(h)))
        return {name: getattr(h, name) for name, _ in vsg_hash_data._fields_}

    def counters(self) -> dict:
        c = vsg_counters()
        _check(lib().vsg_scene_counters(self._h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in vsg_counters._fields_ if name != "reserved"}

    def import_voxels(self, first: int, sdf: np.ndarray, weight: np.ndarray, rgbf: np.ndarray | None = None) -> None:
        """Overwrites the voxels of blocks [first, first + len(sdf)/512)."""
        sdf = np.ascontiguousarray(sdf, np.float32)
        weight = np.ascontiguousarray(weight, np.uint16)
        nb = sdf.size // 512
        r = None if rgbf is None else np.ascontiguousarray(rgbf, np.uint8)
        _check(lib().vsg_import_voxels(self._h, first, nb, sdf.ctypes.data, weight.ctypes.data,
                                       None if r is None else r.ctypes.data))

    def diag_enable(self, on: bool = True) -> None:
        _check(lib().vsg_diag_enable(self._h, 1 if on else 0))

    def diag_read(self, clear: bool = True) -> dict:
        d = vsg_diag()
        _check(lib().vsg_diag_read(self._h, C.byref(d), 1 if clear else 0))
        out = {name: getattr(d, name) for name, _ in vsg_diag._fields_ if name != "march"}
        out.update(zip(("march_evals", "march_super_skips", "march_block_skips", "march_replayed",
                        "march_cell_skips", "march_deferred", "march_tail_windows", "march_tail_skips"), d.march))
        return out

    def export_cell_masks(self, first: int = 0, nblocks: int | None = None):
        """(masks uint8 [n], counts uint8 [n, 8]) of blocks [first, first+n): bit c of a
        block's mask is set when its 4^3-voxel cell c holds a voxel with weight > 0 and
        SDF > 0; counts are those voxels per cell."""
        if nblocks is None:
            nblocks = self.m_params.SDFBlockNum - self.getHeapFreeCount() - first
        masks = np.zeros(nblocks, np.uint8)
        counts = np.zeros((nblocks, 8), np.uint8)
        if nblocks:
            _check(lib().vsg_export_cell_masks(self._h, first, nblocks, masks.ctypes.data, counts.ctypes.data))
        return masks, counts

    # -- streaming (VoxelStreamingManager.cu:36-91,258-358,443-719) -------------
    def block_image_bytes(self) -> int:
        n = C.c_uint32()
        _check(lib().vsg_block_image_bytes(self._h, C.byref(n)))
        return n.value

    def stream_out(self, center, radius: float, capacity: int = 100000, coords_out: np.ndarray | None = None,
                   images_out: np.ndarray | None = None):
        """Removes up to `capacity` blocks at distance >= radius from center (block
        world position = coordinate * 8 * voxel size).  Returns (coords int32 [m, 3],
        images uint8 [m, block_image_bytes()]), views of coords_out / images_out when
        given (C-contiguous, at least capacity rows; pinned host memory makes the copy
        from the device faster); decode_block_images reads the images."""
        c = np.ascontiguousarray(center, np.float32).reshape(3)
        nb = self.block_image_bytes()
        cap = int(capacity)
        coords = np.empty((cap, 3), np.int32) if coords_out is None else coords_out
        images = np.empty((cap, nb), np.uint8) if images_out is None else images_out
        if (coords.dtype != np.int32 or images.dtype != np.uint8 or not coords.flags.c_contiguous
                or not images.flags.c_contiguous or coords.shape[0] < cap or images.shape[0] < cap
                or coords.shape[1:] != (3,) or images.shape[1:] != (nb,)):
            raise VSGError(1, "stream_out: outputs must be C-contiguous int32 [>= capacity, 3] and uint8 "
                              "[>= capacity, block_image_bytes()]")
        n = C.c_uint32()
        _check(lib().vsg_stream_out(self._h, c.ctypes.data, float(radius), cap, coords.ctypes.data,
                                    images.ctypes.data, C.byref(n)))
        return coords[:n.value], images[:n.value]

    def stream_in(self, coords, images) -> np.ndarray:
        """Inserts blocks (stream_out's layout) in order; returns the status per block:
        0 inserted, 1 already allocated, 2 home bucket full, 3 no free block, 4 out of
        the key range."""
        coords = np.ascontiguousarray(coords, np.int32).reshape(-1, 3)
        images = np.ascontiguousarray(images, np.uint8)
        n = len(coords)
        if images.shape != (n, self.block_image_bytes()):
            raise VSGError(1, "stream_in: images must be [n, block_image_bytes()] bytes")
        status = np.zeros(n, np.uint8)
        got = C.c_uint32()
        _check(lib().vsg_stream_in(self._h, n, coords.ctypes.data if n else None,
                                   images.ctypes.data if n else None, status.ctypes.data if n else None,
                                   C.byref(got)))
        return status

    def export_blocks(self):
        """(coords int32 [n,3], hash entry indices int32 [n]) in pointer order."""
        n = C.c_uint32()
        _check(lib().vsg_export_blocks(self._h, None, None, 0, C.byref(n)))
        coords = np.zeros((n.value, 3), np.int32)
        slots = np.zeros(n.value, np.int32)
        if n.value:
            _check(lib().vsg_export_blocks(self._h, coords.ctypes.data, slots.ctypes.data, n.value, C.byref(n)))
        return coords, slots

    def export_voxels(self, first: int = 0, nblocks: int | None = None):
        """(sdf float32 [n*512], weight uint16 [n*512], rgbf uint8 [n*512,4]) of blocks [first, first+n)."""
        if nblocks is None:
            nblocks = self.m_params.SDFBlockNum - self.getHeapFreeCount() - first
        nv = nblocks * 512
        sdf = np.zeros(nv, np.float32)
        w = np.zeros(nv, np.uint16)
        rgbf = np.zeros((nv, 4), np.uint8)
        if nblocks:
            _check(lib().vsg_export_voxels(self._h, first, nblocks, sdf.ctypes.data, w.ctypes.data,
                                           rgbf.ctypes.data))
        return sdf, w, rgbf

    def export_directional(self, first: int = 0, nblocks: int | None = None) -> dict:
        """Direction slots of blocks [first, first+n): id, weight (uint8) and dtsdf
        (float32), each [n*512, 2] (id 0 free, 1..6 = +x, -x, +y, -y, +z, -z)."""
        if nblocks is None:
            nblocks = self.m_params.SDFBlockNum - self.getHeapFreeCount() - first
        nv = nblocks * 512
        ids = np.zeros((nv, 2), np.uint8)
        ws = np.zeros((nv, 2), np.uint8)
        dts = np.zeros((nv, 2), np.float32)
        if nblocks:
            _check(lib().vsg_export_directional(self._h, first, nblocks, ids.ctypes.data, ws.ctypes.data,
                                                dts.ctypes.data))
        return {"id": ids, "weight": ws, "dtsdf": dts}


class RayCastRender:
    """RayCastRender (VoxelSG/include/RayCastRender.h:20-123)."""

    def __init__(self):
        self.m_isAllocated = False
        self.m_width = 0
        self.m_height = 0
        self.m_numExtractedPoints = 0
        self.m_maxExtractedPoints = 10000000
        self._h = C.c_void_p()
        self._ext_h = C.c_void_p()

    def initialize(self, width: int, height: int) -> bool:
        """RayCastRender::initialize (VoxelSG/src/RayCastRender.cpp:88-118)."""
        if self.m_isAllocated:
            print("RayCastRender: Already initialized!", file=sys.stderr)
            return False
        h = C.c_void_p()
        st = lib().vsg_render_create(int(width), int(height), C.byref(h))
        if st != VSG_OK:
            print("RayCastRender: Failed to allocate GPU memory! "
                  + lib().vsg_last_error().decode(errors="replace"), file=sys.stderr)
            return False
        self._h = h
        _check(lib().vsg_render_set_max_points(h, self.m_maxExtractedPoints))
        self.m_width, self.m_height = int(width), int(height)
        self.m_isAllocated = True
        return True

    def set_stream(self, hip_stream: int) -> None:
        """Stream of this renderer's work (vsg_render_set_stream); by default the scene's.
        A render on another stream is ordered after the scene's queued work, and the
        scene's later work after the render."""
        if not self.m_isAllocated:
            raise RuntimeError("RayCastRender: not initialized")
        _check(lib().vsg_render_set_stream(self._h, C.c_void_p(int(hip_stream))))

    def close(self) -> None:
        if self._h.value:
            _check(lib().vsg_render_destroy(self._h))
            self._h = C.c_void_p()
            self.m_isAllocated = False
        if self._ext_h.value:
            _check(lib().vsg_render_destroy(self._ext_h))
            self._ext_h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, voxelScene: VoxelScene, cameraPos, cameraTransform, fx: float, fy: float, cx: float,
               cy: float, minDepth: float, maxDepth: float, voxelSize: float = 0.0) -> None:
        """RayCastRender::render (VoxelSG/src/RayCastRender.cpp:123-183).

        cameraPos: 3 floats; cameraTransform: host 4x4 row-major camera->world.
        voxelSize <= 0 uses the scene's (the reference reads the config's, which
        equals Params::voxelSize).
        """
        if not self.m_isAllocated:
            print("RayCastRender: Not initialized!", file=sys.stderr)
            return
        if voxelScene is None or not voxelScene.isAllocated():
            print("RayCastRender: Invalid voxel scene!", file=sys.stderr)
            return
        cp = np.asarray(cameraPos, dtype=np.float32).reshape(3)
        k = vsg_intrinsics(fx, fy, cx, cy)
        _check(lib().vsg_render_frame(self._h, voxelScene.handle, (C.c_float * 3)(*cp.tolist()),
                                      _pose16(cameraTransform), C.byref(k), float(minDepth), float(maxDepth),
                                      float(voxelSize)))

    def _download(self, depth=False, pos=False, col=False, nrm=False):
        n = self.m_width * self.m_height
        out = {}
        bufs = []
        for flag, name, shape in ((depth, "depth", (n,)), (pos, "pos", (n, 4)), (col, "col", (n, 4)),
                                  (nrm, "nrm", (n, 4))):
            if flag:
                out[name] = np.zeros(shape, np.float32)
                bufs.append(out[name].ctypes.data)
            else:
                bufs.append(None)
        _check(lib().vsg_render_download(self._h, *bufs))
        return out

    def downloadResults(self):
        """(depth float4 [n], colour float4 [n], normal float4 [n]) as the reference fills them
        (RayCastRender.cpp:188-214): depth replicated into xyz, w = 1."""
        r = self._download(depth=True, col=True, nrm=True)
        d = r["depth"]
        depth4 = np.stack([d, d, d, np.ones_like(d)], axis=1)
        return depth4, r["col"], r["nrm"]

    def downloadDepthFloat(self) -> np.ndarray:
        return self._download(depth=True)["depth"]

    def downloadAll(self) -> dict:
        """depth [n], positions [n,4], colours [n,4], normals [n,4]."""
        return self._download(True, True, True, True)

    def savePointCloudPLY(self, filePath: str) -> bool:
        """RayCastRender::savePointCloudPLY (VoxelSG/src/RayCastRender.cpp:221-302): ASCII PLY of
        the hit pixels' world positions and colours."""
        if not self.m_isAllocated:
            print("RayCastRender: Not initialized!", file=sys.stderr)
            return False
        r = self._download(depth=True, pos=True, col=True)
        ok = (r["depth"] < 1e9) & np.isfinite(r["pos"][:, :3]).all(axis=1)
        pts, cols = r["pos"][ok, :3], np.clip(r["col"][ok, :3] * 255.0, 0, 255).astype(np.uint8)
        with open(filePath, "w") as fh:
            fh.write("ply\nformat ascii 1.0\n")
            fh.write(f"element vertex {len(pts)}\n")
            fh.write("property float x\nproperty float y\nproperty float z\n")
            fh.write("property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n")
            for p, c in zip(pts, cols):
                fh.write(f"{p[0]} {p[1]} {p[2]} {c[0]} {c[1]} {c[2]}\n")
        return True

    def set_max_points(self, max_points: int) -> None:
        """Capacity of the surface-point extraction (m_maxExtractedPoints, RayCastRender.cpp:62)."""
        self.m_maxExtractedPoints = int(max_points)
        for h in (self._h, self._ext_h):
            if h:
                _check(lib().vsg_render_set_max_points(h, self.m_maxExtractedPoints))

    def _extractor(self) -> C.c_void_p:
        """The renderer extractions run on: the initialised one, else a 1x1 one of its own
        (the reference extracts without initialize())."""
        if self.m_isAllocated:
            return self._h
        if not self._ext_h:
            _check(lib().vsg_render_create(1, 1, C.byref(self._ext_h)))
            _check(lib().vsg_render_set_max_points(self._ext_h, self.m_maxExtractedPoints))
        return self._ext_h

    def extractSurfacePoints(self, voxelScene: VoxelScene, minSDF: float = -0.01, maxSDF: float = 0.01,
                             minWeight: int = 1) -> int:
        """RayCastRender::extractSurfacePoints (VoxelSG/src/RayCastRender.cpp:307-364): one
        point per voxel of weight >= max(1, minWeight) and SDF in [minSDF, maxSDF] that has a
        zero crossing towards a neighbour (or |sdf| < 0.001), with its normal and colour, in
        hash-entry order then voxel order.  Returns the number of points found, as the
        reference's counter does; the first getMaxPoints() of them are kept
        (m_numExtractedPoints)."""
        if voxelScene is None:
            print("RayCastRender::extractSurfacePoints: Invalid voxelScene!", file=sys.stderr)
            return 0
        stored, total = C.c_uint32(), C.c_uint64()
        _check(lib().vsg_extract_surface_points(self._extractor(), voxelScene.handle, float(minSDF), float(maxSDF),
                                                int(minWeight), C.byref(stored), C.byref(total)))
        self.m_numExtractedPoints = stored.value
        return total.value

    def extractedBuffers(self) -> dict:
        """Device pointers of the extracted points (float4 positions, colours, normals) and
        the number stored."""
        ptrs = [C.c_void_p() for _ in range(3)]
        n = C.c_uint32()
        h = self._h if self.m_isAllocated else self._ext_h
        if not h:
            return {"pos": 0, "col": 0, "nrm": 0, "count": 0}
        _check(lib().vsg_extracted_points(h, *[C.byref(x) for x in ptrs], C.byref(n)))
        return {"pos": ptrs[0].value or 0, "col": ptrs[1].value or 0, "nrm": ptrs[2].value or 0, "count": n.value}

    def downloadExtractedPoints(self) -> dict:
        """The stored points of the last extraction: pos, col, nrm as float32 [n, 4]."""
        n = self.extractedBuffers()["count"]
        out = {k: np.zeros((n, 4), np.float32) for k in ("pos", "col", "nrm")}
        got = C.c_uint32()
        _check(lib().vsg_extracted_points_download(self._h if self.m_isAllocated else self._ext_h, out["pos"].ctypes.data, out["col"].ctypes.data,
                                                   out["nrm"].ctypes.data, n, C.byref(got)))
        assert got.value == n
        return out

    def saveExtractedSurfacePointsPLY(self, filePath: str) -> bool:
        """RayCastRender::saveExtractedSurfacePointsPLY (VoxelSG/src/RayCastRender.cpp:369-446):
        ASCII PLY of the extracted points with finite positions: x y z nx ny nz red green blue,
        colour channels round(clamp(c, 0, 1) * 255) (255 when not finite)."""
        n = self.extractedBuffers()["count"]
        if n == 0:
            print("RayCastRender::saveExtractedSurfacePointsPLY: No points to save!", file=sys.stderr)
            return False
        r = self.downloadExtractedPoints()
        ok = np.isfinite(r["pos"][:, :3]).all(axis=1)
        d = os.path.dirname(os.path.abspath(filePath))
        os.makedirs(d, exist_ok=True)

        def chan(v):
            if not np.isfinite(v):
                return 255
            c = np.float32(min(max(np.float32(v), np.float32(0.0)), np.float32(1.0)))
            return int(np.float32(c * np.float32(255.0)) + np.float32(0.5))

        with open(filePath, "w") as fh:
            fh.write("ply\nformat ascii 1.0\ncomment Generated by RayCastRender::extractSurfacePoints\n")
            fh.write(f"element vertex {int(ok.sum())}\n")
            fh.write("property float x\nproperty float y\nproperty float z\n")
            fh.write("property float nx\nproperty float ny\nproperty float nz\n")
            fh.write("property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n")
            for p, nn, c in zip(r["pos"][ok], r["nrm"][ok], r["col"][ok]):
                fh.write(f"{p[0]:g} {p[1]:g} {p[2]:g} {nn[0]:g} {nn[1]:g} {nn[2]:g} "
                         f"{chan(c[0])} {chan(c[1])} {chan(c[2])}\n")
        return True

    def set_sampling(self, mode: str) -> None:
        """"reference" (the reference's half-voxel-shifted sampling, default) or "trilinear"."""
        _check(lib().vsg_render_set_sampling(self._h, {"reference": 0, "trilinear": 1}[mode]))

    def set_step_budget(self, steps: int) -> None:
        """Steps a ray marches with one lane before the cooperative tail pass takes
        it over (0: no tail pass).  Results do not depend on it."""
        _check(lib().vsg_render_set_step_budget(self._h, int(steps)))

    def adjoint(self, voxelScene: VoxelScene, jac=None, grad_depth=None, grad_sdf=None) -> None:
        """Adjoint of the last render: jac (device float[6*W*H]) receives d depth / d xi for
        pose * exp(xi^); grad_sdf (device float[N*512]) accumulates grad_depth^T d depth / d sdf."""
        _check(lib().vsg_render_adjoint(self._h, voxelScene.handle, _dev_ptr(jac), _dev_ptr(grad_depth),
                                        _dev_ptr(grad_sdf)))

    def set_gn_jacobian(self, mode: str) -> None:
        """"smoothed" (default: implicit-function derivative with the field gradient by central
        differences over +-1 voxel) or "exact" (the adjoint of the forward's computation)."""
        _check(lib().vsg_render_set_gn_jacobian(self._h, {"smoothed": 0, "exact": 1}[mode]))

    def set_gn_noise(self, a: float, b: float = 0.0, z0: float = 0.0) -> None:
        """Residual weights (a / sigma(z))^2 with sigma(z) = a + b (z - z0)^2 of the observed
        depth (a = 0: uniform)."""
        _check(lib().vsg_render_set_gn_noise(self._h, float(a), float(b), float(z0)))

    def set_gn_window(self, on: bool = True) -> None:
        """Projective window of the Gauss-Newton renders: each ray marches only within the
        step's max_residual of its observed depth (pixels without one are not marched)."""
        _check(lib().vsg_render_set_gn_window(self._h, 1 if on else 0))

    def gn_step(self, voxelScene: VoxelScene, observed_depth, fx, fy, cx, cy, minDepth, maxDepth, pose,
                max_residual: float = 0.05, lam: float = 1e-4):
        """One FP64 Gauss-Newton step on the device; returns (new pose float64 4x4, stats dict)."""
        P = (C.c_double * 16)(*np.asarray(pose, np.float64).reshape(16).tolist())
        st = (C.c_double * 8)()
        k = vsg_intrinsics(fx, fy, cx, cy)
        _check(lib().vsg_pose_gn_step(self._h, voxelScene.handle, _dev_ptr(observed_depth), C.byref(k),
                                      float(minDepth), float(maxDepth), float(max_residual), float(lam), P, st))
        return (np.array(P[:], np.float64).reshape(4, 4),
                {"cost": st[0], "count": int(st[1]), "solved": bool(st[2]), "rot": st[3], "trans": st[4]})

    def gn_refine(self, voxelScene: VoxelScene, observed_depth, fx, fy, cx, cy, minDepth, maxDepth, pose,
                  max_residuals, lam: float = 1e-4):
        """len(max_residuals) Gauss-Newton steps with the pose kept on the device between
        them (one read-back); returns (new pose float64 4x4, stats dict of the last step)."""
        P = (C.c_double * 16)(*np.asarray(pose, np.float64).reshape(16).tolist())
        st = (C.c_double * 8)()
        res = [float(r) for r in max_residuals]
        R = (C.c_float * max(1, len(res)))(*res)
        k = vsg_intrinsics(fx, fy, cx, cy)
        _check(lib().vsg_pose_gn_refine(self._h, voxelScene.handle, _dev_ptr(observed_depth), C.byref(k),
                                        float(minDepth), float(maxDepth), len(res), R, float(lam), P, st))
        return (np.array(P[:], np.float64).reshape(4, 4),
                {"cost": st[0], "count": int(st[1]), "solved": bool(st[2]), "rot": st[3], "trans": st[4]})

    def render_device_pose(self, voxelScene: VoxelScene, cameraPos, pose_device, fx: float, fy: float, cx: float,
                           cy: float, minDepth: float, maxDepth: float, voxelSize: float = 0.0) -> None:
        """RayCastRender::render with the pose as the reference passes it: a device
        buffer of 16 floats (torch CUDA tensor or address), read by the kernels."""
        if not self.m_isAllocated:
            print("RayCastRender: Not initialized!", file=sys.stderr)
            return
        if voxelScene is None or not voxelScene.isAllocated():
            print("RayCastRender: Invalid voxel scene!", file=sys.stderr)
            return
        cp = np.asarray(cameraPos, dtype=np.float32).reshape(3)
        k = vsg_intrinsics(fx, fy, cx, cy)
        _check(lib().vsg_render_frame_device_pose(self._h, voxelScene.handle, (C.c_float * 3)(*cp.tolist()),
                                                  _dev_ptr(pose_device), C.byref(k), float(minDepth),
                                                  float(maxDepth), float(voxelSize)))

    def counters(self):
        h, m = C.c_uint32(), C.c_uint32()
        _check(lib().vsg_render_counters(self._h, C.byref(h), C.byref(m)))
        return h.value, m.value

    def buffers(self) -> dict:
        ptrs = [C.c_void_p() for _ in range(4)]
        _check(lib().vsg_render_buffers(self._h, *[C.byref(p) for p in ptrs]))
        return dict(zip(("depth", "pos", "col", "nrm"), (p.value for p in ptrs)))

    def getOutputDepth(self) -> int:
        return self.buffers()["depth"]

    def getOutputColor(self) -> int:
        return self.buffers()["col"]

    def getOutputNormal(self) -> int:
        return self.buffers()["nrm"]

    def getOutputPositions(self) -> int:
        return self.buffers()["pos"]

    def getWidth(self) -> int:
        return self.m_width

    def getHeight(self) -> int:
        return self.m_height

    def isAllocated(self) -> bool:
        return self.m_isAllocated


class PoseTracker:
    """Frame-to-model pose refinement over an image pyramid (no counterpart in the
    reference): Gauss-Newton on the device (RayCastRender.gn_refine) at reduced, then
    full resolution, with the projective window on (RayCastRender.set_gn_window: each
    ray marches only within the iteration's residual window of its observed depth).
    Level d renders every d-th row and column: its image is ceil(W/d) x ceil(H/d) with
    intrinsics fx/d, fy/d, cx/d, cy/d, so its pixel (x, y) is the full image's (d x, d y)
    exactly, and the observed depth is the same subsample.

    levels: (d, residual windows) pairs run in order, one Gauss-Newton iteration per
    window; d a power of two, so that the level's rays are bit for bit the full image's
    ((x' - cx/d) / (fx/d) rounds as (d x' - cx) / fx).  The default (3 at 1/2 with 20, 8
    and 8 cm, 2 at full with 3 cm) reaches the accuracy of ten full-resolution iterations
    with windows 20, 20, 20, 8, 8, 8, 3, 3, 3, 3 cm (median and maximum pose error within
    0.0005 deg / 0.005 mm over 30 frames of the 1024x768 room sequence) in 2.5 ms instead
    of 6.6 ms (24.5 ms without the window).

    The subsampled depths are made on torch's current stream: use the scene's stream
    there (VoxelScene.set_stream) or synchronise the two."""

    LEVELS = ((2, (0.2, 0.08, 0.08)), (1, (0.03, 0.03)))

    def __init__(self, width: int, height: int, levels=LEVELS, sampling: str = "trilinear",
                 jacobian: str = "smoothed", noise=(0.0012, 0.0019, 0.4), window: bool = True):
        self.width, self.height = int(width), int(height)
        self.levels = tuple((int(d), tuple(float(r) for r in rs)) for d, rs in levels)
        self.renders = {}
        for d, _ in self.levels:
            if d < 1 or d & (d - 1):
                raise ValueError("pyramid level divisors must be powers of two")
            if d in self.renders:
                continue
            r = RayCastRender()
            if not r.initialize(-(-self.width // d), -(-self.height // d)):
                raise VSGError(2, "PoseTracker: renderer allocation failed")
            r.set_sampling(sampling)
            r.set_gn_jacobian(jacobian)
            if noise is not None:
                r.set_gn_noise(*noise)
            r.set_gn_window(window)
            self.renders[d] = r

    def refine(self, voxelScene: VoxelScene, observed_depth, fx: float, fy: float, cx: float, cy: float,
               minDepth: float, maxDepth: float, pose, lam: float = 1e-3):
        """observed_depth: device float32 tensor of H*W (or H x W) depths of the frame.
        Returns (refined pose float64 4x4, stats of the last iteration)."""
        obs = observed_depth.reshape(self.height, self.width)
        T = np.asarray(pose, np.float64)
        st = {"cost": 0.0, "count": 0, "solved": False, "rot": 0.0, "trans": 0.0}
        for d, windows in self.levels:
            if not windows:
                continue
            o = obs if d == 1 else obs[::d, ::d].contiguous()
            T, st = self.renders[d].gn_refine(voxelScene, o.reshape(-1), fx / d, fy / d, cx / d, cy / d, minDepth,
                                              maxDepth, T, windows, lam=lam)
        return T, st

    def close(self) -> None:
        for r in self.renders.values():
            r.close()
        self.renders = {}


# ---- marching cubes (VoxelSG/include/MarchingCubes.h, VoxelSG/src/MarchingCubes.cu) ------------

class MarchingCubesParams:
    """MarchingCubesParams (MarchingCubes.h:7-16), same defaults; voxelSize <= 0 selects the
    scene's voxel size (runMarchingCubesCUDA uses voxelSize when it is positive, and the
    default is 1.0)."""

    def __init__(self):
        self.isoValue = 0.0
        self.voxelSize = 1.0
        self.maxTriangles = 0


class MarchingCubesOutput:
    """MarchingCubesOutput (MarchingCubes.h:18-29): device buffers of float3 [3 * maxTriangles]
    (d_vertices, d_normals as device addresses) and the triangle count of the last run."""

    def __init__(self):
        self.d_vertices = 0
        self.d_normals = 0
        self.maxTriangles = 0
        self.triangleCount = 0
        self._h = C.c_void_p()


def allocateMarchingCubesOutput(output: MarchingCubesOutput, maxTriangles: int) -> None:
    """allocateMarchingCubesOutput (MarchingCubes.cu:474-489)."""
    if maxTriangles <= 0:
        print("allocateMarchingCubesOutput: maxTriangles must be positive.", file=sys.stderr)
        return
    freeMarchingCubesOutput(output)
    h = C.c_void_p()
    _check(lib().vsg_mesh_create(int(maxTriangles), C.byref(h)))
    output._h = h
    v, n = C.c_void_p(), C.c_void_p()
    _check(lib().vsg_mesh_buffers(h, C.byref(v), C.byref(n), None))
    output.d_vertices, output.d_normals = v.value or 0, n.value or 0
    output.maxTriangles = int(maxTriangles)
    output.triangleCount = 0


def freeMarchingCubesOutput(output: MarchingCubesOutput) -> None:
    """freeMarchingCubesOutput (MarchingCubes.cu:491-501)."""
    if output._h and output._h.value:
        _check(lib().vsg_mesh_destroy(output._h))
    output._h = C.c_void_p()
    output.d_vertices = output.d_normals = 0
    output.maxTriangles = 0
    output.triangleCount = 0


def runMarchingCubesCUDA(scene: "VoxelScene", params: "Params", paramsMC: MarchingCubesParams,
                         output: MarchingCubesOutput, cross_block: bool = False, min_weight: int = 0) -> int:
    """runMarchingCubesCUDA (MarchingCubes.cu:511-571) over every live block of the scene in
    hash-entry, cube and table order; voxel size paramsMC.voxelSize when positive, else the
    scene's (params.voxelSize).  cross_block adds the cubes across block borders (the
    reference's stay inside one block); min_weight > 0 leaves out cubes with a lighter
    corner.  Returns the triangles found (the reference's counter); min(found,
    maxTriangles) are stored."""
    if not (output._h and output._h.value):
        print("runMarchingCubesCUDA: Output buffers not allocated.", file=sys.stderr)
        return 0
    voxel = paramsMC.voxelSize if paramsMC.voxelSize > 0 else params.voxelSize
    stored, total = C.c_uint32(), C.c_uint64()
    _check(lib().vsg_marching_cubes(output._h, scene._h, float(paramsMC.isoValue), float(voxel), int(bool(cross_block)),
                                    int(min_weight), C.byref(stored), C.byref(total)))
    output.triangleCount = int(total.value)
    return int(total.value)


def copyMarchingCubesTriangleCount(output: MarchingCubesOutput) -> int:
    """copyMarchingCubesTriangleCount (MarchingCubes.cu:503-509): triangles found by the last run."""
    return output.triangleCount


def downloadMarchingCubes(output: MarchingCubesOutput) -> dict:
    """The stored triangles: vertices and normals float32 [n, 3, 3] (triangle, vertex, xyz)."""
    n = min(output.triangleCount, output.maxTriangles)
    verts = np.zeros((n, 3, 3), np.float32)
    nrms = np.zeros((n, 3, 3), np.float32)
    got = C.c_uint32()
    _check(lib().vsg_mesh_download(output._h, verts.ctypes.data, nrms.ctypes.data, n, C.byref(got)))
    assert got.value == n
    return {"verts": verts, "nrms": nrms}


def weld_vertices(verts: np.ndarray):
    """(unique vertices float32 [m, 3] in first-use order, faces int32 [n, 3]) of a triangle
    soup [n, 3, 3]: vertices with equal bits become one."""
    flat = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    keys = flat.view(np.uint32).view(np.dtype((np.void, 12))).reshape(-1)
    _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return flat[first[order]], rank[inv.reshape(-1)].reshape(-1, 3).astype(np.int32)


def write_mesh_ply(path: str, verts: np.ndarray) -> None:
    """Binary little-endian PLY of a triangle soup [n, 3, 3]: welded vertices (float x y z)
    and faces (uchar count 3, int indices)."""
    v, f = weld_vertices(verts)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    face = np.zeros(len(f), dtype=[("n", "u1"), ("i", "<i4", (3,))])
    face["n"] = 3
    face["i"] = f
    with open(path, "wb") as fh:
        fh.write((f"ply\nformat binary_little_endian 1.0\ncomment Generated by VoxelScene::saveToMesh\n"
                  f"element vertex {len(v)}\nproperty float x\nproperty float y\nproperty float z\n"
                  f"element face {len(f)}\nproperty list uchar int vertex_indices\nend_header\n").encode())
        fh.write(np.ascontiguousarray(v, "<f4").tobytes())
        fh.write(face.tobytes())


def read_mesh_ply(path: str):
    """(vertices float32 [m, 3], faces int32 [n, 3]) of a PLY written by write_mesh_ply."""
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().split("\n")
    nv = int(next(line for line in head if line.startswith("element vertex")).split()[-1])
    nf = int(next(line for line in head if line.startswith("element face")).split()[-1])
    v = np.frombuffer(data, "<f4", 3 * nv, end).reshape(nv, 3)
    face = np.frombuffer(data, [("n", "u1"), ("i", "<i4", (3,))], nf, end + 12 * nv)
    assert np.all(face["n"] == 3)
    return v.copy(), face["i"].copy()


class CameraParams:
    """CustomDepthMapGenerator::CameraParams (VoxelSG/include/CustomDepthMapGenerator.h:18-27),
    with the reference's defaults."""

    def __init__(self, fx=400.0, fy=400.0, cx=200.0, cy=240.0, nearPlane=0.1, farPlane=10.0, width=400, height=480):
        self.fx, self.fy, self.cx, self.cy = float(fx), float(fy), float(cx), float(cy)
        self.nearPlane, self.farPlane = float(nearPlane), float(farPlane)
        self.width, self.height = int(width), int(height)


class GeneratedDepthMap:
    """CustomDepthMapGenerator::GeneratedDepthMap (CustomDepthMapGenerator.h:37-45): depthmap
    [H*W], pointmap / normalmap [H*W, 3] float32, colormap [H*W, 3] uint8 (numpy arrays for
    host results, torch CUDA tensors for device results)."""

    def __init__(self, width=0, height=0, depthmap=None, pointmap=None, normalmap=None, colormap=None, valid=0):
        self.width, self.height = width, height
        self.depthmap, self.pointmap, self.normalmap, self.colormap = depthmap, pointmap, normalmap, colormap
        self.valid = valid


class CustomDepthMapGenerator:
    """CustomDepthMapGenerator (VoxelSG/include/CustomDepthMapGenerator.h:15-144) over the
    GPU generator of the C ABI (vsg_depthgen_*): the reference's generateFromPointCloud runs
    on the CPU (its CUDA path is a TODO, CustomDepthMapGenerator.cpp:593-611)."""

    PERSPECTIVE, ORTHOGRAPHIC, CUSTOM_CAMERA = 0, 1, 2

    def __init__(self):
        self.cameraParams_ = CameraParams()
        self.projectionMethod_ = self.PERSPECTIVE
        self._h = C.c_void_p()
        self._wh = (0, 0)
        self._stream = None

    def close(self) -> None:
        if self._h.value:
            _check(lib().vsg_depthgen_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setCameraParams(self, *args) -> None:
        """setCameraParams(CameraParams) or setCameraParams(fx, fy, cx, cy, nearPlane, farPlane,
        width, height) (CustomDepthMapGenerator.cpp:31-51)."""
        if len(args) == 1 and isinstance(args[0], CameraParams):
            self.cameraParams_ = args[0]
        else:
            self.cameraParams_ = CameraParams(*args)

    def getCameraParams(self) -> CameraParams:
        return self.cameraParams_

    def setProjectionMethod(self, method) -> None:
        """PERSPECTIVE (default), ORTHOGRAPHIC or CUSTOM_CAMERA (projects as PERSPECTIVE)."""
        if isinstance(method, str):
            method = {"perspective": 0, "orthographic": 1, "custom_camera": 2}[method.lower()]
        self.projectionMethod_ = int(method)

    def getProjectionMethod(self) -> int:
        return self.projectionMethod_

    def set_stream(self, hip_stream: int) -> None:
        """Stream of the generator's work (e.g. a scene's, VoxelScene.stream())."""
        self._stream = int(hip_stream)
        if self._h.value:
            _check(lib().vsg_depthgen_set_stream(self._h, self._stream))

    def _handle(self) -> C.c_void_p:
        wh = (self.cameraParams_.width, self.cameraParams_.height)
        if self._h.value and self._wh != wh:
            self.close()
        if not self._h.value:
            _check(lib().vsg_depthgen_create(wh[0], wh[1], C.byref(self._h)))
            self._wh = wh
            if self._stream is not None:
                _check(lib().vsg_depthgen_set_stream(self._h, self._stream))
        _check(lib().vsg_depthgen_set_projection(self._h, self.projectionMethod_))
        return self._h

    def generate_device(self, points, normals=None, colors=None, out: dict | None = None) -> GeneratedDepthMap:
        """generateFromPointCloud on torch CUDA tensors: points float32 [n, 3], normals float32
        [m, 3] (m <= n), colours uint8 [k, 3] (k <= n).  Returns device maps (the inputs of
        VoxelScene.integrateFromScanData); `out` may hold preallocated "pointmap",
        "depthmap", "normalmap", "colormap" tensors.  Asynchronous on the generator's
        stream; valid is filled by valid_pixels()."""
        import torch
        cp = self.cameraParams_
        npix = cp.width * cp.height
        dev = points.device
        if out is None:
            out = {"pointmap": torch.empty((npix, 3), dtype=torch.float32, device=dev),
                   "depthmap": torch.empty(npix, dtype=torch.float32, device=dev),
                   "normalmap": torch.empty((npix, 3), dtype=torch.float32, device=dev),
                   "colormap": torch.empty((npix, 3), dtype=torch.uint8, device=dev)}
        n = int(points.shape[0]) if points is not None else 0
        nn = int(normals.shape[0]) if normals is not None else 0
        nc = int(colors.shape[0]) if colors is not None else 0
        k = vsg_intrinsics(cp.fx, cp.fy, cp.cx, cp.cy)
        _check(lib().vsg_depthgen_from_points(self._handle(), C.byref(k), _dev_ptr(points) if n else None, n,
                                              _dev_ptr(normals) if nn else None, nn, _dev_ptr(colors) if nc else None,
                                              nc, _dev_ptr(out["pointmap"]), _dev_ptr(out.get("depthmap")),
                                              _dev_ptr(out["normalmap"]), _dev_ptr(out["colormap"])))
        return GeneratedDepthMap(cp.width, cp.height, out.get("depthmap"), out["pointmap"], out["normalmap"],
                                 out["colormap"])

    def valid_pixels(self) -> int:
        """Pixels with depth > 0 of the last generation (synchronising)."""
        c = C.c_uint32()
        _check(lib().vsg_depthgen_valid_pixels(self._handle(), C.byref(c)))
        return c.value

    def generateFromPointCloud(self, points, normals=None, colors=None) -> GeneratedDepthMap:
        """generateFromPointCloud (VoxelSG/src/CustomDepthMapGenerator.cpp:81-222) on host
        arrays (numpy): camera-frame points [n, 3], normals [m, 3] (empty: finite-difference
        normals), colours uint8 [k, 3].  Returns host maps."""
        import torch
        dev = torch.device("cuda")
        pts = torch.as_tensor(np.ascontiguousarray(points, np.float32).reshape(-1, 3)).to(dev)
        nrm = None if normals is None or len(normals) == 0 else \
            torch.as_tensor(np.ascontiguousarray(normals, np.float32).reshape(-1, 3)).to(dev)
        col = None if colors is None or len(colors) == 0 else \
            torch.as_tensor(np.ascontiguousarray(colors, np.uint8).reshape(-1, 3)).to(dev)
        torch.cuda.synchronize()
        g = self.generate_device(pts, nrm, col)
        valid = self.valid_pixels()
        return GeneratedDepthMap(g.width, g.height, g.depthmap.cpu().numpy(), g.pointmap.cpu().numpy(),
                                 g.normalmap.cpu().numpy(), g.colormap.cpu().numpy(), valid)

    def generateFromCombinedPointClouds(self, src0, src45) -> GeneratedDepthMap:
        """generateFromCombinedPointClouds (CustomDepthMapGenerator.cpp:613-643): src0's points,
        normals and colours followed by src45's.  Each source is (points, normals, colors)."""
        parts = [p for p in (src0, src45) if p is not None and p[0] is not None and len(p[0])]
        if not parts:
            print("CustomDepthMapGenerator: Both point clouds are empty!", file=sys.stderr)
            return GeneratedDepthMap()
        # each array is appended as it is, as the reference appends its vectors
        shapes = ((np.float32, 3), (np.float32, 3), (np.uint8, 3))
        cat = [np.concatenate([np.asarray(p[j] if p[j] is not None else np.zeros((0, w)), dt).reshape(-1, w)
                               for p in parts]) for j, (dt, w) in enumerate(shapes)]
        return self.generateFromPointCloud(*cat)


def mc_table() -> np.ndarray:
    """The marching-cubes triangle table, int8 [256, 32] (edge triples, then -1)."""
    out = np.zeros((256, 32), np.int8)
    _check(lib().vsg_mc_table(out.ctypes.data))
    return out


def diag_convert(values) -> np.ndarray:
    """The kernels' float -> int32 conversion (truncation, saturation, NaN -> 0) on the device."""
    x = np.ascontiguousarray(values, np.float32).reshape(-1)
    out = np.zeros(x.size, np.int32)
    _check(lib().vsg_diag_convert(x.ctypes.data, out.ctypes.data, x.size))
    return out


def timing_enable(on: bool = True) -> None:
    _check(lib().vsg_timing_enable(1 if on else 0))


def timing_reset() -> None:
    _check(lib().vsg_timing_reset())


def timing_read() -> dict:
    ms = (C.c_double * len(KERNELS))()
    n = (C.c_uint32 * len(KERNELS))()
    _check(lib().vsg_timing_read(ms, n))
    return {k: (ms[i], n[i]) for i, k in enumerate(KERNELS)}


# ---- streaming: VoxelStreamingManager (VoxelSG/include/VoxelStreamingManager.h:120-301) ----

def decode_block_images(images: np.ndarray, store: str = "fp32", directional: bool = False) -> dict:
    """Voxels of block images (vsg_block_image_bytes' layout): sdf float32 [m, 512],
    weight uint16 [m, 512], rgbf uint8 [m, 512, 4], and with directional the slot ids
    and weights uint8 [m, 512, 2] and dtsdf float32 [m, 512, 2]."""
    im = np.ascontiguousarray(images, np.uint8)
    m = im.shape[0]
    wb = 8 if store == "fp32" else 4
    vox = im[:, :512 * wb]
    if wb == 8:
        words = vox.view(np.uint32).reshape(m, 512, 2)
        sdf = words[:, :, 0].view(np.float32).copy()
        weight = (words[:, :, 1] & 0xFFFF).astype(np.uint16)
    else:
        words = vox.view(np.uint32).reshape(m, 512)
        sdf = (words & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float32)
        weight = (words >> 16).astype(np.uint16)
    off = 512 * wb
    out = {"sdf": sdf, "weight": weight, "rgbf": im[:, off:off + 2048].reshape(m, 512, 4).copy()}
    if directional:
        meta = im[:, off + 2048:off + 4096].reshape(m, 512, 4)
        out["dir_id"] = meta[:, :, 0::2].copy()
        out["dir_w"] = meta[:, :, 1::2].copy()
        out["dtsdf"] = im[:, off + 4096:off + 8192].copy().view(np.float32).reshape(m, 512, 2)
    return out


class _Chunk:
    """One chunk of the streaming grid: one image per block coordinate, in the order the
    coordinates first arrived; a coordinate that arrives again keeps its place and takes
    the newer image.  Coordinates are packed as the scene's keys (21 bits per axis,
    injective over the coordinates a scene can hold).  Rows [n0, n) of the buffers are
    live; the buffers grow by doubling, so an update copies only what it adds, and taking
    blocks from the front moves n0 instead of the rows."""

    def __init__(self, image_bytes: int):
        self.n0 = self.n = 0
        self.keys = np.zeros(16, np.int64)
        self.coords = np.zeros((16, 3), np.int32)
        self.images = np.zeros((16, image_bytes), np.uint8)

    def __len__(self) -> int:
        return self.n - self.n0

    def live_keys(self) -> np.ndarray:
        """Packed coordinates of the blocks the chunk holds, in order."""
        return self.keys[self.n0:self.n]

    @staticmethod
    def pack(coords) -> np.ndarray:
        c = np.asarray(coords, np.int64).reshape(-1, 3) & 0x1FFFFF
        return (c[:, 0] << 42) | (c[:, 1] << 21) | c[:, 2]

    def _room(self, m: int) -> None:
        if self.n + m <= len(self.keys):
            return
        live = len(self)
        cap = max(16, 2 * (live + m))
        keys, coords = np.empty(cap, np.int64), np.empty((cap, 3), np.int32)
        images = np.empty((cap, self.images.shape[1]), np.uint8)
        keys[:live] = self.keys[self.n0:self.n]
        coords[:live] = self.coords[self.n0:self.n]
        images[:live] = self.images[self.n0:self.n]
        self.keys, self.coords, self.images = keys, coords, images
        self.n0, self.n = 0, live

    def put(self, coords, images, rows) -> None:
        """Blocks rows of (coords, images): each image is copied once, from `images` into
        the chunk."""
        rows = np.asarray(rows, np.int64)
        k = self.pack(coords[rows])
        src = rows
        if len(np.unique(k)) < len(k):  # repeated within the batch: first place, last image
            u, first = np.unique(k, return_index=True)
            last = len(k) - 1 - np.unique(k[::-1], return_index=True)[1]
            o = np.argsort(first)
            k, rows, src = k[first[o]], rows[first[o]], rows[last[o]]
        live = self.keys[self.n0:self.n]
        if len(live):
            srt = np.argsort(live)
            j = np.minimum(np.searchsorted(live[srt], k), len(srt) - 1)
            hit = live[srt[j]] == k
            if hit.any():
                self.images[self.n0 + srt[j[hit]]] = images[src[hit]]
                k, rows, src = k[~hit], rows[~hit], src[~hit]
        m = len(k)
        if m:
            self._room(m)
            self.keys[self.n:self.n + m] = k
            self.coords[self.n:self.n + m] = coords[rows]
            np.take(images, src, axis=0, out=self.images[self.n:self.n + m])
            self.n += m

    def take_front(self, m: int):
        """Removes the first m blocks; returns views of their (coords, images), valid until
        the next put."""
        m = min(m, len(self))
        sl = slice(self.n0, self.n0 + m)
        self.n0 += m
        return self.coords[sl], self.images[sl]


class VoxelStreamingManager:
    """VoxelStreamingManager (VoxelSG/include/VoxelStreamingManager.h:120-301,
    VoxelSG/src/VoxelStreamingManager.cpp): blocks far from a sphere centre move from
    the scene (HBM) to a host chunk grid and come back when the sphere reaches them.

    The chunk grid has gridDimensions chunks of voxelExtents metres from minGridPos
    (chunk units).  A block belongs to the chunk of its world position (coordinate *
    8 * voxel size) rounded half away from zero (worldToChunks, .cpp:905-922); blocks
    outside the grid are dropped as the reference drops them (updateChunkGrid,
    .cpp:875-879) and counted in droppedOutsideGrid.  The four stages keep the
    reference's names; multiThreaded is accepted and ignored (every stage runs
    synchronously on the caller's thread).

    Two departures from the reference's chunk lists (chunkData, VoxelStreamingManager.h:
    78-118): a chunk holds one image per block coordinate, the newest (a block streamed
    out, re-created by integration and streamed out again replaces its older image,
    where the reference appends a second copy to the chunk's list); and a chunk larger
    than what one stream-in may still collect (m_maxNumberOfSDFBlocksIntegrateFromGlobalHash)
    is collected in parts over several calls, where the reference returns at that chunk
    (.cpp:1080-1085) and never collects it or any chunk after it."""

    def __init__(self, scene: VoxelScene, params: Params, voxelExtents, gridDimensions, minGridPos,
                 initialChunkListSize: int = 1000, streamingEnabled: bool = False):
        self.scene = scene
        self.params = params
        self.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = 100000  # VoxelStreamingManager.cpp:113
        self.initialize(voxelExtents, gridDimensions, minGridPos, initialChunkListSize, streamingEnabled)

    def initialize(self, voxelExtents, gridDimensions, minGridPos, initialChunkListSize: int = 1000,
                   streamingEnabled: bool = False) -> None:
        self.m_voxelExtents = np.asarray(voxelExtents, np.float32).reshape(3)
        self.m_gridDimensions = np.asarray(gridDimensions, np.int64).reshape(3)
        self.m_minGridPos = np.asarray(minGridPos, np.int64).reshape(3)
        self.m_maxGridPos = self.m_minGridPos + self.m_gridDimensions
        self.m_initialChunkListSize = int(initialChunkListSize)
        # chunk index -> _Chunk (one image per block coordinate, first-stream-out order)
        self.m_grid: dict[int, _Chunk] = {}
        self.droppedOutsideGrid = 0
        self.nStreamedOut = 0
        self.nStreamedIn = 0
        self._out = (np.zeros((0, 3), np.int32), np.zeros((0, 0), np.uint8))
        self._in = (np.zeros((0, 3), np.int32), np.zeros((0, 0), np.uint8))
        self._stage = None  # stream-out's host buffer, kept between calls
        # seconds spent in the scene's stream_out / stream_in calls (kernels and copies) and
        # in the chunk grid (host)
        self.seconds = {"scene_out": 0.0, "scene_in": 0.0, "grid": 0.0}

    # -- chunk geometry (VoxelStreamingManager.cpp:905-990) ----------------------
    def worldToChunks(self, posWorld) -> np.ndarray:
        p = np.asarray(posWorld, np.float32).reshape(-1, 3) / self.m_voxelExtents
        s = np.sign(p).astype(np.float32)
        return np.trunc(p + s * np.float32(0.5)).astype(np.int64)

    def blockWorldPos(self, coords) -> np.ndarray:
        return (np.asarray(coords, np.int64).reshape(-1, 3) * 8).astype(np.float32) * np.float32(self.params.voxelSize)

    def isValidChunk(self, chunk) -> np.ndarray:
        c = np.asarray(chunk, np.int64).reshape(-1, 3)
        return np.all((c >= self.m_minGridPos) & (c < self.m_maxGridPos), axis=1)

    def linearizeChunkPos(self, chunk) -> np.ndarray:
        p = np.asarray(chunk, np.int64).reshape(-1, 3) - self.m_minGridPos
        d = self.m_gridDimensions
        return p[:, 2] * d[0] * d[1] + p[:, 1] * d[0] + p[:, 0]

    def isChunkInSphere(self, chunk, center, radius: float) -> bool:
        cw = np.asarray(chunk, np.float32).reshape(3) * self.m_voxelExtents
        chunk_radius = np.float32(0.5) * np.float32(self.m_voxelExtents.max()) * np.float32(np.sqrt(np.float32(3.0)))
        d = cw - np.asarray(center, np.float32).reshape(3)
        dist = np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
        return bool(dist - chunk_radius <= np.float32(radius))

    # -- stream-out ---------------------------------------------------------------
    def streamOutFindBlocksOnGPU(self, sphereCenter, radius: float, useParts: bool = False,
                                 multiThreaded: bool = False) -> int:
        """Removes the blocks at distance >= radius from sphereCenter from the scene
        (at most m_maxNumberOfSDFBlocksIntegrateFromGlobalHash per call) and holds
        them for streamOutCopyToChunkGrid, in a host buffer the next call reuses (the
        reference's single output buffer).  Returns how many."""
        cap = self.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash
        nb = self.scene.block_image_bytes()
        if self._stage is None or self._stage[0].shape[0] < cap or self._stage[1].shape[1] != nb:
            self._stage = (np.empty((cap, 3), np.int32), np.empty((cap, nb), np.uint8))
        t0 = time.perf_counter()
        coords, images = self.scene.stream_out(sphereCenter, radius, cap, self._stage[0], self._stage[1])
        self.seconds["scene_out"] += time.perf_counter() - t0
        self._out = (coords, images)  # views of the staging buffer
        self.nStreamedOut += len(coords)
        return len(coords)

    def streamOutCopyToChunkGrid(self, multiThreaded: bool = False) -> None:
        coords, images = self._out
        t0 = time.perf_counter()
        self.updateChunkGrid(coords, images)
        self.seconds["grid"] += time.perf_counter() - t0
        self._out = (np.zeros((0, 3), np.int32), np.zeros((0, 0), np.uint8))

    def updateChunkGrid(self, coords, images) -> None:
        if len(coords) == 0:
            return
        chunks = self.worldToChunks(self.blockWorldPos(coords))
        valid = self.isValidChunk(chunks)
        self.droppedOutsideGrid += int((~valid).sum())
        if not valid.any():
            return
        coords = np.asarray(coords, np.int32).reshape(-1, 3)
        images = np.asarray(images, np.uint8).reshape(len(valid), -1)
        keep = np.flatnonzero(valid)
        idx = self.linearizeChunkPos(chunks[keep])
        order = np.argsort(idx, kind="stable")
        cut = np.flatnonzero(np.diff(idx[order])) + 1
        for part in np.split(order, cut):
            ci = int(idx[part[0]])
            chunk = self.m_grid.get(ci)
            if chunk is None:
                chunk = self.m_grid[ci] = _Chunk(images.shape[1])
            chunk.put(coords, images, keep[part])

    # -- stream-in ------------------------------------------------------------------
    def streamInFindChunk(self, sphereCenter, radius: float, useParts: bool = False) -> int:
        """Collects the blocks of the streamed-out chunks that the sphere touches (the
        chunk box around the centre's chunk, isChunkInSphere), removing them from the
        grid (VoxelStreamingManager.cpp:993-1130), at most
        m_maxNumberOfSDFBlocksIntegrateFromGlobalHash of them: a chunk that does not fit
        whole gives what fits and keeps the rest for the next call.  With useParts, one
        chunk (or the part that fits) per call.  Returns how many."""
        cam = self.worldToChunks(sphereCenter)[0]
        rad = np.ceil(np.float32(radius) / self.m_voxelExtents).astype(np.int64)
        lo = np.maximum(cam - rad, self.m_minGridPos)
        hi = np.minimum(cam + rad, self.m_maxGridPos - 1)
        cs, ims, got = [], [], 0
        cap = self.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash
        for x in range(lo[0], hi[0] + 1):
            for y in range(lo[1], hi[1] + 1):
                for z in range(lo[2], hi[2] + 1):
                    ci = int(self.linearizeChunkPos((x, y, z))[0])
                    chunk = self.m_grid.get(ci)
                    if chunk is None or not self.isChunkInSphere((x, y, z), sphereCenter, radius):
                        continue
                    c, im = chunk.take_front(cap - got)
                    cs.append(c)
                    ims.append(im)
                    got += len(c)
                    if len(chunk) == 0:
                        del self.m_grid[ci]
                    if useParts or got == cap:
                        return self._set_in(cs, ims)
        return self._set_in(cs, ims)

    def _set_in(self, cs, ims) -> int:
        nb = self.scene.block_image_bytes()
        self._in = (np.concatenate(cs) if cs else np.zeros((0, 3), np.int32),
                    np.concatenate(ims) if ims else np.zeros((0, nb), np.uint8))
        return len(self._in[0])

    def streamInCopyToGPUBuffer(self, sphereCenter, radius: float, useParts: bool = False,
                                multiThreaded: bool = False) -> int:
        t0 = time.perf_counter()
        n = self.streamInFindChunk(sphereCenter, radius, useParts)
        self.seconds["grid"] += time.perf_counter() - t0
        return n

    def streamInInsertToHashTable(self, multiThreaded: bool = False) -> np.ndarray:
        """Inserts the collected blocks; those that could not be inserted (bucket full,
        no free block) go back to the chunk grid.  Returns the status per block."""
        coords, images = self._in
        self._in = (np.zeros((0, 3), np.int32), np.zeros((0, 0), np.uint8))
        if len(coords) == 0:
            return np.zeros(0, np.uint8)
        t0 = time.perf_counter()
        status = self.scene.stream_in(coords, images)
        self.seconds["scene_in"] += time.perf_counter() - t0
        self.nStreamedIn += int((status == 0).sum())
        back = (status == 2) | (status == 3)
        if back.any():
            self.updateChunkGrid(coords[back], images[back])
        return status

    # -- one-call forms (VoxelStreamingManager.h: streamOut / streamIn) -------------
    def streamOut(self, scene: VoxelScene, cameraPos, radius: float) -> int:
        """Moves every block at distance >= radius from cameraPos to the chunk grid."""
        total = 0
        while True:
            n = self.streamOutFindBlocksOnGPU(cameraPos, radius)
            self.streamOutCopyToChunkGrid()
            total += n
            if n < self.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash:
                return total

    def streamIn(self, scene: VoxelScene, cameraPos, radius: float) -> np.ndarray:
        """Moves the blocks of the chunks the sphere touches back into the scene, in as
        many collect + insert passes as the cap needs (a pass that fills the cap and
        inserts something is followed by another).  Returns the status per block over
        all passes."""
        out = []
        while True:
            n = self.streamInCopyToGPUBuffer(cameraPos, radius)
            out.append(self.streamInInsertToHashTable())
            if n < self.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash or not np.any(out[-1] == 0):
                return np.concatenate(out)

    def numStreamedBlocks(self) -> int:
        return sum(len(c) for c in self.m_grid.values())
