"""Renders with the pose in device memory (RayCastRender.h:66-70), the device-resident
Gauss-Newton loop, and the scene's device arrays (VoxelScene::getHashData), through
the C ABI.

- vsg_render_frame_device_pose gives bit-identical depth, positions, colours and
  counts to vsg_render_frame with the same pose on the host, also when the pose
  buffer is overwritten right after the call is queued on the same (explicit)
  stream (the kernels read it in stream order), and the adjoint of such a render
  equals the host-pose render's.
- vsg_pose_gn_refine (pose kept on the device between steps) ends at the same pose,
  bit for bit, as the same number of vsg_pose_gn_step calls (pose read back to the
  host after each).
- vsg_scene_hash_data's block array and bucket counts, copied from the device, equal
  the exported blocks and the per-bucket occupancy they imply.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from fusion_util import GpuRun
from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu


def _fused(torch, workload, frames):
    wl = synth.WORKLOADS[workload]
    sc = synth.SynthScene(workload)
    g = GpuRun(torch, wl)
    for i in frames:
        g.integrate(sc.render(i), sc.intrinsics)
    return wl, sc, g


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_device_pose_render_matches_host_pose(torch_cuda):
    torch = torch_cuda
    wl, sc, g = _fused(torch, 1, range(0, 20, 2))
    # an explicit stream shared by torch and the scene orders torch's writes after the
    # library's kernels (torch's default stream and the library's stream 0 are not
    # ordered with each other)
    s = torch.cuda.Stream()
    g.scene.set_stream(s.cuda_stream)
    f = sc.render(21)
    T = f["pose"].astype(np.float32)
    fx, fy, cx, cy = sc.intrinsics
    mn, mx = synth.depth_range(f["depth"])
    a, b = vsg.RayCastRender(), vsg.RayCastRender()
    assert a.initialize(sc.W, sc.H) and b.initialize(sc.W, sc.H)
    a.render(g.scene, T[:3, 3], T, fx, fy, cx, cy, mn, mx)
    ra = a.downloadAll()
    with torch.cuda.stream(s):
        dT = torch.from_numpy(T.reshape(16).copy()).to("cuda", non_blocking=False)
    s.synchronize()
    b.render_device_pose(g.scene, T[:3, 3], dT, fx, fy, cx, cy, mn, mx)
    # overwrite the buffer on the same stream right after the render was queued: the
    # kernels must have read the old values (stream order), march and tail pass alike
    with torch.cuda.stream(s):
        dT.fill_(0.0)
    rb = b.downloadAll()
    for key in ("depth", "pos", "col", "nrm"):
        assert np.array_equal(_bits(ra[key]), _bits(rb[key])), (key, int((_bits(ra[key]) != _bits(rb[key])).sum()))
    assert a.counters() == b.counters() and a.counters()[0] > 100000
    # the adjoint re-reads the pose: put it back first
    with torch.cuda.stream(s):
        dT.copy_(torch.from_numpy(T.reshape(16).copy()).to("cuda"))
    s.synchronize()
    n = sc.W * sc.H
    ja = torch.zeros(6 * n, dtype=torch.float32, device="cuda")
    jb = torch.zeros(6 * n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    a.adjoint(g.scene, jac=ja)
    b.adjoint(g.scene, jac=jb)
    s.synchronize()
    assert torch.equal(ja, jb) and float(ja.abs().max()) > 0
    g.scene.set_stream(0)


@pytest.mark.parametrize("jacobian", ["smoothed", "exact"])
def test_device_gauss_newton_loop_matches_host_steps(torch_cuda, jacobian):
    torch = torch_cuda
    wl, sc, g = _fused(torch, 1, range(0, 64, 2))
    rr = vsg.RayCastRender()
    assert rr.initialize(sc.W, sc.H)
    rr.set_sampling("trilinear")
    rr.set_gn_jacobian(jacobian)
    rr.set_gn_noise(0.0012, 0.0019, 0.4)
    fx, fy, cx, cy = sc.intrinsics
    rng = np.random.default_rng(3)
    sched = [0.2, 0.2, 0.2, 0.08, 0.08, 0.08, 0.03, 0.03, 0.03, 0.03]
    for i in (33, 47, 61):
        f = sc.render(i)
        Tgt = f["pose"].astype(np.float64)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        xi = np.concatenate([axis * np.radians(1.5), rng.normal(size=3) * 0.008])
        T0 = Tgt @ O.se3_exp(xi)
        obs = torch.from_numpy(f["depth"].reshape(-1)).cuda()
        mn, mx = synth.depth_range(f["depth"])
        Th = T0.copy()
        for res in sched:
            Th, st_h = rr.gn_step(g.scene, obs, fx, fy, cx, cy, mn, mx, Th, max_residual=res, lam=1e-3)
        Td, st_d = rr.gn_refine(g.scene, obs, fx, fy, cx, cy, mn, mx, T0, sched, lam=1e-3)
        assert np.array_equal(Th.view(np.uint64), Td.view(np.uint64)), (i, Th - Td)
        assert st_h == st_d
        E = np.linalg.inv(Tgt) @ Td
        print(f"frame {i} ({jacobian}): refined to {np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1))):.4f}"
              f" deg {np.linalg.norm(E[:3, 3]) * 1e3:.3f} mm from 1.5 deg")
    # one step: the same as vsg_pose_gn_step; zero steps: pose unchanged
    Tz, _ = rr.gn_refine(g.scene, obs, fx, fy, cx, cy, mn, mx, T0, [], lam=1e-3)
    assert np.array_equal(Tz, T0)


def _hip():
    L = C.CDLL("libamdhip64.so")
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemcpy.restype = C.c_int
    L.hipDeviceSynchronize.restype = C.c_int
    return L


def test_hash_data_arrays(torch_cuda):
    torch = torch_cuda
    wl, sc, g = _fused(torch, 0, range(0, 30, 3))
    h = g.scene.getHashData()
    coords, slots = g.scene.export_blocks()
    n = len(coords)
    assert h["num_buckets"] == wl.buckets and h["num_blocks"] == wl.blocks and h["voxel_bytes"] == 8
    assert h["index_entries"] >= 4 * wl.blocks and h["index_entries"] & (h["index_entries"] - 1) == 0
    hip = _hip()
    assert hip.hipDeviceSynchronize() == 0
    nal = np.zeros(1, np.uint32)
    assert hip.hipMemcpy(nal.ctypes.data, h["num_allocated_device"], 4, 2) == 0
    assert nal[0] == n
    blk = np.zeros((n, 4), np.int32)
    assert hip.hipMemcpy(blk.ctypes.data, h["blocks"], blk.nbytes, 2) == 0
    assert np.array_equal(blk[:, :3], coords) and np.array_equal(blk[:, 3], slots)
    used = np.zeros(wl.buckets, np.uint8)
    assert hip.hipMemcpy(used.ctypes.data, h["bucket_used"], used.nbytes, 2) == 0
    assert np.array_equal(np.bincount(slots // 10, minlength=wl.buckets), used)
    # the first live bucket's record: keys of its blocks in slot order, then their pointers
    b = int(slots[0] // 10)
    rec = np.zeros(16, np.uint64)
    assert hip.hipMemcpy(rec.ctypes.data, h["buckets"] + 128 * b, 128, 2) == 0
    ptrs = rec[10:15].view(np.int32)
    for j in range(int(used[b])):
        p = int(ptrs[j])
        x, y, z = (int(c) for c in coords[p])
        key = ((x & 0x1FFFFF) << 42) | ((y & 0x1FFFFF) << 21) | (z & 0x1FFFFF)
        assert int(rec[j]) == key and slots[p] == 10 * b + j
    # a voxel word of the first block equals the exported voxel
    sdf, w, _ = g.scene.export_voxels(0, 1)
    word = np.zeros(2, np.uint32)
    assert hip.hipMemcpy(word.ctypes.data, h["voxels"] + 8 * 77, 8, 2) == 0
    assert word[0] == sdf[77:78].view(np.uint32)[0] and (word[1] & 0xFFFF) == w[77]
