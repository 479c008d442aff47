"""Depth maps from point clouds (CustomDepthMapGenerator::generateFromPointCloud,
CustomDepthMapGenerator.cpp:81-222) on the GPU vs the oracle, through the C ABI.

Bar: pointmap, depth, normalmap and colormap bit-identical, the same valid-pixel count.
The clouds are the room workload's frames as camera-frame points, a second sensor's
points (a later frame's, moved into this camera's frame) appended, shuffled, with a
perturbed copy of a fifth of the points placed later in the array so that pixels are
contested and the last-writer rule decides them.
"""
import numpy as np
import pytest

import oracle as O
from fusion_util import GpuRun
from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu


def _cloud(sc, i, seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for j in (i, i + 3):
        f = sc.render(j)
        ok = f["depth"].reshape(-1) > 0
        p = f["points"].reshape(-1, 3)[ok].astype(np.float64)
        if j != i:  # into camera i's frame
            Ti, Tj = sc.render(i)["pose"].astype(np.float64), f["pose"].astype(np.float64)
            M = np.linalg.inv(Ti) @ Tj
            p = p @ M[:3, :3].T + M[:3, 3]
        parts.append((p.astype(np.float32), f["normals"].reshape(-1, 3)[ok], f["colors"].reshape(-1, 3)[ok]))
    pts = np.concatenate([q[0] for q in parts])
    nrm = np.concatenate([q[1] for q in parts])
    col = np.concatenate([q[2] for q in parts])
    perm = rng.permutation(len(pts))
    pts, nrm, col = pts[perm], nrm[perm], col[perm]
    k = len(pts) // 5
    dup = rng.choice(len(pts), k, replace=False)
    extra = pts[dup] * np.float32(1.0005)
    pts = np.concatenate([pts, extra]).astype(np.float32)
    nrm = np.concatenate([nrm, -nrm[dup]]).astype(np.float32)
    col = np.concatenate([col, 255 - col[dup]]).astype(np.uint8)
    return pts, nrm, col


def _gen(sc, projection=0):
    g = vsg.CustomDepthMapGenerator()
    fx, fy, cx, cy = sc.intrinsics
    g.setCameraParams(fx, fy, cx, cy, 0.1, 10.0, sc.W, sc.H)
    g.setProjectionMethod(projection)
    return g


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _compare(r, o):
    assert r.valid == o["valid"], (r.valid, o["valid"])
    assert _same(r.depthmap, o["depth"])
    assert _same(r.pointmap, o["pointmap"])
    assert _same(r.normalmap, o["normalmap"])
    assert _same(r.colormap, o["colormap"])


@pytest.mark.parametrize("case", ["full", "fd_normals", "partial", "orthographic"])
def test_depth_maps_match_oracle(torch_cuda, case):
    sc = synth.SynthScene(1)
    pts, nrm, col = _cloud(sc, 10)
    intr = sc.intrinsics
    proj = 0
    if case == "fd_normals":
        nrm = np.zeros((0, 3), np.float32)
    elif case == "partial":
        nrm, col = nrm[: len(pts) // 2], col[: len(pts) // 3]
    elif case == "orthographic":
        proj = 1
        pts = (pts * np.float32(100.0)).astype(np.float32)  # centimetres: spreads over the image
    g = _gen(sc, proj)
    r = g.generateFromPointCloud(pts, nrm, col)
    o = O.depthmap_from_points(pts, nrm, col, sc.W, sc.H, intr, orthographic=proj == 1)
    _compare(r, o)
    print(f"{case}: {len(pts)} points, {o['valid']} valid pixels, bit-identical")
    assert o["valid"] > 1000


def test_combined_clouds_and_empty_cloud(torch_cuda):
    sc = synth.SynthScene(1)
    a = _cloud(sc, 20, seed=1)
    b = _cloud(sc, 40, seed=2)
    g = _gen(sc)
    r = g.generateFromCombinedPointClouds(a, b)
    o = O.depthmap_from_points(*[np.concatenate([x, y]) for x, y in zip(a, b)], sc.W, sc.H, sc.intrinsics)
    _compare(r, o)
    e = g.generateFromPointCloud(np.zeros((0, 3), np.float32))
    assert e.valid == 0 and not e.depthmap.any() and not e.pointmap.any() and not e.normalmap.any()


def test_device_maps_feed_integration(torch_cuda):
    """Generator on the scene's stream -> integrate, with no host round trip, gives the
    scene that integrating the oracle's maps gives."""
    torch = torch_cuda
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    fx, fy, cx, cy = sc.intrinsics
    a, b = GpuRun(torch, wl), GpuRun(torch, wl)
    g = _gen(sc)
    g.set_stream(a.scene.stream())
    for i in (0, 6):
        pts, nrm, col = _cloud(sc, i, seed=i)
        T = sc.render(i)["pose"]
        d = g.generate_device(torch.from_numpy(pts).cuda(), torch.from_numpy(nrm).cuda(), torch.from_numpy(col).cuda())
        a.scene.integrateFromScanData(d.pointmap, d.colormap, d.normalmap, sc.W, sc.H, wl.trunc, T, fx, fy, cx, cy)
        torch.cuda.synchronize()
        o = O.depthmap_from_points(pts, nrm, col, sc.W, sc.H, sc.intrinsics)
        f = {"points": o["pointmap"].reshape(sc.H, sc.W, 3), "colors": o["colormap"].reshape(sc.H, sc.W, 3),
             "normals": o["normalmap"].reshape(sc.H, sc.W, 3), "depth": o["depth"].reshape(sc.H, sc.W), "pose": T}
        b.integrate(f, sc.intrinsics)
    ca, sa = a.blocks()
    cb, sb = b.blocks()
    assert np.array_equal(ca, cb) and np.array_equal(sa, sb) and len(ca) > 1000
    for x, y in zip(a.voxels(len(ca)), b.voxels(len(cb))):
        assert np.array_equal(x, y)
