"""The reference's C++ class API (include/vsg_voxelscene.hpp) driven like the
reference's voxelSG.cpp loop, from a C++ executable, gives the same scene and
renders as the Python mirror over the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

from fusion_util import GpuRun
from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(synth.__file__), "lib", "vsg_shim_example")


@pytest.mark.parametrize("workload,frames", [(0, 6), (1, 3)])
def test_cpp_shim_matches_python_path(torch_cuda, workload, frames):
    out = subprocess.run([EXE, str(workload), str(frames)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    tok = out.stdout.split()
    got = {tok[i]: tok[i + 1] for i in range(0, len(tok) - 1, 2)}
    wl = synth.WORKLOADS[workload]
    sc = synth.SynthScene(workload)
    g = GpuRun(torch_cuda, wl)
    for i in range(frames):
        f = sc.render(i)
        g.integrate(f, sc.intrinsics)
    r = g.raycast(f, sc.intrinsics)
    d = r["depth"]
    k = np.arange(d.size)
    ok = d < 1e9
    checksum = float(np.sum(d[ok].astype(np.float64) * ((k[ok] % 97) + 1)))
    assert int(got["frames"]) == frames
    assert int(got["blocks"]) == g.scene.counters()["num_allocated"]
    assert (int(got["hits"]), int(got["misses"])) == (r["hits"], r["misses"])
    assert abs(float(got["checksum"]) - checksum) <= 1e-9 * abs(checksum)


def test_cpp_generator_path_matches_python(torch_cuda):
    """The C++ CustomDepthMapGenerator (device maps on the scene's stream, then integrate)
    gives the scene and renders the Python mirror gives for the same clouds: each frame's
    valid pixels, in pixel order, with their normals and colours."""
    frames = 3
    out = subprocess.run([EXE, "1", str(frames), "-", "gen"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    tok = out.stdout.split()
    got = {tok[i]: tok[i + 1] for i in range(0, len(tok) - 1, 2)}
    torch = torch_cuda
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    g = GpuRun(torch, wl)
    gen = vsg.CustomDepthMapGenerator()
    fx, fy, cx, cy = sc.intrinsics
    gen.setCameraParams(fx, fy, cx, cy, 0.1, 10.0, sc.W, sc.H)
    gen.set_stream(g.scene.stream())
    for i in range(frames):
        f = sc.render(i)
        ok = f["depth"].reshape(-1) > 0
        cloud = [torch.from_numpy(np.ascontiguousarray(f[k].reshape(-1, 3)[ok])).cuda()
                 for k in ("points", "normals", "colors")]
        d = gen.generate_device(*cloud)
        g.scene.integrateFromScanData(d.pointmap, d.colormap, d.normalmap, sc.W, sc.H, wl.trunc, f["pose"], fx, fy,
                                      cx, cy)
        torch.cuda.synchronize()
    r = g.raycast(f, sc.intrinsics)
    dd = r["depth"]
    k = np.arange(dd.size)
    hit = dd < 1e9
    checksum = float(np.sum(dd[hit].astype(np.float64) * ((k[hit] % 97) + 1)))
    assert int(got["blocks"]) == g.scene.counters()["num_allocated"]
    assert (int(got["hits"]), int(got["misses"])) == (r["hits"], r["misses"])
    assert abs(float(got["checksum"]) - checksum) <= 1e-9 * abs(checksum)


@pytest.mark.parametrize("cap", [100000, 400])
def test_cpp_streaming_manager_matches_python(torch_cuda, cap):
    """The C++ VoxelStreamingManager (stream-out beyond 1.2 m of the camera, then stream-in
    within it, before every integration from the second frame on) gives the scene,
    renders and chunk-grid contents the Python VoxelStreamingManager gives: blocks that
    integration re-creates beyond the radius replace their older images in both grids,
    and with a cap of 400 blocks per pass both collect chunks in parts."""
    frames, radius = 6, 1.2
    out = subprocess.run([EXE, "1", str(frames), "-", "-", "-", f"stream:{radius}:{cap}"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    tok = out.stdout.split()
    got = {tok[i]: tok[i + 1] for i in range(0, len(tok) - 1, 2)}
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    g = GpuRun(torch_cuda, wl)
    mgr = vsg.VoxelStreamingManager(g.scene, g.params, (1.0, 1.0, 1.0), (64, 64, 64), (-32, -32, -32))
    mgr.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = cap
    for i in range(frames):
        f = sc.render(i)
        if i > 0:
            cam = f["pose"][:3, 3].astype(np.float32)
            mgr.streamOut(g.scene, cam, radius)
            mgr.streamIn(g.scene, cam, radius)
        g.integrate(f, sc.intrinsics)
    r = g.raycast(f, sc.intrinsics)
    d = r["depth"]
    k = np.arange(d.size)
    ok = d < 1e9
    checksum = float(np.sum(d[ok].astype(np.float64) * ((k[ok] % 97) + 1)))
    assert mgr.numStreamedBlocks() > 0
    assert int(got["streamed"]) == mgr.numStreamedBlocks()
    keys = np.concatenate([c.live_keys() for c in mgr.m_grid.values()])
    assert len(np.unique(keys)) == len(keys)
    assert int(got["blocks"]) == g.scene.counters()["num_allocated"]
    assert (int(got["hits"]), int(got["misses"])) == (r["hits"], r["misses"])
    assert abs(float(got["checksum"]) - checksum) <= 1e-9 * abs(checksum)
