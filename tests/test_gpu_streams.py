"""A renderer on a stream of its own (vsg_render_set_stream): the render is ordered
after the scene's queued work and the scene's later work after the render.  Renders,
hit and miss counts and the final scene state are bit-identical to the run with the
renderer on the scene's stream, with frames queued back to back (no host
synchronisation between them) and with other scene calls (counters, exports) between
a render and the next integration.
"""
import numpy as np
import pytest

from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _run(torch, sc, wl, frames, own_stream: bool, probe: bool):
    dev = torch.device("cuda:0")
    s_scene = torch.cuda.Stream()
    s_render = torch.cuda.Stream()
    scene = vsg.VoxelScene(vsg.Params(wl.buckets, wl.blocks, wl.voxel, wl.trunc), wl.store)
    scene.set_stream(s_scene.cuda_stream)
    rr = vsg.RayCastRender()
    assert rr.initialize(sc.W, sc.H)
    if own_stream:
        rr.set_stream(s_render.cuda_stream)
    fr = [sc.render(i) for i in frames]
    # every frame resident before the loop
    pts = [torch.from_numpy(f["points"]).to(dev) for f in fr]
    rgb = [torch.from_numpy(f["colors"]).to(dev) for f in fr]
    nrm = [torch.from_numpy(f["normals"]).to(dev) for f in fr]
    torch.cuda.synchronize()
    fx, fy, cx, cy = sc.intrinsics
    renders, counts = [], []
    for k, f in enumerate(fr):
        T = f["pose"]
        scene.integrateFromScanData(pts[k], rgb[k], nrm[k], sc.W, sc.H, wl.trunc, T, fx, fy, cx, cy)
        mn, mx = synth.depth_range(f["depth"])
        rr.render(scene, T[:3, 3], T, fx, fy, cx, cy, mn, mx)
        if probe and k % 3 == 1:
            # a scene call between a render and the next integration
            counts.append(("scene", k, scene.counters()))
        if probe and k == 7:
            counts.append(("blocks", k, len(scene.export_blocks()[0])))
        if k % 5 == 4:
            renders.append(rr.downloadAll())
            counts.append(("render", k, rr.counters()))
    torch.cuda.synchronize()
    coords, slots = scene.export_blocks()
    sdf, w, rgbf = scene.export_voxels()
    return renders, counts, (coords, slots, sdf, w, rgbf)


def test_render_on_own_stream_matches_one_stream(torch_cuda):
    torch = torch_cuda
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    frames = range(0, 30)
    for probe in (False, True):
        ref_r, ref_c, ref_s = _run(torch, sc, wl, frames, own_stream=False, probe=probe)
        r, c, st = _run(torch, sc, wl, frames, own_stream=True, probe=probe)
        assert len(r) == len(ref_r) == 6
        for a, b in zip(r, ref_r):
            for key in ("depth", "pos", "col", "nrm"):
                assert np.array_equal(_bits(a[key]), _bits(b[key])), (probe, key)
        # render counts and, with probes, the scene's counters and block count read
        # between a render and the next integration
        assert c == ref_c, probe
        for a, b in zip(st, ref_s):
            assert np.array_equal(np.asarray(a), np.asarray(b)), probe
