"""Shared helpers of the fusion tests: run a frame sequence through the HIP path
(via the C ABI) or through the oracle, and compare the results."""
from __future__ import annotations

import numpy as np

import oracle as O
from vsg_amd import synth, vsg


def frames_of(scene: synth.SynthScene, idx) -> list[dict]:
    return [scene.render(i) for i in idx]


class GpuRun:
    """A VoxelScene + RayCastRender on cuda:0 fed from host frames."""

    def __init__(self, torch, wl: synth.Workload, store: str | None = None, buckets: int | None = None,
                 blocks: int | None = None, directional: bool = False):
        self.torch = torch
        self.wl = wl
        self.params = vsg.Params(buckets or wl.buckets, blocks or wl.blocks, wl.voxel, wl.trunc)
        self.scene = vsg.VoxelScene(self.params, store or wl.store, directional=directional)
        self.scene.set_stream(torch.cuda.current_stream().cuda_stream)
        self.render = None

    def upload(self, f: dict):
        t = self.torch
        return (t.from_numpy(f["points"]).cuda(), t.from_numpy(f["colors"]).cuda(), t.from_numpy(f["normals"]).cuda())

    def integrate(self, f: dict, intr) -> None:
        p, c, n = self.upload(f)
        H, W = f["depth"].shape
        fx, fy, cx, cy = intr
        self.scene.integrateFromScanData(p, c, n, W, H, self.wl.trunc, f["pose"], fx, fy, cx, cy)
        self.torch.cuda.synchronize()  # keep the uploaded tensors alive until the kernels ran

    def raycast(self, f: dict, intr) -> dict:
        H, W = f["depth"].shape
        if self.render is None:
            self.render = vsg.RayCastRender()
            assert self.render.initialize(W, H)
        mn, mx = synth.depth_range(f["depth"])
        fx, fy, cx, cy = intr
        T = f["pose"]
        self.render.render(self.scene, T[:3, 3], T, fx, fy, cx, cy, mn, mx)
        out = self.render.downloadAll()
        out["hits"], out["misses"] = self.render.counters()
        return out

    def blocks(self):
        return self.scene.export_blocks()

    def voxels(self, n: int):
        return self.scene.export_voxels(0, n)


class OracleRun:
    def __init__(self, wl: synth.Workload, store: str | None = None, buckets: int | None = None,
                 blocks: int | None = None, fp64: bool = False, directional: bool = False):
        self.wl = wl
        self.o = O.OracleScene(buckets or wl.buckets, blocks or wl.blocks, wl.voxel,
                               store_fp16=(store or wl.store) == "fp16", fp64=fp64, directional=directional)

    def integrate(self, f: dict, intr) -> dict:
        return self.o.integrate(f["points"], f["colors"], self.wl.trunc, f["pose"], intr, normals=f["normals"])

    def raycast(self, f: dict, intr, pixels=None) -> dict:
        H, W = f["depth"].shape
        mn, mx = synth.depth_range(f["depth"])
        T = f["pose"].astype(np.float64)
        return self.o.render(W, H, T[:3, 3], T, intr, mn, mx, pixels=pixels)


def compare_state(g: GpuRun, o: OracleRun, sdf_tol: float = 1e-5) -> dict:
    """Block set, pointers, hash slots, weights, colours, flags exact; SDF within sdf_tol."""
    gc, gs = g.blocks()
    oc, op, osl = o.o.export_blocks()
    assert len(gc) == len(oc), f"allocated blocks: gpu {len(gc)} vs oracle {len(oc)}"
    assert np.array_equal(op, np.arange(len(oc))), "oracle pointers are not 0..n-1"
    assert np.array_equal(gc, oc), "block coordinates per pointer differ"
    assert np.array_equal(gs, osl), "hash entry (bucket, slot) per block differs"
    n = len(gc)
    gsdf, gw, grgbf = g.voxels(n)
    osdf, ow, ocol, ofl = o.o.export_voxels(n)
    assert np.array_equal(gw.astype(np.int64), ow.astype(np.int64)), \
        f"weights differ at {np.count_nonzero(gw != ow)} voxels"
    assert np.array_equal(grgbf[:, :3], ocol), "colours differ"
    assert np.array_equal(grgbf[:, 3], ofl), "flags differ"
    err = np.abs(gsdf.astype(np.float64) - osdf)
    assert err.max(initial=0.0) <= sdf_tol, f"sdf error {err.max()} > {sdf_tol}"
    return {"blocks": n, "sdf_max_err": float(err.max(initial=0.0)),
            "sdf_bit_equal": bool(np.array_equal(gsdf, osdf.astype(np.float32)))}


def compare_raycast(gr: dict, orr: dict, pixels=None, tol: float = 1e-4, frac: float = 0.999) -> dict:
    gd = gr["depth"] if pixels is None else gr["depth"][pixels]
    od = orr["depth"] if pixels is None else orr["depth"][pixels]
    gv, ov = gd < 1e9, od < 1e9
    both = gv | ov
    n = int(both.sum())
    ok = gv & ov & (np.abs(gd - od) <= tol)
    good = int(ok.sum())
    assert n == 0 or good >= frac * n, f"raycast: {good}/{n} valid pixels within {tol} m"
    return {"valid": n, "within_tol": good, "max_err": float(np.abs(gd[gv & ov] - od[gv & ov]).max(initial=0.0)),
            "exact": int((gv & ov & (gd == od)).sum())}
