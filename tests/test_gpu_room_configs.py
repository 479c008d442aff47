"""The 640x480 room sequence at its configured lengths, HIP path vs the oracle.

- Plain method, full size: all 300 frames with the FP32 store; counters every frame,
  the whole store (block set, pointers, hash slots, weights, colours, flags, SDF) at
  frames 60, 150 and 300, and the raycast of the last frame on 8192 seeded pixels.
- Precision: the same 300 frames with the FP16 store (FP32 arithmetic), bit-identical
  to the oracle's FP16 mode, and both stores against an FP64 oracle: SDF error on the
  voxels both have weighted, raycast depth error on pixels valid in both, split by
  whether the FP64 march accepted the same sample pair (the error of a pixel whose
  accepted pair changes is a bracket or silhouette change, not storage precision).
- Differentiation: 64 frames fused at their poses, the Jacobian of a held-out frame's
  render against the oracle's FP64 adjoint and central differences (h = 1e-4) on 4096
  seeded pixels, and the SDF gradient against the oracle.

Bounds sit below what the oracle pair measures on these frames (CPU run of the same
code, 300 frames, 8192 pixels): 99.0 % of pixels within 0.5 mm of FP64 for both
stores, 99.9 % of the same-bracket pixels, FP16 vs FP32 99.99 %, SDF RMS 0.022 mm.
"""
import numpy as np
import pytest

import oracle as O
from fusion_util import GpuRun, OracleRun, compare_raycast, compare_state
from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu


def _bit_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def test_room_all_300_frames_match_oracle(torch_cuda):
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    assert sc.nframes == 300
    g, o = GpuRun(torch_cuda, wl), OracleRun(wl)
    for i in range(sc.nframes):
        f = sc.render(i)
        g.integrate(f, sc.intrinsics)
        oc = o.integrate(f, sc.intrinsics)
        gc = g.scene.counters()
        assert (gc["num_allocated"], gc["new_blocks"], gc["valid_pixels"], gc["updated_voxels"], gc["heap_fail"]) == \
            (oc["num_allocated"], oc["new_blocks"], oc["valid_pixels"], oc["updated"], oc["heap_fail"]), (i, gc, oc)
        if i in (59, 149, 299):
            print(f"frame {i + 1}: {compare_state(g, o)}", flush=True)
    rng = np.random.default_rng(13)
    pix = rng.choice(sc.W * sc.H, 8192, replace=False).astype(np.int32)
    gr, orr = g.raycast(f, sc.intrinsics), o.raycast(f, sc.intrinsics, pixels=pix)
    info = compare_raycast(gr, orr, pixels=pix)
    for key in ("depth", "pos", "col"):
        assert _bit_equal(gr[key][pix], orr[key][pix]), key
    print(f"frame 300 raycast: {info}")


def _common_sdf_error(gpu_run, o64):
    """|SDF - SDF_fp64| (metres) over voxels of common blocks weighted in both."""
    coords, _ = gpu_run.blocks()
    sdf, w, _ = gpu_run.voxels(len(coords))
    c64, _, _ = o64.export_blocks()
    s64, w64, _, _ = o64.export_voxels()
    key = {tuple(c): i for i, c in enumerate(c64)}
    pairs = np.array([(i, key[tuple(c)]) for i, c in enumerate(coords) if tuple(c) in key])
    a = sdf.reshape(-1, 512)[pairs[:, 0]].astype(np.float64)
    b = s64.reshape(-1, 512)[pairs[:, 1]]
    m = (w.reshape(-1, 512)[pairs[:, 0]] > 0) & (w64.reshape(-1, 512)[pairs[:, 1]] > 0)
    return np.abs(a - b)[m], len(coords), len(c64), len(pairs)


def _depth_stats(d, d64, same):
    both = (d < 1e9) & (d64 < 1e9)
    err = np.abs(d.astype(np.float64) - d64)
    return {"valid_both": int(both.sum()), "valid_either": int(((d < 1e9) | (d64 < 1e9)).sum()),
            "within_0.5mm": float(np.mean(err[both] <= 5e-4)),
            "same_bracket": int((both & same).sum()),
            "same_within_0.5mm": float(np.mean(err[both & same] <= 5e-4)),
            "same_max_mm": float(err[both & same].max() * 1e3),
            "p999_mm": float(np.quantile(err[both], 0.999) * 1e3), "max_mm": float(err[both].max() * 1e3)}


def test_room_fp16_store_against_fp16_oracle_and_fp64(torch_cuda):
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    g16, g32 = GpuRun(torch_cuda, wl, store="fp16"), GpuRun(torch_cuda, wl, store="fp32")
    r16, r64 = OracleRun(wl, store="fp16"), OracleRun(wl, fp64=True)
    o16, o64 = r16.o, r64.o
    for i in range(sc.nframes):
        f = sc.render(i)
        g16.integrate(f, sc.intrinsics)
        g32.integrate(f, sc.intrinsics)
        # visible-only integration equals the all-blocks update (pinned on the CPU) and is faster
        o16.integrate(f["points"], f["colors"], wl.trunc, f["pose"], sc.intrinsics, visible_only=True)
        o64.integrate(f["points"], f["colors"], wl.trunc, f["pose"], sc.intrinsics, visible_only=True)
        if i % 100 == 99:
            print(f"frame {i + 1} integrated", flush=True)

    # FP16 store: bit-identical to the oracle's FP16 mode
    info = compare_state(g16, r16, sdf_tol=0.0)
    assert info["sdf_bit_equal"]
    rng = np.random.default_rng(12)
    pix = rng.choice(sc.W * sc.H, 8192, replace=False).astype(np.int32)
    T = f["pose"].astype(np.float64)
    mn, mx = synth.depth_range(f["depth"])
    or16 = o16.render(sc.W, sc.H, T[:3, 3], T, sc.intrinsics, mn, mx, pixels=pix)
    gr16, gr32 = g16.raycast(f, sc.intrinsics), g32.raycast(f, sc.intrinsics)
    for key in ("depth", "pos", "col"):
        assert _bit_equal(gr16[key][pix], or16[key][pix]), key

    # against FP64
    d64 = o64.render(sc.W, sc.H, T[:3, 3], T, sc.intrinsics, mn, mx, pixels=pix)["depth"][pix].astype(np.float64)
    b64 = o64.march_brackets(sc.W, T[:3, 3], T, sc.intrinsics, mn, mx, pix)
    b16 = o16.march_brackets(sc.W, T[:3, 3], T, sc.intrinsics, mn, mx, pix)
    # same accepted sample pair as the FP64 march (the FP32 store's pairs differ from the
    # FP16 store's only where their depths differ, 0.01 % of the pixels)
    same = np.abs(b16[:, 0] - b64[:, 0]) < wl.voxel / 4
    s16 = _depth_stats(gr16["depth"][pix], d64, same)
    s32 = _depth_stats(gr32["depth"][pix], d64, same)
    e16, n16, n64, ncommon = _common_sdf_error(g16, o64)
    e32, _, _, _ = _common_sdf_error(g32, o64)
    sdf = {k: {"rms_mm": float(np.sqrt(np.mean(e ** 2)) * 1e3), "p99_mm": float(np.quantile(e, 0.99) * 1e3),
               "p999_mm": float(np.quantile(e, 0.999) * 1e3), "max_mm": float(e.max() * 1e3)}
           for k, e in (("fp16", e16), ("fp32", e32))}
    d16, d32 = gr16["depth"][pix], gr32["depth"][pix]
    both = (d16 < 1e9) & (d32 < 1e9)
    f16_vs_32 = float(np.mean(np.abs(d16 - d32)[both] <= 5e-4))
    print(f"blocks fp16 {n16} fp64 {n64} common {ncommon}; sdf vs fp64 {sdf}")
    print(f"depth vs fp64: fp16 {s16}; fp32 {s32}; fp16 vs fp32 within 0.5 mm {f16_vs_32:.5f}")
    assert abs(n16 - n64) <= 0.001 * n64 and ncommon >= 0.999 * n64
    for s in (s16, s32):
        assert s["valid_both"] >= 0.998 * s["valid_either"]
        assert s["within_0.5mm"] >= 0.985
        assert s["same_bracket"] >= 0.98 * s["valid_both"]
        assert s["same_within_0.5mm"] >= 0.997 and s["same_max_mm"] < 5.0
    assert s16["within_0.5mm"] >= s32["within_0.5mm"] - 0.002
    for k in ("fp16", "fp32"):
        assert sdf[k]["rms_mm"] <= 0.05 and sdf[k]["p99_mm"] <= 0.05 and sdf[k]["p999_mm"] <= 0.2
    assert f16_vs_32 >= 0.999 and np.count_nonzero((d16 < 1e9) != (d32 < 1e9)) <= 0.001 * pix.size


def _room_fused(torch, n=64):
    """The differentiation configuration's model: the even frames 0 .. 2n-2 fused at their poses."""
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    g, o = GpuRun(torch, wl), OracleRun(wl)
    for i in range(0, 2 * n, 2):
        f = sc.render(i)
        g.integrate(f, sc.intrinsics)
        o.integrate(f, sc.intrinsics)
    return wl, sc, g, o


def _render_depth(g, sc, T, mn, mx):
    fx, fy, cx, cy = sc.intrinsics
    T = np.asarray(T, np.float64)
    if g.render is None:
        g.render = vsg.RayCastRender()
        assert g.render.initialize(sc.W, sc.H)
    g.render.render(g.scene, T[:3, 3].astype(np.float32), T.astype(np.float32), fx, fy, cx, cy, mn, mx)
    return g.render.downloadDepthFloat()


def test_room_gradients_against_oracle_and_central_differences(torch_cuda):
    """Held-out frame 63 of the room against the model fused from the 64 even frames.

    Jacobian d depth / d xi (pose * exp(xi^)): equal to the oracle's FP64 adjoint of the
    same float forward on >= 99 % of the hit pixels (relative 2e-3).  Central
    differences: the sampled field jumps at half-voxel planes, so a difference over
    [-h, h] that contains a jump measures the jump / 2h, not the slope; differences at
    h and h/2 agree only where no jump lies within +-h (their gap grows as 1/h
    otherwise).  On the (pixel, component) pairs where they agree, the Jacobian equals
    them; those pairs are the majority.  SDF gradient (vector-Jacobian product with
    seeded weights): equal to the oracle's on >= 99 % of the touched voxels."""
    torch = torch_cuda
    wl, sc, g, o = _room_fused(torch)
    f = sc.render(63)
    T = f["pose"].astype(np.float64)
    mn, mx = synth.depth_range(f["depth"])
    d0 = _render_depth(g, sc, T, mn, mx)
    n = sc.W * sc.H
    jac = torch.zeros(6 * n, dtype=torch.float32, device="cuda")
    g.render.adjoint(g.scene, jac=jac)
    torch.cuda.synchronize()
    J = jac.view(n, 6).cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(6)
    pix = rng.choice(n, 4096, replace=False).astype(np.int32)
    Jo, hit, _ = o.o.render_adjoint(sc.W, sc.H, T[:3, 3], T, sc.intrinsics, mn, mx, pix)
    assert np.array_equal(hit, d0[pix] < 1e9) and hit.sum() > 3500
    err = np.abs(J[pix][hit] - Jo[hit]).max(axis=1)
    good = err <= 2e-3 * np.abs(Jo[hit]).max(axis=1) + 1e-5
    print(f"room jacobian vs oracle FP64 adjoint: {good.mean():.4f} of {hit.sum()} hit pixels")
    assert good.mean() >= 0.99

    hp = pix[hit]
    h = 1e-4
    fd = {}
    for step in (h, h / 2):
        out = np.zeros((hp.size, 6))
        for k in range(6):
            e = np.zeros(6)
            e[k] = step
            dp = _render_depth(g, sc, T @ O.se3_exp(e), mn, mx)[hp].astype(np.float64)
            dm = _render_depth(g, sc, T @ O.se3_exp(-e), mn, mx)[hp].astype(np.float64)
            out[:, k] = (dp - dm) / (2 * step)
        fd[step] = out
    Jh = J[hp]
    tol = 0.02 * np.abs(Jh) + 2e-3 + 3e-7 / h
    smooth = np.abs(fd[h] - fd[h / 2]) <= tol
    agree = np.abs(Jh - fd[h]) <= tol
    print(f"room central differences h={h:g}: smooth pairs {smooth.mean():.4f}, agreement on them "
          f"{agree[smooth].mean():.4f}, on all pairs {agree.mean():.4f}")
    assert smooth.mean() >= 0.5 and agree[smooth].mean() >= 0.98

    w = rng.uniform(-1, 1, pix.size)
    gd = np.zeros(n, np.float32)
    gd[pix] = w
    nb = g.scene.counters()["num_allocated"]
    gsdf = torch.zeros(wl.blocks * 512, dtype=torch.float32, device="cuda")
    _render_depth(g, sc, T, mn, mx)
    g.render.adjoint(g.scene, grad_depth=torch.from_numpy(gd).cuda(), grad_sdf=gsdf)
    torch.cuda.synchronize()
    got = gsdf.cpu().numpy()[:nb * 512].astype(np.float64)
    _, _, ref = o.o.render_adjoint(sc.W, sc.H, T[:3, 3], T, sc.intrinsics, mn, mx, pix, grad_depth=w)
    nz = np.abs(ref) > 1e-6
    rel = np.abs(got[nz] - ref[nz]) / (np.abs(ref[nz]) + 1e-3)
    print(f"room sdf gradient: {nz.sum()} voxels, {np.mean(rel < 1e-2):.4f} within 1 %")
    assert nz.sum() > 10000 and np.mean(rel < 1e-2) >= 0.99 and np.abs(got[~nz]).max() < 1e-3
