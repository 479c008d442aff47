"""The 1024x768 room configuration (8x8x3 m, 60 primitives, voxel 2 mm, truncation
1 cm, 2^22 buckets and 2^22 blocks, FP16 store) with its tracking loop: before each
frame is integrated, its pose is refined on the device from a prior perturbed by up
to 1 deg / 1 cm (FP64 Gauss-Newton against the model fused so far, standard
trilinear sampling).  The fused state after 12 frames equals the oracle's given the
same poses and the reference-sampling raycast equals it on 4096 seeded pixels.  The
pose errors before and after refinement are printed, not asserted: against a model
fused from a few frames with this sensor noise (sigma 1-6 cm at 2-6 m) and 2 mm voxels
the depth-only alignment's optimum lies away from the true pose (its cost decreases
while the error grows), so this loop shows the data flow, not tracking accuracy.
test_c4_tracking_against_a_tracking_model tracks the same frames against a second scene
fused at the refined poses with 5 mm voxels and 8 cm truncation (wider than the noise)
and bounds the pose error."""
import numpy as np
import pytest

import oracle as O
from fusion_util import GpuRun, OracleRun, compare_raycast, compare_state
from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu


def _pose_err(Tgt, T):
    E = np.linalg.inv(Tgt) @ T
    return (float(np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1)))),
            float(np.linalg.norm(E[:3, 3])))


def test_c4_tracking_loop_and_parity(torch_cuda):
    torch = torch_cuda
    wl = synth.WORKLOADS[4]
    sc = synth.SynthScene(4)
    fx, fy, cx, cy = sc.intrinsics
    g, o = GpuRun(torch, wl), OracleRun(wl)
    tracker = vsg.RayCastRender()
    assert tracker.initialize(sc.W, sc.H)
    tracker.set_sampling("trilinear")
    rng = np.random.default_rng(5)
    errs, prior = [], []
    for i in range(12):
        f = sc.render(i)
        Tgt = f["pose"]
        T = Tgt
        if i > 0:
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            tdir = rng.normal(size=3)
            tdir /= np.linalg.norm(tdir)
            xi = np.concatenate([axis * np.radians(rng.uniform(0, 1.0)), tdir * rng.uniform(0, 0.01)])
            T = Tgt @ O.se3_exp(xi)
            obs = torch.from_numpy(f["depth"].reshape(-1)).cuda()
            mn, mx = synth.depth_range(f["depth"])
            prior.append(_pose_err(Tgt, T))
            for it in range(30):
                res = 0.05 if it < 10 else 0.02
                T, _ = tracker.gn_step(g.scene, obs, fx, fy, cx, cy, mn, mx, T, max_residual=res, lam=1e-3)
            errs.append(_pose_err(Tgt, T))
        ft = dict(f)
        ft["pose"] = T.astype(np.float32).astype(np.float64)  # both sides integrate the same float pose
        g.integrate(ft, sc.intrinsics)
        o.integrate(ft, sc.intrinsics)
    info = compare_state(g, o)
    pix = np.random.default_rng(6).choice(sc.W * sc.H, 4096, replace=False).astype(np.int32)
    gr, orr = g.raycast(ft, sc.intrinsics), o.raycast(ft, sc.intrinsics, pixels=pix)
    rc = compare_raycast(gr, orr, pixels=pix)
    errs, prior = np.array(errs), np.array(prior)
    print(f"c4: {info}; raycast {rc}")
    print(f"c4 prior pose error deg {np.round(prior[:, 0], 3)} mm {np.round(prior[:, 1] * 1e3, 2)}")
    print(f"c4 refined pose error deg {np.round(errs[:, 0], 4)} mm {np.round(errs[:, 1] * 1e3, 3)}")
    assert rc["exact"] == rc["valid"]


def test_c4_tracking_against_a_tracking_model(torch_cuda):
    torch = torch_cuda
    wl = synth.WORKLOADS[4]
    sc = synth.SynthScene(4)
    fx, fy, cx, cy = sc.intrinsics
    g = GpuRun(torch, wl)
    tp = vsg.Params(1 << 20, 1 << 20, 0.005, 0.08)
    ts = vsg.VoxelScene(tp, "fp16")
    ts.set_stream(torch.cuda.current_stream().cuda_stream)
    tracker = vsg.RayCastRender()
    assert tracker.initialize(sc.W, sc.H)
    tracker.set_sampling("trilinear")
    tracker.set_gn_noise(0.0012, 0.0019, 0.4)
    rng = np.random.default_rng(5)
    errs, prior = [], []
    for i in range(30):
        f = sc.render(i)
        Tgt = f["pose"].astype(np.float64)
        T = Tgt
        if i > 0:
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            tdir = rng.normal(size=3)
            tdir /= np.linalg.norm(tdir)
            T = Tgt @ O.se3_exp(np.concatenate([axis * np.radians(rng.uniform(0, 1.0)), tdir * rng.uniform(0, 0.01)]))
            obs = torch.from_numpy(f["depth"].reshape(-1)).cuda()
            mn, mx = synth.depth_range(f["depth"])
            prior.append(_pose_err(Tgt, T))
            for it in range(10):
                res = 0.2 if it < 3 else (0.08 if it < 6 else 0.03)
                T, _ = tracker.gn_step(ts, obs, fx, fy, cx, cy, mn, mx, T, max_residual=res, lam=1e-3)
            errs.append(_pose_err(Tgt, T))
        ft = dict(f)
        ft["pose"] = T.astype(np.float32)
        g.integrate(ft, sc.intrinsics)
        p, c, n = g.upload(ft)
        ts.integrateFromScanData(p, c, n, sc.W, sc.H, tp.truncation, ft["pose"], fx, fy, cx, cy)
        torch.cuda.synchronize()
    errs, prior = np.array(errs), np.array(prior)
    print(f"c4 tracking model: prior median {np.median(prior[:, 0]):.3f} deg {np.median(prior[:, 1]) * 1e3:.2f} mm; "
          f"refined median {np.median(errs[:, 0]):.4f} deg {np.median(errs[:, 1]) * 1e3:.3f} mm, "
          f"max {errs[:, 0].max():.4f} deg {errs[:, 1].max() * 1e3:.3f} mm")
    assert np.median(errs[:, 0]) <= 0.05 and np.median(errs[:, 1]) <= 0.002
    assert errs[:, 0].max() <= 0.1 and errs[:, 1].max() <= 0.004
    assert np.median(errs[:, 0]) < np.median(prior[:, 0]) / 5


def _fused_c4(torch, frames, voxel=None, trunc=None):
    wl = synth.WORKLOADS[4]
    sc = synth.SynthScene(4)
    fx, fy, cx, cy = sc.intrinsics
    p = vsg.Params(1 << 20, 1 << 20, voxel, trunc) if voxel else vsg.Params(wl.buckets, wl.blocks, wl.voxel, wl.trunc)
    scene = vsg.VoxelScene(p, "fp16")
    scene.set_stream(torch.cuda.current_stream().cuda_stream)
    for i in frames:
        f = sc.render(i)
        pts, col, nrm = (torch.from_numpy(f[k]).cuda() for k in ("points", "colors", "normals"))
        scene.integrateFromScanData(pts, col, nrm, sc.W, sc.H, p.truncation, f["pose"].astype(np.float32), fx, fy, cx,
                                    cy)
    torch.cuda.synchronize()
    return sc, scene


@pytest.mark.parametrize("sampling", ["trilinear", "reference"])
def test_pyramid_level_renders_are_the_full_image_subsampled(torch_cuda, sampling):
    """A render at 1/d resolution with intrinsics / d (what PoseTracker's levels draw)
    gives, bit for bit, the full render's every d-th row and column: (x' - cx/d)/(fx/d)
    rounds as (d x' - cx)/fx for d a power of two."""
    torch = torch_cuda
    sc, scene = _fused_c4(torch, range(0, 12, 2))
    fx, fy, cx, cy = sc.intrinsics
    f = sc.render(13)
    T = f["pose"].astype(np.float32)
    mn, mx = synth.depth_range(f["depth"])
    out = {}
    for d in (1, 2, 4):
        r = vsg.RayCastRender()
        assert r.initialize(sc.W // d, sc.H // d)
        r.set_sampling(sampling)
        r.render(scene, T[:3, 3], T, fx / d, fy / d, cx / d, cy / d, mn, mx)
        out[d] = r.downloadAll()
        r.close()
    full = out[1]
    assert int((full["depth"] < 1e9).sum()) > sc.W * sc.H // 3
    for d in (2, 4):
        for k in ("depth", "pos"):
            a = np.ascontiguousarray(full[k].reshape(sc.H, sc.W, -1)[::d, ::d], np.float32).view(np.uint32)
            b = np.ascontiguousarray(out[d][k].reshape(sc.H // d, sc.W // d, -1), np.float32).view(np.uint32)
            assert np.array_equal(a, b), (d, k, int((a != b).sum()))
    scene.close()


def test_gn_window_wider_than_the_depth_range_changes_nothing(torch_cuda):
    """With the projective window on and residual windows wider than the depth range,
    every ray marches [minDepth, maxDepth] as without it, so Gauss-Newton steps and
    refinements end at the same poses bit for bit (observed depths made positive
    everywhere, since the window leaves pixels without one unmarched)."""
    torch = torch_cuda
    sc, ts = _fused_c4(torch, range(0, 20, 2), voxel=0.005, trunc=0.08)
    fx, fy, cx, cy = sc.intrinsics
    f = sc.render(21)
    d = f["depth"].copy()
    d[~(d > 0)] = 2.5
    obs = torch.from_numpy(d.reshape(-1)).cuda()
    mn, mx = synth.depth_range(f["depth"])
    T0 = f["pose"].astype(np.float64) @ O.se3_exp(np.array([0.01, -0.005, 0.004, 0.006, 0.004, -0.003]))
    out = []
    for win in (False, True):
        r = vsg.RayCastRender()
        assert r.initialize(sc.W, sc.H)
        r.set_sampling("trilinear")
        r.set_gn_window(win)
        T1, s1 = r.gn_step(ts, obs, fx, fy, cx, cy, mn, mx, T0, max_residual=1e4, lam=1e-3)
        T2, s2 = r.gn_refine(ts, obs, fx, fy, cx, cy, mn, mx, T0, [1e4] * 4, lam=1e-3)
        out.append((T1, s1, T2, s2, r.downloadAll()["depth"]))
        r.close()
    (a1, as1, a2, as2, ad), (b1, bs1, b2, bs2, bd) = out
    assert np.array_equal(a1.view(np.uint64), b1.view(np.uint64)) and as1 == bs1
    assert np.array_equal(a2.view(np.uint64), b2.view(np.uint64)) and as2 == bs2
    assert np.array_equal(ad.view(np.uint32), bd.view(np.uint32))
    ts.close()


@pytest.mark.parametrize("window", [True, False])
def test_pose_tracker_pyramid_matches_full_resolution_refinement(torch_cuda, window):
    """PoseTracker's default pyramid (3 iterations at 1/2, 2 at full) ends within
    0.01 deg / 0.3 mm of ten full-resolution iterations with residual windows
    (and the same projective-window setting), frame by frame, against the 5 mm / 8 cm
    tracking model fused at the true poses; both
    recover the 0.5-1 deg / 5-10 mm priors to 0.05 deg / 1.5 mm, and with the window (the
    tracker's default) the pyramid's median error is within 0.05 mm of the
    full-resolution one (without it, five iterations stop short of ten: printed)."""
    torch = torch_cuda
    sc, ts = _fused_c4(torch, range(0, 40, 2), voxel=0.005, trunc=0.08)
    fx, fy, cx, cy = sc.intrinsics
    tracker = vsg.PoseTracker(sc.W, sc.H, window=window)
    full = vsg.RayCastRender()
    assert full.initialize(sc.W, sc.H)
    full.set_sampling("trilinear")
    full.set_gn_noise(0.0012, 0.0019, 0.4)
    full.set_gn_window(window)
    windows = [0.2] * 3 + [0.08] * 3 + [0.03] * 4
    rng = np.random.default_rng(11)
    rows = []
    for i in range(41, 56, 2):
        f = sc.render(i)
        Tgt = f["pose"].astype(np.float64)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        tdir = rng.normal(size=3)
        tdir /= np.linalg.norm(tdir)
        T0 = Tgt @ O.se3_exp(np.concatenate([axis * np.radians(rng.uniform(0.5, 1.0)), tdir * rng.uniform(0.005, 0.01)]))
        obs = torch.from_numpy(f["depth"]).cuda()
        mn, mx = synth.depth_range(f["depth"])
        Tp, _ = tracker.refine(ts, obs, fx, fy, cx, cy, mn, mx, T0)
        Tf, _ = full.gn_refine(ts, obs.reshape(-1), fx, fy, cx, cy, mn, mx, T0, windows, lam=1e-3)
        rows.append((*_pose_err(Tgt, Tp), *_pose_err(Tgt, Tf), *_pose_err(Tf, Tp)))
    R = np.array(rows)
    print(f"window {window}: pyramid vs full: max pose difference {R[:, 4].max():.5f} deg {R[:, 5].max() * 1e3:.4f} mm; "
          f"pyramid error max {R[:, 0].max():.4f} deg {R[:, 1].max() * 1e3:.3f} mm, median "
          f"{np.median(R[:, 1]) * 1e3:.3f} mm (full resolution {np.median(R[:, 3]) * 1e3:.3f} mm)")
    assert R[:, 4].max() <= 0.01 and R[:, 5].max() <= 3e-4
    assert R[:, 0].max() <= 0.05 and R[:, 1].max() <= 1.5e-3
    assert R[:, 2].max() <= 0.05 and R[:, 3].max() <= 1.5e-3
    if window:
        assert np.median(R[:, 1]) <= np.median(R[:, 3]) + 5e-5
    tracker.close()
    full.close()
    ts.close()
