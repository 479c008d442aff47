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
