#!/usr/bin/env python3
"""Fused integrate + raycast throughput of the VoxelSG path on one MI355X.

Workload: the 640x480 room sequence (300 float32 depth frames, f=525, 4x4x3 m room
with 20 primitives, random-walk camera, voxel 4 mm, truncation 0.02 m, 2^20
buckets, 2^19 blocks).  One step = one pass over the whole sequence from an empty
scene: per frame VoxelScene::integrateFromScanData, then RayCastRender::render at
the frame's pose over the driver's depth range.  All frames are resident in HBM
before the timed region starts; nothing is copied inside it.

Prints one JSON line: frames/s (value), the roofline of the dominant kernel (HIP
events on the scene's stream around every launch in the timed region, algorithmic
bytes from per-frame counters of an untimed instrumented pass), the fused-frame
roofline, and the oracle's CPU throughput on a bounded sample of the same frames.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(REPO, "real-time-depth-fusion-directional-tsdf_amd")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_pkg():
    spec = importlib.util.spec_from_file_location("vsg_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vsg_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def kernel_bytes(st: dict, npix: int, b_sw: int) -> dict:
    """Algorithmic HBM bytes of one frame, per kernel (compulsory traffic of the data flow)."""
    return {
        # pointmap read, claim flag written, buckets with live entries probed (10 x 16 B each)
        "alloc_claim": 12 * npix + 4 * npix + 160 * st["alloc_buckets"],
        # live-block list (16 B per block) read, visible list (4 B) written
        "compact": 16 * st["num_allocated"] + 4 * st["visible_blocks"],
        # pointmap + colour read once, list + block coordinate per block, voxel word and colour
        # read-modify-write per updated voxel
        "integrate": 15 * npix + 20 * st["visible_blocks"] + st["updated_voxels"] * (2 * b_sw + 8),
        # outputs (depth 4 + position 16 + colour 16 + normal 16), probed buckets, voxel words of
        # the blocks the march samples
        "raycast": 52 * npix + 160 * st["ray_buckets"] + 512 * b_sw * st["ray_blocks"],
    }


def frame_bytes(st: dict, npix: int, b_sw: int) -> int:
    """Bytes(frame) of the fused path: inputs, outputs, live entries, probed buckets, voxel
    read-modify-writes, sampled voxel blocks."""
    return (27 * npix + 52 * npix + 16 * st["num_allocated"] + 160 * st["any_buckets"]
            + st["updated_voxels"] * (2 * b_sw + 8) + 512 * b_sw * st["ray_blocks"])


def host_info() -> dict:
    """Cores this process may use, the CPU model and its clock (from /proc/cpuinfo)."""
    info = {"nproc": len(os.sched_getaffinity(0)), "cpu_model": None, "mhz": None}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "model name" and info["cpu_model"] is None:
                info["cpu_model"] = v.strip()
            elif k == "cpu MHz" and info["mhz"] is None:
                info["mhz"] = float(v)
    except OSError:
        pass
    return info


def cpu_run(synth, config: int, threads: int, frames: int, raycast_frames, progress: bool = False) -> dict:
    """The oracle (float, the reference's algorithm: allocation in canonical order,
    compaction by a full table scan, every allocated block integrated, 10-slot bucket
    probes on every lookup, a ray march every half voxel with no empty-space skipping) on
    frames 0..frames-1 of a configuration's sequence with `threads` OpenMP threads.
    Every frame is integrated, its allocation, compaction and integration timed apart;
    the frames in raycast_frames are then raycast over the whole image at their pose
    (depth range by the driver's rule, timed with the raycast).  Frame synthesis is
    outside the timed stages."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    used = O.set_threads(threads)
    wl = synth.WORKLOADS[config]
    sc = synth.SynthScene(config)
    o = O.OracleScene(wl.buckets, wl.blocks, wl.voxel)
    st = {"alloc": [], "compact": [], "integrate": [], "raycast": []}
    ctr = {"num_allocated": [], "visible": [], "updated": [], "hits": [], "misses": []}
    rset = set(int(i) for i in raycast_frames)
    for i in range(frames):
        f = sc.render(i)
        c = o.integrate(f["points"], f["colors"], wl.trunc, f["pose"], sc.intrinsics)
        a, k, g = o.stage_times()
        st["alloc"].append(a * 1e3)
        st["compact"].append(k * 1e3)
        st["integrate"].append(g * 1e3)
        for n in ("num_allocated", "visible", "updated"):
            ctr[n].append(c[n])
        if i in rset:
            T = f["pose"].astype(np.float64)
            t0 = time.perf_counter()
            mn, mx = synth.depth_range(f["depth"])
            r = o.render(sc.W, sc.H, T[:3, 3], T, sc.intrinsics, mn, mx)
            st["raycast"].append((time.perf_counter() - t0) * 1e3)
            ctr["hits"].append(r["hits"])
            ctr["misses"].append(r["misses"])
        if progress:
            print(f"[cpu] config {config} threads {used} frame {i}: alloc {a * 1e3:.2f} compact {k * 1e3:.2f} "
                  f"integrate {g * 1e3:.2f} ms" + (f" raycast {st['raycast'][-1]:.0f} ms" if i in rset else ""),
                  file=sys.stderr, flush=True)
    summary = {k: {"mean": round(float(np.mean(v)), 3), "p95": round(float(np.quantile(v, 0.95)), 3)}
               for k, v in st.items() if v}
    frame_ms = sum(summary[k]["mean"] for k in ("alloc", "compact", "integrate")) + summary["raycast"]["mean"]
    return {"config": config, "workload": wl.name, "threads": used, "frames": frames,
            "raycast_frames": sorted(rset), "frames_per_s": round(1e3 / frame_ms, 5), "ms": summary,
            "counters_mean": {k: int(np.mean(v)) for k, v in ctr.items() if v},
            "per_frame_ms": {k: [round(x, 3) for x in v] for k, v in st.items()}}


def cpu_baseline(synth) -> dict:
    """The oracle timed on the host's cores, bounded to about 10-20 s: the 320x240 sphere
    sequence (30 frames integrated, full-image raycasts after frames 9, 19 and 29) and
    the first 30 frames of the 640x480 room sequence (integrated, full-image raycasts
    after frames 14 and 29), all cores.  value = the room sample's frames/s (mean
    allocation + compaction + integration + raycast per frame).  The full table (both
    sequences at 1 thread and all cores, a raycast after every frame) comes from
    bench.py --cpu-plan runs."""
    host = host_info()
    n = host["nproc"]
    c0 = cpu_run(synth, 0, n, 30, (9, 19, 29))
    c1 = cpu_run(synth, 1, n, 30, (14, 29))
    for r in (c0, c1):
        r.pop("per_frame_ms")
    return {"value": c1["frames_per_s"], "unit": "frames/s", "cores": c1["threads"], "kind": "port",
            "sample": ("oracle (float, the reference's algorithm) on frames 0-29 of the 640x480 room sequence, "
                       "every frame integrated, full-image raycasts after frames 14 and 29; per-stage ms below; "
                       "also the 320x240 sphere sequence (30 frames, raycasts after 9, 19, 29)"),
            "host": host, "room_640x480": c1, "sphere_320x240": c0,
            "plan": ("profiles/cpu_baseline.md: both sequences at 1 thread and all cores, a raycast after every "
                     "frame (bench.py --cpu-plan)")}


def error_stats(err: np.ndarray) -> dict:
    err = np.abs(np.asarray(err, np.float64))
    if err.size == 0:
        return {"n": 0}
    return {"n": int(err.size), "max_mm": round(float(err.max()) * 1e3, 5),
            "rms_mm": round(float(np.sqrt(np.mean(err ** 2))) * 1e3, 5),
            "p50_mm": round(float(np.quantile(err, 0.5)) * 1e3, 5),
            "p99_mm": round(float(np.quantile(err, 0.99)) * 1e3, 5),
            "p999_mm": round(float(np.quantile(err, 0.999)) * 1e3, 5),
            "within_0.5mm": round(float(np.mean(err <= 5e-4)), 5)}


def precision_report(torch, vsg, sc, frames, wl, pts, rgb, nrm, nframes: int, npix_check: int) -> dict:
    """SDF and raycast-depth error of the FP32 and FP16 stores against the FP64 oracle on
    the first `nframes` frames: SDF on voxels updated in both (matched by block
    coordinate), depth of the last frame's raycast on `npix_check` seeded pixels valid
    in both."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    W, H = sc.W, sc.H
    fx, fy, cx, cy = sc.intrinsics
    o = O.OracleScene(wl.buckets, wl.blocks, wl.voxel, fp64=True)
    for i in range(nframes):
        o.integrate(frames[i]["points"], frames[i]["colors"], wl.trunc, frames[i]["pose"], sc.intrinsics)
    oc, _, _ = o.export_blocks()
    osdf, ow, _, _ = o.export_voxels(len(oc))
    rng = np.random.default_rng(11)
    pix = rng.choice(W * H, npix_check, replace=False).astype(np.int32)
    f = frames[nframes - 1]
    T = f["pose"].astype(np.float64)
    mn, mx = f["range"]
    od = o.render(W, H, T[:3, 3], T, sc.intrinsics, mn, mx, pixels=pix)["depth"][pix]
    okey = {tuple(c): k for k, c in enumerate(oc)}
    out = {"frames": nframes, "pixels": npix_check}
    gdepth = {}
    for store in ("fp32", "fp16"):
        scene = vsg.VoxelScene(vsg.Params(wl.buckets, wl.blocks, wl.voxel, wl.trunc), store)
        scene.set_stream(torch.cuda.current_stream().cuda_stream)
        for i in range(nframes):
            T_i = frames[i]["pose"]
            scene.integrateFromScanData(pts[i], rgb[i], nrm[i], W, H, wl.trunc, T_i, fx, fy, cx, cy)
        rr = vsg.RayCastRender()
        rr.initialize(W, H)
        rr.render(scene, T[:3, 3], T, fx, fy, cx, cy, mn, mx)
        gd = rr.downloadDepthFloat()[pix]
        gdepth[store] = gd
        gc, _ = scene.export_blocks()
        gsdf, gw, _ = scene.export_voxels(0, len(gc))
        gi, oi = [], []
        for k, c in enumerate(gc):
            j = okey.get(tuple(c))
            if j is not None:
                gi.append(k)
                oi.append(j)
        gi, oi = np.array(gi), np.array(oi)
        gv = (gi[:, None] * 512 + np.arange(512)[None, :]).reshape(-1)
        ov = (oi[:, None] * 512 + np.arange(512)[None, :]).reshape(-1)
        both = (gw[gv] > 0) & (ow[ov] > 0)
        valid = (gd < 1e9) & (od < 1e9)
        out[store] = {"sdf": error_stats(gsdf[gv][both].astype(np.float64) - osdf[ov][both]),
                      "depth": error_stats(gd[valid].astype(np.float64) - od[valid]),
                      "blocks_gpu": int(len(gc)), "blocks_fp64": int(len(oc)), "blocks_common": int(len(gi))}
        rr.close()
        scene.close()
    # what the FP16 store alone changes: the same march over FP16 vs FP32 voxels
    v = (gdepth["fp16"] < 1e9) & (gdepth["fp32"] < 1e9)
    out["fp16_vs_fp32_depth"] = error_stats(gdepth["fp16"][v].astype(np.float64) - gdepth["fp32"][v])
    return out


def depthgen_bench(torch, vsg, sc, frames, reps: int = 20) -> dict:
    """Depth maps from a point cloud (CustomDepthMapGenerator::generateFromPointCloud): the
    valid points of frames 10 and 13 of the sequence (the second moved into frame 10's
    camera), shuffled, 737k points -> 640x480 maps.  Algorithmic bytes: 27 B read per
    point (position, normal, colour), 31 B written per pixel (point, depth, normal,
    colour)."""
    rng = np.random.default_rng(0)
    parts = []
    T10 = frames[10]["pose"].astype(np.float64)
    for j in (10, 13):
        f = frames[j]
        ok = f["depth"].reshape(-1) > 0
        p = f["points"].reshape(-1, 3)[ok].astype(np.float64)
        M = np.linalg.inv(T10) @ f["pose"].astype(np.float64)
        parts.append(((p @ M[:3, :3].T + M[:3, 3]).astype(np.float32), f["normals"].reshape(-1, 3)[ok],
                      f["colors"].reshape(-1, 3)[ok]))
    pts, nrm, col = (np.concatenate([q[k] for q in parts]) for k in range(3))
    perm = rng.permutation(len(pts))
    dev = torch.device("cuda:0")
    P, N, Cc = (torch.from_numpy(np.ascontiguousarray(a[perm])).to(dev) for a in (pts, nrm, col))
    g = vsg.CustomDepthMapGenerator()
    fx, fy, cx, cy = sc.intrinsics
    g.setCameraParams(fx, fy, cx, cy, 0.1, 10.0, sc.W, sc.H)
    d = g.generate_device(P, N, Cc)
    torch.cuda.synchronize()
    out = {"pointmap": d.pointmap, "depthmap": d.depthmap, "normalmap": d.normalmap, "colormap": d.colormap}
    vsg.timing_enable(True)
    vsg.timing_reset()
    for _ in range(reps):
        g.generate_device(P, N, Cc, out=out)
    valid = g.valid_pixels()
    ms, n = vsg.timing_read()["depthgen"]
    vsg.timing_enable(False)
    g.close()
    npix = sc.W * sc.H
    b = len(pts) * 27 + npix * 31
    gbs = b / (ms / n * 1e-3) / 1e9
    log(f"depth maps from points: {len(pts)} points -> {valid} pixels in {ms / n:.4f} ms")
    return {"points": int(len(pts)), "valid_pixels": int(valid), "avg_ms": round(ms / n, 4),
            "algorithmic_bytes": int(b), "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 5)}


def shim_bench(nframes: int, passes: int) -> dict | None:
    """The C++ drop-in classes (include/vsg_voxelscene.hpp) on the same sequence and passes,
    timed by lib/vsg_shim_example's bench form (frames resident in device memory, pose as
    a device pointer)."""
    exe = os.path.join(PKG_DIR, "lib", "vsg_shim_example")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe, "bench", str(nframes), str(passes)], capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"vsg_shim_example bench failed ({out.returncode}): {out.stderr[-500:]}")
    tok = out.stdout.split()[1:]  # "bench" then name / value pairs
    got = {tok[i]: tok[i + 1] for i in range(0, len(tok) - 1, 2)}
    return {"value": float(got["fps"]), "ms_per_frame": float(got["ms_per_frame"]), "passes": passes}


def se3_exp(xi) -> np.ndarray:
    """exp of the twist xi = (omega, upsilon) as a 4x4 (Rodrigues; V = I + B [w] + C [w]^2)."""
    w, u = np.asarray(xi[:3], np.float64), np.asarray(xi[3:], np.float64)
    th = float(np.linalg.norm(w))
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        R, V = np.eye(3) + K, np.eye(3) + 0.5 * K
    else:
        A, B, C = np.sin(th) / th, (1 - np.cos(th)) / th ** 2, (th - np.sin(th)) / th ** 3
        R = np.eye(3) + A * K + B * K @ K
        V = np.eye(3) + B * K + C * K @ K
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ u
    return T


GN_SCHEDULE = [0.2] * 3 + [0.08] * 3 + [0.03] * 4  # residual window (m) of the 10 Gauss-Newton iterations


def pose_refinement_bench(torch, vsg, synth, sc, frames, wl) -> dict:
    """Differentiable pose refinement (the differentiation configuration): the even frames
    0..126 fused at their poses, the 64 odd frames 1..127 perturbed by seeded uniform
    rotations <= 2 deg and translations <= 2 cm, 10 Gauss-Newton iterations each on the
    device (vsg_pose_gn_refine; FP64 6x6 solve; residual window 20 -> 8 -> 3 cm; smoothed-gradient
    Jacobian, residuals weighted by the sensor's noise model, standard trilinear
    sampling)."""
    dev = torch.device("cuda:0")
    scene = vsg.VoxelScene(vsg.Params(wl.buckets, wl.blocks, wl.voxel, wl.trunc), "fp32")
    fx, fy, cx, cy = sc.intrinsics
    for i in range(0, 128, 2):
        f = frames[i]
        p, c, n = (torch.from_numpy(f[k]).to(dev) for k in ("points", "colors", "normals"))
        scene.integrateFromScanData(p, c, n, sc.W, sc.H, wl.trunc, f["pose"], fx, fy, cx, cy)
        torch.cuda.synchronize()
    rr = vsg.RayCastRender()
    assert rr.initialize(sc.W, sc.H)
    rr.set_sampling("trilinear")
    rr.set_gn_noise(0.0012, 0.0019, 0.4)  # the sequence's depth noise model
    rng = np.random.default_rng(7)
    errs, t_gn = [], 0.0
    for i in range(1, 128, 2):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        tdir = rng.normal(size=3)
        tdir /= np.linalg.norm(tdir)
        xi = np.concatenate([axis * np.radians(rng.uniform(0, 2.0)), tdir * rng.uniform(0, 0.02)])
        f = frames[i]
        Tgt = f["pose"].astype(np.float64)
        T = Tgt @ se3_exp(xi)
        obs = torch.from_numpy(f["depth"].reshape(-1)).to(dev)
        mn, mx = synth.depth_range(f["depth"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        T, _ = rr.gn_refine(scene, obs, fx, fy, cx, cy, mn, mx, T, GN_SCHEDULE, lam=1e-3)
        t_gn += time.perf_counter() - t0
        errs.append(_pose_err(Tgt, T))
    scene.close()
    rr.close()
    E = np.array(errs)
    ok = (E[:, 0] <= 0.1) & (E[:, 1] <= 1.0)
    log(f"pose refinement: {int(ok.sum())}/64 within 0.1 deg / 1 mm")
    return {"frames": 64, "iterations": 10, "within_0.1deg_1mm": int(ok.sum()),
            "median_deg": round(float(np.median(E[:, 0])), 4), "median_mm": round(float(np.median(E[:, 1])), 3),
            "max_deg": round(float(E[:, 0].max()), 3), "max_mm": round(float(E[:, 1].max()), 2),
            "ms_per_frame": round(t_gn / 64 * 1e3, 2), "ms_per_iteration": round(t_gn / 640 * 1e3, 3)}


def _stage(ms) -> dict:
    a = np.asarray(ms, np.float64)
    return {"mean": round(float(a.mean()), 3), "p95": round(float(np.quantile(a, 0.95)), 3)} if a.size else {}


def _pose_summary(R: np.ndarray) -> dict:
    ok = (R[:, 0] <= 0.1) & (R[:, 1] <= 1.0)
    return {"median": {"deg": round(float(np.median(R[:, 0])), 4), "mm": round(float(np.median(R[:, 1])), 3)},
            "p95": {"deg": round(float(np.quantile(R[:, 0], 0.95)), 4), "mm": round(float(np.quantile(R[:, 1], 0.95)), 3)},
            "max": {"deg": round(float(R[:, 0].max()), 3), "mm": round(float(R[:, 1].max()), 2)},
            "final": {"deg": round(float(R[-1, 0]), 4), "mm": round(float(R[-1, 1]), 3)},
            "within_0.1deg_1mm": int(ok.sum()), "frames": int(len(R))}


def large_run_bench(torch, vsg, synth, nframes: int, stream_radius: float = 0.0, trace_path: str = "") -> dict:
    """The large combined configuration: 1024x768 room (8x8x3 m, 60 primitives), voxel 2 mm,
    truncation 1 cm, 2^22 buckets and blocks, FP16 store.  Every frame after the first is
    tracked from a prior perturbed by a seeded uniform <= 1 deg / 1 cm: Gauss-Newton on
    the device over an image pyramid (vsg.PoseTracker: 3 iterations at 1/2 resolution,
    2 at full; projective window, smoothed Jacobian, noise-weighted residuals, standard
    trilinear sampling, one pose read-back per level), then integrated at the refined
    pose and raycast there (the reference's sampling).

    The pose that is integrated comes from tracking against a companion model of the same
    frames (voxel 5 mm, truncation 8 cm: wider than the sensor noise, 1-2.6 cm at 2-4 m),
    integrated at the same refined poses.  The same prior is also refined against the
    2 mm / 1 cm model itself, whose result is reported beside it and not used.  With
    stream_radius > 0, every 25 frames the 2 mm scene's blocks beyond that distance from
    the camera go to a VoxelStreamingManager chunk grid (1 m chunks) and the chunks
    within it come back.  Per-frame times of the stages (a synchronisation after each)
    and the pose errors; with trace_path, one line per frame goes there."""
    wl = synth.WORKLOADS[4]
    sc = synth.SynthScene(4, nframes=nframes)
    fx, fy, cx, cy = sc.intrinsics
    dev = torch.device("cuda:0")
    scene = vsg.VoxelScene(vsg.Params(wl.buckets, wl.blocks, wl.voxel, wl.trunc), "fp16")
    stream = torch.cuda.current_stream().cuda_stream
    scene.set_stream(stream)
    mgr = None
    t_stream, max_blocks, n_stream = [], 0, 0
    if stream_radius > 0:
        mgr = vsg.VoxelStreamingManager(scene, scene.m_params, (1.0, 1.0, 1.0), (64, 64, 64), (-32, -32, -32))
    track_voxel, track_trunc = 0.005, 0.08
    tscene = vsg.VoxelScene(vsg.Params(1 << 20, 1 << 20, track_voxel, track_trunc), "fp16")
    tscene.set_stream(stream)
    trackers = [vsg.PoseTracker(sc.W, sc.H) for _ in range(2)]
    rr = vsg.RayCastRender()
    assert rr.initialize(sc.W, sc.H)
    rng = np.random.default_rng(5)
    ms = {k: [] for k in ("track", "track_2mm_model", "integrate", "integrate_tracking_model", "raycast")}
    prior, refined, refined2 = [], [], []
    trace = open(trace_path, "w") if trace_path else None
    for i in range(nframes):
        f = sc.render(i)
        Tgt = f["pose"].astype(np.float64)
        p, c, n = (torch.from_numpy(f[k]).to(dev) for k in ("points", "colors", "normals"))
        obs = torch.from_numpy(f["depth"]).to(dev)
        mn, mx = synth.depth_range(f["depth"])
        torch.cuda.synchronize()
        T = Tgt
        if i > 0:
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            tdir = rng.normal(size=3)
            tdir /= np.linalg.norm(tdir)
            T0 = Tgt @ se3_exp(np.concatenate([axis * np.radians(rng.uniform(0, 1.0)), tdir * rng.uniform(0, 0.01)]))
            prior.append(_pose_err(Tgt, T0))
            t0 = time.perf_counter()
            T, st = trackers[0].refine(tscene, obs, fx, fy, cx, cy, mn, mx, T0, lam=1e-3)
            t1 = time.perf_counter()
            T2, st2 = trackers[1].refine(scene, obs, fx, fy, cx, cy, mn, mx, T0, lam=1e-3)
            t2 = time.perf_counter()
            ms["track"].append((t1 - t0) * 1e3)
            ms["track_2mm_model"].append((t2 - t1) * 1e3)
            refined.append(_pose_err(Tgt, T))
            refined2.append(_pose_err(Tgt, T2))
            if trace:
                trace.write(f"{i} prior {prior[-1][0]:.4f} {prior[-1][1]:.3f} track {refined[-1][0]:.4f} "
                            f"{refined[-1][1]:.3f} n {st['count']} track2mm {refined2[-1][0]:.4f} "
                            f"{refined2[-1][1]:.3f} n {st2['count']}\n")
        Tf = T.astype(np.float32)
        t0 = time.perf_counter()
        scene.integrateFromScanData(p, c, n, sc.W, sc.H, wl.trunc, Tf, fx, fy, cx, cy)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        tscene.integrateFromScanData(p, c, n, sc.W, sc.H, track_trunc, Tf, fx, fy, cx, cy)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        rr.render(scene, Tf[:3, 3], Tf, fx, fy, cx, cy, mn, mx)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        ms["integrate"].append((t1 - t0) * 1e3)
        ms["integrate_tracking_model"].append((t2 - t1) * 1e3)
        ms["raycast"].append((t3 - t2) * 1e3)
        if mgr is not None and i % 25 == 24:
            max_blocks = max(max_blocks, scene.counters()["num_allocated"])
            ts0 = time.perf_counter()
            mgr.streamOut(scene, Tf[:3, 3], stream_radius)
            mgr.streamIn(scene, Tf[:3, 3], stream_radius)
            t_stream.append((time.perf_counter() - ts0) * 1e3)
        if i % 100 == 99:
            log(f"large run: frame {i + 1}/{nframes}, track {np.mean(ms['track']):.2f} ms, "
                f"refined median {np.median(np.array(refined)[:, 1]):.3f} mm")
    if trace:
        trace.close()
    blocks = scene.counters()["num_allocated"]
    tblocks = tscene.counters()["num_allocated"]
    streaming = None
    if mgr is not None:
        keys = [c.live_keys() for c in mgr.m_grid.values()]
        images = int(mgr.numStreamedBlocks())
        distinct = int(len(np.unique(np.concatenate(keys)))) if keys else 0
        streaming = {"radius": stream_radius, "every_frames": 25, "max_blocks_before": int(max_blocks),
                     "images_in_chunk_grid": images, "distinct_blocks_in_chunk_grid": distinct,
                     "dropped_outside_grid": int(mgr.droppedOutsideGrid),
                     "round_trip_ms": _stage(t_stream),
                     "blocks_out_per_round_trip": int(mgr.nStreamedOut / max(len(t_stream), 1)),
                     "blocks_in_per_round_trip": int(mgr.nStreamedIn / max(len(t_stream), 1)),
                     "image_bytes": int(scene.block_image_bytes()),
                     "ms_per_round_trip_by_part": {k: round(v / max(len(t_stream), 1) * 1e3, 1)
                                                   for k, v in mgr.seconds.items()}}
    scene.close()
    tscene.close()
    for t in trackers:
        t.close()
    rr.close()
    P, R, R2 = np.array(prior), np.array(refined), np.array(refined2)
    per_frame = {k: _stage(v) for k, v in ms.items()}
    used = np.array(ms["track"][:]) + np.array(ms["integrate"][1:]) + np.array(ms["integrate_tracking_model"][1:]) \
        + np.array(ms["raycast"][1:])
    log(f"large run: {nframes} frames, tracking-model pose median {np.median(R[:, 0]):.3f} deg "
        f"{np.median(R[:, 1]):.2f} mm, 2 mm model {np.median(R2[:, 0]):.3f} deg {np.median(R2[:, 1]):.2f} mm")
    return {"workload": "room 1024x768, voxel 2 mm, trunc 0.01 m, 2^22 buckets, FP16 store", "frames": nframes,
            "blocks": int(blocks),
            "tracking_model": {"voxel": track_voxel, "trunc": track_trunc, "blocks": int(tblocks)},
            **({"streaming": streaming} if streaming else {}),
            "ms_per_frame": {**per_frame, "track_integrate_raycast": _stage(used)},
            "gn_iterations": {f"1/{d}" if d > 1 else "full": len(w) for d, w in vsg.PoseTracker.LEVELS},
            "prior_error": _pose_summary(P),
            "pose_error_tracking_model": _pose_summary(R),
            "pose_error_2mm_model": _pose_summary(R2),
            "note": "wall clock per stage with a synchronisation after each; track_integrate_raycast = track + "
                    "integrate of both models + raycast (the 2 mm model's tracking is reported, not used)"}


def _pose_err(Tgt, T):
    E = np.linalg.inv(Tgt) @ T
    return (float(np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1)))),
            float(np.linalg.norm(E[:3, 3]) * 1e3))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed passes over the sequence")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before timing")
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--store", choices=["fp32", "fp16"], default="fp32")
    ap.add_argument("--compare-fp16", action="store_true", default=True,
                    help="also time the FP16 store (reported beside the FP32 headline)")
    ap.add_argument("--no-compare-fp16", dest="compare_fp16", action="store_false")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-plan", default="",
                    help="CONFIG:THREADS:FIRST-LAST (THREADS 0 = all cores): only run the oracle on frames "
                         "0..LAST of configuration CONFIG (0: 320x240 sphere, 1: 640x480 room), raycasting after "
                         "each of frames FIRST..LAST, and print its per-stage times as one JSON line (no GPU)")
    ap.add_argument("--precision-frames", type=int, default=30)
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-extract", action="store_true",
                    help="skip timing the surface-point extraction of the fused scene")
    ap.add_argument("--no-mesh", action="store_true",
                    help="skip timing the marching cubes of the fused scene")
    ap.add_argument("--large-stream-radius", type=float, default=0.0,
                    help="stream the large run's blocks beyond this distance (m) out every 25 frames (0: off)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip timing a stream-out / stream-in round trip of the fused scene")
    ap.add_argument("--no-shim", action="store_true",
                    help="skip timing the C++ drop-in classes on the same sequence (lib/vsg_shim_example bench)")
    ap.add_argument("--no-depthgen", action="store_true",
                    help="skip timing the depth maps from a point cloud")
    ap.add_argument("--no-pose", action="store_true",
                    help="skip the pose-refinement run (64 held-out frames, 10 iterations each)")
    ap.add_argument("--large-frames", type=int, default=1000,
                    help="frames of the 1024x768 tracked run (0: skip; the configuration has 1000)")
    ap.add_argument("--large-trace", default="",
                    help="write the large run's per-frame pose errors to this file")
    ap.add_argument("--step-budget", type=int, default=None,
                    help="raycast steps per lane before the cooperative tail pass (library default if unset)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first run a plain copy kernel (torch.neg, float32) over this many MiB, 4 times: "
                         "its FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc scale the counters of the "
                         "fusion kernels to bytes")
    args = ap.parse_args()
    if args.cpu_plan:
        load_pkg()
        from vsg_amd import synth
        cfg, thr, rng = args.cpu_plan.split(":")
        first, last = (int(v) for v in rng.split("-"))
        thr = int(thr) or host_info()["nproc"]
        res = cpu_run(synth, int(cfg), thr, last + 1, range(first, last + 1), progress=True)
        print(json.dumps({"cpu_plan": args.cpu_plan, "host": host_info(), **res}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path is single-GPU: "
              f"each GPU would own a share of the hash buckets)", file=sys.stderr)
        return 2

    import torch
    load_pkg()
    from vsg_amd import synth, vsg

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1, nframes=args.frames)
    W, H, nf = sc.W, sc.H, sc.nframes
    npix = W * H
    fx, fy, cx, cy = sc.intrinsics

    log(f"rendering {nf} synthetic frames {W}x{H}")
    t0 = time.perf_counter()
    frames = []
    for i in range(nf):
        f = sc.render(i)
        f["range"] = synth.depth_range(f["depth"])
        frames.append(f)
    log(f"frames ready in {time.perf_counter() - t0:.1f} s; uploading to HBM")
    dev = torch.device("cuda:0")
    pts = torch.from_numpy(np.stack([f["points"] for f in frames])).to(dev)
    rgb = torch.from_numpy(np.stack([f["colors"] for f in frames])).to(dev)
    nrm = torch.from_numpy(np.stack([f["normals"] for f in frames])).to(dev)
    poses = [f["pose"] for f in frames]
    ranges = [f["range"] for f in frames]
    torch.cuda.synchronize()

    if args.calibration_copy_mib > 0:
        n = args.calibration_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(4):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log(f"calibration copy: 4 x {args.calibration_copy_mib} MiB read + {args.calibration_copy_mib} MiB written")
        del src, dst

    stream = torch.cuda.Stream()

    def make(store):
        scene = vsg.VoxelScene(vsg.Params(wl.buckets, wl.blocks, wl.voxel, wl.trunc), store)
        scene.set_stream(stream.cuda_stream)
        rr = vsg.RayCastRender()
        assert rr.initialize(W, H)
        if args.step_budget is not None:
            rr.set_step_budget(args.step_budget)
        return scene, rr

    def run_pass(scene, rr, per_frame=None):
        scene.reset()
        for i in range(nf):
            T = poses[i]
            scene.integrateFromScanData(pts[i], rgb[i], nrm[i], W, H, wl.trunc, T, fx, fy, cx, cy)
            rr.render(scene, T[:3, 3], T, fx, fy, cx, cy, ranges[i][0], ranges[i][1])
            if per_frame is not None:
                per_frame(i)

    def measure(store):
        scene, rr = make(store)
        b_sw = 8 if store == "fp32" else 4
        # untimed instrumented pass: per-frame counters of the byte model
        log(f"[{store}] instrumented pass for the byte model")
        scene.diag_enable(True)
        stats = []

        def collect(i):
            c = scene.counters()
            d = scene.diag_read(clear=True)
            h, m = rr.counters()
            stats.append({**c, **d, "hits": h, "misses": m})

        run_pass(scene, rr, collect)
        scene.diag_enable(False)
        for w in range(args.warmup):
            log(f"[{store}] warmup pass {w + 1}/{args.warmup}")
            run_pass(scene, rr)
        torch.cuda.synchronize()
        log(f"[{store}] timing {args.steps} passes x {nf} frames")
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            run_pass(scene, rr)
        torch.cuda.synchronize()
        t_total = time.perf_counter() - t_start
        # per-kernel durations: one more pass with every launch bracketed by HIP events on
        # its stream (after the timed passes, so the event records do not slow them)
        vsg.timing_enable(True)
        vsg.timing_reset()
        run_pass(scene, rr)
        torch.cuda.synchronize()
        ktimes = vsg.timing_read()
        ext = None
        if not args.no_extract:
            # surface-point extraction of the fused scene (reference defaults: SDF in
            # [-0.01, 0.01], weight >= 1), outside the timed region
            n_alloc = scene.counters()["num_allocated"]
            rr.extractSurfacePoints(scene)  # first call allocates the outputs
            torch.cuda.synchronize()
            vsg.timing_reset()
            reps = 10
            t0 = time.perf_counter()
            for _ in range(reps):
                npts = rr.extractSurfacePoints(scene)
            wall = (time.perf_counter() - t0) / reps
            ems, en = vsg.timing_read()["extract"]
            ebytes = n_alloc * 512 * b_sw + min(npts, rr.m_maxExtractedPoints) * (4 + 48)
            egbs = ebytes / (ems / en * 1e-3) / 1e9
            ext = {"points": npts, "blocks": n_alloc, "avg_ms": round(ems / en, 4), "wall_ms": round(wall * 1e3, 4),
                   "algorithmic_bytes": int(ebytes), "gbs": round(egbs, 1), "frac": round(egbs / HBM_PEAK_GBS, 5)}
            log(f"[{store}] extraction: {npts} points from {n_alloc} blocks in {ems / en:.3f} ms")
        mesh = None
        if not args.no_mesh:
            # marching cubes of the fused scene, outside the timed region: the reference's
            # cubes (inside blocks, no weight test) and saveToMesh's (cross-block cubes,
            # weight >= 1); algorithmic bytes: the voxel store once and 72 bytes per
            # triangle (3 float3 vertices, 3 float3 normals)
            n_alloc = scene.counters()["num_allocated"]
            out = vsg.MarchingCubesOutput()
            vsg.allocateMarchingCubesOutput(out, 1 << 23)
            mcp = vsg.MarchingCubesParams()
            mcp.voxelSize = 0.0
            mesh = {}
            for name, cross, mw in (("in_block", False, 0), ("cross_block_w1", True, 1)):
                vsg.runMarchingCubesCUDA(scene, scene.m_params, mcp, out, cross_block=cross, min_weight=mw)
                torch.cuda.synchronize()
                vsg.timing_reset()
                reps = 10
                t0 = time.perf_counter()
                for _ in range(reps):
                    ntri = vsg.runMarchingCubesCUDA(scene, scene.m_params, mcp, out, cross_block=cross, min_weight=mw)
                wall = (time.perf_counter() - t0) / reps
                mms, mn = vsg.timing_read()["mesh"]
                mbytes = n_alloc * 512 * b_sw + min(ntri, out.maxTriangles) * 72
                mgbs = mbytes / (mms / mn * 1e-3) / 1e9
                mesh[name] = {"triangles": ntri, "blocks": n_alloc, "avg_ms": round(mms / mn, 4),
                              "wall_ms": round(wall * 1e3, 4), "algorithmic_bytes": int(mbytes),
                              "gbs": round(mgbs, 1), "frac": round(mgbs / HBM_PEAK_GBS, 5)}
                log(f"[{store}] marching cubes {name}: {ntri} triangles in {mms / mn:.3f} ms")
            vsg.freeMarchingCubesOutput(out)
        streaming = None
        if not args.no_stream:
            # streaming round trip of the fused scene: the blocks beyond the median block
            # distance from the last camera go to pinned host memory and come back (wall
            # clock of the synchronising calls, PCIe copies of the block images included)
            torch.cuda.synchronize()
            coords, _ = scene.export_blocks()
            cam = frames[nf - 1]["pose"][:3, 3]
            dist = np.linalg.norm(coords.astype(np.float64) * 8 * wl.voxel - cam, axis=1)
            radius = float(np.median(dist))
            nbi = scene.block_image_bytes()
            hc = torch.empty((len(coords), 3), dtype=torch.int32, pin_memory=True).numpy()
            him = torch.empty((len(coords), nbi), dtype=torch.uint8, pin_memory=True).numpy()
            t0 = time.perf_counter()
            oc, oim = scene.stream_out(cam.astype(np.float32), radius, capacity=len(coords), coords_out=hc,
                                       images_out=him)
            t_out = time.perf_counter() - t0
            t0 = time.perf_counter()
            st_in = scene.stream_in(oc, oim)
            t_in = time.perf_counter() - t0
            assert np.all(st_in == 0) and scene.counters()["num_allocated"] == len(coords)
            streaming = {"blocks": int(len(oc)), "of_blocks": int(len(coords)), "image_bytes": int(oim.nbytes),
                      "out_ms": round(t_out * 1e3, 3), "in_ms": round(t_in * 1e3, 3),
                      "out_gbs": round(oim.nbytes / t_out / 1e9, 2), "in_gbs": round(oim.nbytes / t_in / 1e9, 2)}
            log(f"[{store}] streaming: {len(oc)} of {len(coords)} blocks out in {t_out * 1e3:.2f} ms, "
                f"in in {t_in * 1e3:.2f} ms")
        vsg.timing_enable(False)
        scene.close()
        rr.close()
        kb = {k: 0 for k in ("alloc_claim", "compact", "integrate", "raycast")}
        fb = 0
        for st in stats:
            for k, v in kernel_bytes(st, npix, b_sw).items():
                kb[k] += v
            fb += frame_bytes(st, npix, b_sw)
        log(f"[{store}] {t_total:.3f} s for {args.steps * nf} frames")
        return {"t_total": t_total, "ktimes": ktimes, "kbytes": kb, "fbytes": fb, "stats": stats, "extract": ext,
                "mesh": mesh, "stream": streaming}

    res = measure(args.store)
    t_total, ktimes = res["t_total"], res["ktimes"]
    frames_done = args.steps * nf
    fps = frames_done / t_total
    # dominant kernel: most total time in the timed region
    kern_ms = {k: v[0] for k, v in ktimes.items() if v[1] > 0}
    dom = max(kern_ms, key=kern_ms.get)
    launches = ktimes[dom][1]
    avg_ms = ktimes[dom][0] / launches
    bytes_per_launch = res["kbytes"].get(dom, 0) / nf
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a
    # copy kernel of known size) on this workload, recorded in profiles/
    traffic = None
    tr_path = os.path.join(REPO, "profiles", f"{dom}_room640_traffic.md")
    if nf == 300 and args.store == "fp32" and os.path.exists(tr_path):
        for line in open(tr_path):
            if line.startswith("bytes_per_launch:"):
                traffic = int(line.split(":", 1)[1])
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": dom,
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(avg_ms, 4)}
    fused_gbs = res["fbytes"] * args.steps / t_total / 1e9
    per_kernel = {}
    for k, (ms, n) in ktimes.items():
        if n == 0 or k in ("extract", "mesh"):
            continue
        e = {"avg_ms": round(ms / n, 4), "launches": n, "share": round(ms / sum(kern_ms.values()), 4)}
        if k in res["kbytes"]:
            gbs = res["kbytes"][k] / nf / (ms / n * 1e-3) / 1e9
            e["gbs"] = round(gbs, 1)
            e["frac"] = round(gbs / HBM_PEAK_GBS, 5)
        per_kernel[k] = e
    st = res["stats"]
    extra = {
        "ms_per_frame": round(t_total / frames_done * 1e3, 4),
        "fused_roofline": {"achieved": round(fused_gbs, 2), "frac": round(fused_gbs / HBM_PEAK_GBS, 5),
                           "bytes_per_frame": int(res["fbytes"] / nf)},
        "kernels": per_kernel,
        "scene_end": {"blocks": st[-1]["num_allocated"], "visible_mean": int(np.mean([s["visible_blocks"] for s in st])),
                      "updated_mean": int(np.mean([s["updated_voxels"] for s in st])),
                      "ray_blocks_mean": int(np.mean([s["ray_blocks"] for s in st])),
                      "hit_rate": round(float(np.mean([s["hits"] / npix for s in st])), 4)},
        "march_per_ray": {k[6:]: round(float(np.mean([s[k] for s in st])) / npix, 2)
                          for k in ("march_evals", "march_super_skips", "march_block_skips", "march_replayed",
                                    "march_cell_skips", "march_deferred", "march_tail_windows", "march_tail_skips")},
    }
    if res["extract"] is not None:
        extra["surface_points"] = res["extract"]
    if res["mesh"] is not None:
        extra["marching_cubes"] = res["mesh"]
    if res["stream"] is not None:
        extra["streaming"] = res["stream"]
    if not args.no_shim and nf == 300:
        log(f"C++ drop-in path: {args.steps} passes over the same sequence")
        sh = shim_bench(nf, args.steps)
        if sh is not None:
            sh["ratio_to_python"] = round(sh["value"] / fps, 4)
            extra["cpp_shim"] = sh
    if not args.no_depthgen:
        extra["depth_from_points"] = depthgen_bench(torch, vsg, sc, frames)
    if not args.no_pose and nf >= 128:
        extra["pose_refinement"] = pose_refinement_bench(torch, vsg, synth, sc, frames, wl)
    if args.large_frames > 0:
        log(f"large run: {args.large_frames} frames 1024x768 with tracking")
        extra["large_run"] = large_run_bench(torch, vsg, synth, args.large_frames, args.large_stream_radius,
                                             args.large_trace)
    if args.compare_fp16 and args.store == "fp32":
        r16 = measure("fp16")
        extra["fp16_store"] = {"value": round(frames_done / r16["t_total"], 3),
                               "speedup_vs_fp32": round(t_total / r16["t_total"], 4),
                               "ms_per_frame": round(r16["t_total"] / frames_done * 1e3, 4)}
        if r16["extract"] is not None:
            extra["fp16_store"]["surface_points"] = r16["extract"]
        if r16["mesh"] is not None:
            extra["fp16_store"]["marching_cubes"] = r16["mesh"]
    if not args.no_precision:
        log(f"precision: FP32 and FP16 stores vs the FP64 oracle on {args.precision_frames} frames")
        extra["precision_vs_fp64"] = precision_report(torch, vsg, sc, frames, wl, pts, rgb, nrm,
                                                      args.precision_frames, 20000)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle on 30 room frames and 30 sphere frames, all cores")
        cpu = cpu_baseline(synth)
    out = {
        "metric": "frames/s 640x480 integrate+raycast, % HBM roofline; fp16 vs fp32 at ≤0.5mm err",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.store == "fp32" else "f32 math, f16 store",
        "data": "synthetic",
        "config": {"workload": f"room 640x480, {nf} frames, voxel 4 mm, trunc 0.02 m, 2^20 buckets, "
                               f"{args.store} store; step = one pass over the sequence from an empty scene",
                   "frames_per_step": nf, "store": args.store},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
