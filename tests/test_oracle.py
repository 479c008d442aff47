"""Pins the oracle (the CPU restatement of the reference path) against facts fixed
independently of its code: hand-computed hash vectors, brute-force rounding rules,
dense sampling of DDA segments, numpy's IEEE half conversion, hand-enumerable
allocation cases, closed-form SDF values of a plane, and a vectorised numpy
evaluation of the trilinear sampler over the exported store.

The reference ships no tests or fixtures (it has none), so these checks are the pins.
"""
import math

import numpy as np
import pytest

import oracle as O
from vsg_amd import synth

# ---------------------------------------------------------------------------------
# A2: hash
# ---------------------------------------------------------------------------------


def _hash_ref(x, y, z, B):
    # two's-complement 32-bit products, xor, unsigned remainder
    m = 0xFFFFFFFF
    h = ((x * 73856093) & m) ^ ((y * 19349663) & m) ^ ((z * 83492791) & m)
    return h % B


def test_hash_known_answers():
    # (1,1,1): 73856093 ^ 19349663 ^ 83492791 = 28855157 (no wrap-around)
    assert O.hash_bucket(1, 1, 1, 1 << 22) == 28855157 % (1 << 22)
    assert O.hash_bucket(1, 1, 1, 200000) == 28855157 % 200000
    assert O.hash_bucket(0, 0, 0, 200000) == 0
    # (-1,0,0): (-73856093) mod 2^32 = 4221111203
    assert O.hash_bucket(-1, 0, 0, 1 << 20) == 4221111203 % (1 << 20)
    assert O.hash_bucket(-1, 0, 0, 200000) == 4221111203 % 200000
    rng = np.random.default_rng(0)
    for _ in range(2000):
        x, y, z = (int(v) for v in rng.integers(-(1 << 20), 1 << 20, 3))
        B = int(rng.choice([1, 7, 211, 200000, 1 << 20, 1 << 22]))
        assert O.hash_bucket(x, y, z, B) == _hash_ref(x, y, z, B)


# ---------------------------------------------------------------------------------
# Q1 / Q2: world -> voxel -> block rounding
# ---------------------------------------------------------------------------------


@pytest.mark.parametrize("v", [1.0, 0.25])
def test_allocator_rounding_bruteforce(v):
    # q = p / v exact for these v; allocator: round half away from zero, block = floor(vi / 8)
    for k in range(-20 * 64, 20 * 64 + 1):
        q = k / 64.0
        vox, blk = O.alloc_world_to_block([q * v, -q * v, 0.0], v)
        exp = int(math.copysign(math.floor(abs(q) + 0.5), q)) if q != 0 else 0
        exp2 = int(math.copysign(math.floor(abs(q) + 0.5), -q)) if q != 0 else 0
        assert vox[0] == exp and vox[1] == exp2, q
        assert blk[0] == exp // 8 and blk[1] == exp2 // 8, q


@pytest.mark.parametrize("v", [1.0, 0.25])
def test_raycaster_rounding_bruteforce(v):
    # raycaster: floor(q + 0.5) for q > 0, floor(q - 0.5) for q < 0; local = vi mod 8
    for k in range(-20 * 64, 20 * 64 + 1):
        q = k / 64.0
        vox, blk, loc = O.ray_world_to_voxel([q * v, 0.0, 0.0], v)
        exp = math.floor(q + 0.5) if q > 0 else (math.floor(q - 0.5) if q < 0 else 0)
        assert vox[0] == exp and blk[0] == exp // 8 and loc[0] == exp % 8, q


def test_negative_side_asymmetry():
    # For negative coordinates the raycaster reads one voxel below the allocator's
    v = 1.0
    a, _ = O.alloc_world_to_block([-2.4, 2.4, 0.0], v)
    r, _, _ = O.ray_world_to_voxel([-2.4, 2.4, 0.0], v)
    assert a[0] == -2 and r[0] == -3
    assert a[1] == 2 and r[1] == 2


# ---------------------------------------------------------------------------------
# A5: DDA over blocks
# ---------------------------------------------------------------------------------


def _block_of(p, v):
    # block k spans [8k - 0.5, 8k + 7.5) voxels on every axis
    return np.floor((np.asarray(p, np.float64) / v + 0.5) / 8.0).astype(np.int64)


def test_dda_dense_sampling():
    rng = np.random.default_rng(1)
    v = 0.004
    for trial in range(300):
        a = rng.uniform(-1.0, 1.0, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        b = a + d * rng.uniform(0.0, 0.25)
        got = O.dda(a, b, v)
        sa, sb = _block_of(np.float32(a), v), _block_of(np.float32(b), v)
        if np.array_equal(sa, sb):
            assert len(got) == 0  # start == end block: nothing (VoxelScene.cu:345-348)
            continue
        assert np.array_equal(got[0], sa) and np.array_equal(got[-1], sb)
        steps = np.abs(np.diff(got.astype(np.int64), axis=0)).sum(axis=1)
        assert np.all(steps == 1), "blocks must be 6-connected"
        # every block densely sampled away from block faces is visited, and every visited
        # block comes within 1e-3 voxel of the segment
        t = np.linspace(0.0, 1.0, 20001)
        pts = a[None, :] + (b - a)[None, :] * t[:, None]
        q = pts / v + 0.5
        frac = q / 8.0 - np.floor(q / 8.0)
        interior = np.all((frac > 1e-3 / 8) & (frac < 1 - 1e-3 / 8), axis=1)
        sampled = {tuple(x) for x in _block_of(pts[interior], v)}
        visited = {tuple(int(c) for c in x) for x in got}
        assert sampled <= visited, (trial, sampled - visited)
        for blk in visited - sampled:
            lo = (np.array(blk) * 8 - 0.5) * v - 1e-3 * v
            hi = (np.array(blk) * 8 + 7.5) * v + 1e-3 * v
            inside = np.all((pts >= lo) & (pts <= hi), axis=1)
            assert inside.any(), (trial, blk)


def test_dda_axis_parallel():
    v = 0.01
    got = O.dda([0.001, 0.002, 0.1], [0.001, 0.002, 0.5], v)
    assert np.all(got[:, 0] == 0) and np.all(got[:, 1] == 0)
    assert list(got[:, 2]) == list(range(1, 7))


# ---------------------------------------------------------------------------------
# FP16 store conversion
# ---------------------------------------------------------------------------------


def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(2)
    vals = np.concatenate([
        rng.normal(0, 0.02, 20000), rng.normal(0, 1, 5000), rng.uniform(-70000, 70000, 2000),
        np.exp(rng.uniform(-30, 12, 5000)) * rng.choice([-1, 1], 5000),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e-8, 5.96e-8, 2.98e-8, 6.1e-5, np.inf, -np.inf]]).astype(np.float32)
    for x in vals:
        h = O.f32_to_f16(float(x))
        with np.errstate(over="ignore"):
            exp = int(np.array([x], np.float32).astype(np.float16).view(np.uint16)[0])
        assert h == exp, (x, hex(h), hex(exp))
        back = O.f16_to_f32(h)
        assert back == float(np.array([exp], np.uint16).view(np.float16)[0].astype(np.float32)) or math.isnan(back)


# ---------------------------------------------------------------------------------
# A15: depth range rule
# ---------------------------------------------------------------------------------


def test_depth_range_rule():
    rng = np.random.default_rng(3)
    d = rng.uniform(0.3, 3.0, (48, 64)).astype(np.float32)
    d[rng.random(d.shape) < 0.3] = 0.0
    v = d[d > 0]
    mn, mx = np.float32(v.min()), np.float32(v.max())
    margin = np.float32((mx - mn) * np.float32(0.1))
    exp = (float(max(np.float32(0.01), np.float32(mn - margin))), float(np.float32(mx + margin)))
    assert O.depth_range(d) == exp
    assert synth.depth_range(d) == exp


# ---------------------------------------------------------------------------------
# A4/A6: allocation on hand-enumerable scenes
# ---------------------------------------------------------------------------------

EYE = np.eye(4)


def _plane_frame(W, H, d, fx, fy, cx, cy):
    u, vv = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
    z = np.full((H, W), d, np.float32)
    pts = np.stack([(u - np.float32(cx)) * z / np.float32(fx), (vv - np.float32(cy)) * z / np.float32(fy), z], -1)
    return pts.astype(np.float32)


def _plane_blocks_by_hand(pts, v, d, tau):
    exp = set()
    for p in pts.reshape(-1, 3).astype(np.float64):
        bx, by = _block_of(p, v)[:2]
        zs = {_block_of([0, 0, d - tau], v)[2], _block_of([0, 0, d + tau], v)[2]}
        assert zs == {9, 10}  # boundary at 8*10*v - v/2 = 0.795
        for bz in range(min(zs), max(zs) + 1):
            exp.add((int(bx), int(by), int(bz)))
    return exp


def test_plane_block_set_by_hand():
    # identity pose: the band of every pixel runs along z from d-tau to d+tau at fixed (x, y)
    v, tau, d = 0.01, 0.02, 0.8
    pts = _plane_frame(4, 4, d, 10.0, 10.0, -5.5, -1.5)  # x blocks 5..8, y blocks 1..4
    B = 1000003
    o = O.OracleScene(B, 1000, v)
    o.alloc(pts, tau, EYE)
    got = {tuple(int(x) for x in c) for c in o.export_blocks()[0]}
    exp = _plane_blocks_by_hand(pts, v, d, tau)
    assert len(exp) == 32
    # precondition: no two expected blocks share a bucket (one insertion per bucket per frame)
    assert len({_hash_ref(*b, B) for b in exp}) == len(exp)
    assert got == exp


def test_colliding_blocks_are_deferred():
    # the XOR hash maps several blocks of a plane centred on the axis to one bucket: only
    # the first request of each bucket (in pixel, step order) is inserted per frame
    v, tau, d = 0.01, 0.02, 0.8
    pts = _plane_frame(4, 4, d, 10.0, 10.0, 1.5, 1.5)  # blocks -2..1 in x and y
    want = _plane_blocks_by_hand(pts, v, d, tau)
    for B in (1 << 16, 1000003):
        o = O.OracleScene(B, 1000, v)
        o.alloc(pts, tau, EYE)
        n_buckets = len({_hash_ref(*b, B) for b in want})
        assert n_buckets < len(want)
        assert o.num_allocated() == n_buckets
        for _ in range(10):
            o.alloc(pts, tau, EYE)
        assert {tuple(int(x) for x in c) for c in o.export_blocks()[0]} == want


def test_single_bucket_one_new_block_per_frame():
    v, tau, d = 0.01, 0.02, 0.8
    pts = _plane_frame(4, 4, d, 10.0, 10.0, 1.5, 1.5)
    o = O.OracleScene(1, 1000, v)
    first = O.dda(pts[0, 0].astype(np.float64) - [0, 0, tau], pts[0, 0].astype(np.float64) + [0, 0, tau], v)[0]
    o.alloc(pts, tau, EYE)
    coords, ptrs, slots = o.export_blocks()
    assert len(coords) == 1 and tuple(coords[0]) == tuple(first) and ptrs[0] == 0 and slots[0] == 0
    for k in range(2, 12):
        o.alloc(pts, tau, EYE)
        assert o.num_allocated() == min(k, 10)  # 32 distinct blocks requested, the bucket holds 10


def test_heap_exhaustion_order():
    v, tau, d = 0.01, 0.02, 0.8
    pts = _plane_frame(4, 4, d, 10.0, 10.0, 1.5, 1.5)
    full = O.OracleScene(1 << 16, 1000, v)
    full.alloc(pts, tau, EYE)
    order = [tuple(c) for c in full.export_blocks()[0]]
    o = O.OracleScene(1 << 16, 3, v)
    o.alloc(pts, tau, EYE)
    assert [tuple(c) for c in o.export_blocks()[0]] == order[:3]
    assert o.num_allocated() == 3 and o.heap_counter() == 0xFFFFFFFF
    o.alloc(pts, tau, EYE)
    assert o.num_allocated() == 3


def test_per_bucket_insert_invariant():
    # over frames of workload 0 with few buckets, a bucket gains at most one block per frame
    wl = synth.WORKLOADS[0]
    sc = synth.SynthScene(0)
    B = 997
    o = O.OracleScene(B, 1 << 16, wl.voxel)
    prev = np.zeros(B, np.int64)
    for i in range(0, 30, 5):
        f = sc.render(i)
        o.alloc(f["points"], wl.trunc, f["pose"])
        _, _, slots = o.export_blocks()
        cnt = np.bincount(slots // 10, minlength=B)
        assert np.all(cnt - prev <= 1) and np.all(cnt <= 10)
        prev = cnt


# ---------------------------------------------------------------------------------
# A8: integration closed forms
# ---------------------------------------------------------------------------------


def _expected_plane_update(o, pts, tau, fx, fy, cx, cy):
    """float64 recomputation of projection and Euclidean SDF for every allocated voxel."""
    coords, ptrs, _ = o.export_blocks()
    v = o.voxel
    H, W = pts.shape[:2]
    l = np.arange(512)
    lx, ly, lz = l % 8, (l // 8) % 8, l // 64
    vw = (coords[:, None, :] * 8 + np.stack([lx, ly, lz], -1)[None]) * np.float64(np.float32(v))
    vw = vw.reshape(-1, 3)
    sx = vw[:, 0] * fx / vw[:, 2] + cx
    sy = vw[:, 1] * fy / vw[:, 2] + cy
    px, py = np.trunc(sx + 0.5), np.trunc(sy + 0.5)
    inimg = (px >= 0) & (px < W) & (py >= 0) & (py < H)
    amb = (np.abs(sx + 0.5 - np.round(sx + 0.5)) < 1e-4) | (np.abs(sy + 0.5 - np.round(sy + 0.5)) < 1e-4)
    pix = np.where(inimg, py * W + px, 0).astype(np.int64)
    D = pts.reshape(-1, 3).astype(np.float64)[pix]
    sdf = np.linalg.norm(vw - D, axis=1)
    sdf = np.where(np.linalg.norm(vw, axis=1) < np.linalg.norm(D, axis=1), -sdf, sdf)
    upd = inimg & (np.abs(sdf) <= tau)
    amb |= np.abs(np.abs(sdf) - tau) < 1e-6
    return upd, sdf, amb


def test_plane_integrate_closed_form():
    v, tau, d = 0.004, 0.02, 0.6
    W, H, fx, cx, cy = 64, 48, 80.0, 31.5, 23.5
    pts = _plane_frame(W, H, d, fx, fx, cx, cy)
    rgb = np.full((H, W, 3), 90, np.uint8)
    o = O.OracleScene(1 << 16, 1 << 14, v)
    o.integrate(pts, rgb, tau, EYE, (fx, fx, cx, cy))
    sdf, w, col, fl = o.export_voxels()
    upd, esdf, amb = _expected_plane_update(o, pts, tau, fx, fx, cx, cy)
    ok = ~amb
    assert np.array_equal(w[ok] == 1, upd[ok]) and np.all(w <= 1)
    assert np.abs(sdf[ok & upd] - esdf[ok & upd]).max() < 2e-6
    assert np.all(col[w == 1] == 90) and np.all(fl[w == 1] & 1)
    # sign: the rule compares distances to the camera; away from the plane (more than
    # the lateral pixel footprint can change) that is the side of the plane
    coords = o.export_blocks()[0]
    z = ((coords[:, None, 2] * 8 + (np.arange(512) // 64)[None]) * np.float32(v)).reshape(-1)
    sel = (w == 1) & (np.abs(z - d) > 5e-3)
    assert sel.sum() > 1000
    assert np.all(np.sign(sdf[sel]) == np.sign(z[sel] - d))


def test_repeated_frames_weight_and_colour():
    v, tau, d = 0.004, 0.02, 0.6
    W, H, fx, cx, cy = 16, 12, 20.0, 7.5, 5.5
    pts = _plane_frame(W, H, d, fx, fx, cx, cy)
    o = O.OracleScene(1 << 14, 1 << 12, v)
    o.integrate(pts, np.full((H, W, 3), 100, np.uint8), tau, EYE, (fx, fx, cx, cy))
    s1, w1, _, _ = o.export_voxels()
    n1 = o.num_allocated()  # later frames may add blocks deferred by bucket collisions
    o.integrate(pts, np.full((H, W, 3), 201, np.uint8), tau, EYE, (fx, fx, cx, cy))
    s2, w2, c2, _ = o.export_voxels(n1)
    m = w1 == 1
    assert m.sum() > 100
    assert np.all(w2[m] == 2) and np.all(c2[m] == 150)  # (100*1 + 201)/2 = 150.5 truncates
    assert np.abs(s2[m] - s1[m]).max() < 1e-7
    for _ in range(300):
        o.integrate(pts, np.full((H, W, 3), 100, np.uint8), tau, EYE, (fx, fx, cx, cy))
    s3, w3, _, _ = o.export_voxels(n1)
    assert np.all(w3[m] == 255) and np.abs(s3[m] - s1[m]).max() < 1e-6


def test_visible_only_integration_matches_all_blocks():
    wl = synth.WORKLOADS[0]
    sc = synth.SynthScene(0)
    a = O.OracleScene(wl.buckets, wl.blocks, wl.voxel)
    b = O.OracleScene(wl.buckets, wl.blocks, wl.voxel)
    culled = 0
    for i in range(0, 30, 2):
        f = sc.render(i)
        ca = a.integrate(f["points"], f["colors"], wl.trunc, f["pose"], sc.intrinsics)
        cb = b.integrate(f["points"], f["colors"], wl.trunc, f["pose"], sc.intrinsics, visible_only=True)
        assert ca["updated"] == cb["updated"]
        culled += ca["visible"] - cb["visible"]
    assert culled > 0
    for x, y in zip(a.export_voxels(), b.export_voxels()):
        assert np.array_equal(x, y)


# ---------------------------------------------------------------------------------
# A12/A13: trilinear sampler against a numpy evaluation over the exported store
# ---------------------------------------------------------------------------------


def _numpy_trilinear(coords, sdf, w, p, v):
    v = np.float32(v)
    lut = {}
    for k, c in enumerate(coords):
        lut[tuple(int(x) for x in c)] = k
    p = np.asarray(p, np.float32)
    pd = p - v / np.float32(2)
    q = pd / v
    f = q - np.floor(q)

    def rvox(x):  # raycaster rounding of the quotient x / v
        qq = np.float32(x) / v
        return int(math.floor(qq + 0.5)) if qq > 0 else (int(math.floor(qq - 0.5)) if qq < 0 else 0)

    num = den = 0.0
    for i in range(8):
        c = [pd[a] + (v if (i >> a) & 1 else np.float32(0)) for a in range(3)]
        vi = [rvox(x) for x in c]
        blk = tuple(x // 8 for x in vi)
        if blk not in lut:
            continue
        idx = lut[blk] * 512 + (vi[0] % 8) + 8 * (vi[1] % 8) + 64 * (vi[2] % 8)
        if w[idx] == 0:
            continue
        wt = np.prod([f[a] if (i >> a) & 1 else 1 - f[a] for a in range(3)])
        num += wt * sdf[idx]
        den += wt
    return (den > 1e-6), (num / den if den > 1e-6 else 0.0)


def test_trilinear_matches_numpy_store_evaluation():
    wl = synth.WORKLOADS[0]
    sc = synth.SynthScene(0)
    o = O.OracleScene(wl.buckets, wl.blocks, wl.voxel)
    f = sc.render(0)
    o.integrate(f["points"], f["colors"], wl.trunc, f["pose"], sc.intrinsics)
    coords, _, _ = o.export_blocks()
    sdf, w, _, _ = o.export_voxels()
    rng = np.random.default_rng(4)
    pts = f["points"].reshape(-1, 3)
    valid = np.flatnonzero(pts[:, 2] > 0)
    T = f["pose"].astype(np.float64)
    n_ok = 0
    for k in rng.choice(valid, 400, replace=False):
        P = pts[k].astype(np.float64)
        world = T[:3, :3] @ P + T[:3, 3] + rng.normal(0, 0.01, 3)
        ok, dist, _ = o.trilinear(world)
        eok, edist = _numpy_trilinear(coords, sdf, w, world, wl.voxel)
        assert ok == eok
        if ok:
            n_ok += 1
            assert abs(dist - edist) < 1e-6
    assert n_ok > 200


# ---------------------------------------------------------------------------------
# A10-A14: raycast invariants
# ---------------------------------------------------------------------------------


def test_raycast_plane_within_half_voxel_bias():
    # the frac(p/v - 1/2) weights sample the field up to half a voxel off, and the
    # refined hit lies inside the one-step (v/2) bracket where the sampled field
    # changes sign, so a plane's hit depth lies within v of the plane
    v, tau, d = 0.004, 0.02, 0.6
    W, H, fx, cx, cy = 64, 48, 80.0, 31.5, 23.5
    pts = _plane_frame(W, H, d, fx, fx, cx, cy)
    o = O.OracleScene(1 << 16, 1 << 14, v)
    for _ in range(4):  # later frames insert the blocks bucket collisions deferred
        o.integrate(pts, np.full((H, W, 3), 90, np.uint8), tau, EYE, (fx, fx, cx, cy))
    r = o.render(W, H, [0, 0, 0], EYE, (fx, fx, cx, cy), 0.3, 0.9)
    hit = r["depth"] < 1e9
    assert r["hits"] == hit.sum() and r["hits"] + r["misses"] == W * H
    assert hit.mean() > 0.9
    err = np.abs(r["depth"][hit] - d)
    assert err.max() < v + 2e-4 and np.median(err) < 0.5 * v
    # positions lie on the camera ray at the reported depth; colours are the fused colour
    u = np.arange(W * H) % W
    assert np.allclose(r["pos"][hit, 0], (u[hit] - cx) / fx * r["depth"][hit], atol=1e-5)
    # interpolated colour is truncated to uchar: a weighted mean of 90s may give 89.99998 -> 89
    c = np.round(r["col"][hit, :3] * 255.0)
    assert np.all((c == 90) | (c == 89)) and (c == 90).mean() > 0.5 and np.all(r["nrm"] == 1e10)


def test_raycast_empty_space_and_guards():
    v = 0.004
    o = O.OracleScene(1 << 10, 64, v)
    r = o.render(8, 8, [0, 0, 0], EYE, (8.0, 8.0, 3.5, 3.5), 0.3, 0.9, init=7.0)
    assert not r["ran"] and np.all(r["depth"] == 7.0)  # no block: outputs untouched
    pts = _plane_frame(8, 8, 0.6, 8.0, 8.0, 3.5, 3.5)
    o.integrate(pts, np.zeros((8, 8, 3), np.uint8), 0.02, EYE, (8.0, 8.0, 3.5, 3.5))
    r = o.render(8, 8, [0, 0, 0], EYE, (8.0, 8.0, 3.5, 3.5), 0.7, 0.9)  # range behind the plane
    assert r["ran"] and r["hits"] == 0 and r["misses"] == 64 and np.all(r["depth"] == 1e10)
    r = o.render(8, 8, [0, 0, 0], EYE, (8.0, 8.0, 3.5, 3.5), 0.0, 0.9)  # minDepth 0: no march at all
    assert r["hits"] == 0 and r["misses"] == 0 and np.all(r["depth"] == 1e10)
