"""Pins the oracle's ray refinement and acceptance rule (RayCastRender.cu:220-271,
302-397) with designed SDF fields whose answers are worked out by hand.

Set-up: a fused plane allocates the blocks around the optical axis; every voxel of
every allocated block is then overwritten with a field that depends only on the
voxel's z index k (world z = k * v).  Pixel (32, 24) with cx = 32, cy = 24 looks
along +z from the origin, so its samples sit at x = y = 0, where the trilinear
sampler's x and y weights are exactly 1/2 and cancel: the sampled distance at ray
position t depends only on q = (t - v/2) / v = n + w (n integer, w in [0, 1)).  The
sampler reads the voxels at z = q and q + 1 through the raycaster's rounding
floor(. + 1/2) and weighs the second by w, so on this ray

    w <  1/2:  f(q) = (1 - w) s(n)     + w s(n + 1)
    w >= 1/2:  f(q) = (1 - w) s(n + 1) + w s(n + 2)

(voxels of weight 0 dropped, the rest renormalised), with jumps at w = 1/2 and at
integer q.  `field` below is that restriction, written from these two lines;
`_numpy_trilinear` in test_oracle.py pins the sampler itself on fused data.

The march starts at q = n0 + 1/4 and steps v/2, so samples alternate between w = 1/4
and w = 3/4: the sample at q = k - 1/4 reads (k, k+1) as .25 s(k) + .75 s(k+1), the
one at q = k + 1/4 reads (k, k+1) as .75 s(k) + .25 s(k+1).  Each scenario lists its
sample sequence; a test asserts it, so an arithmetic slip in a comment shows up.

Each of these variants of the rule fails at least one assertion: two or four
regula-falsi steps, a swapped bracket update, the colour of another iterate, a 2v
or 4v threshold, no lastWeight reset on an invalid sample, an accepted
positive-to-negative change, an inclusive range end.
"""
import math

import numpy as np
import pytest

import oracle as O

V = np.float32(0.004)
W, H, FX, CX, CY = 64, 48, 80.0, 32.0, 24.0
INTR = (FX, FX, CX, CY)
AXIS = 24 * W + 32
EYE = np.eye(4)
MM = 1e-3


def _plane(d):
    u, vv = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
    z = np.full((H, W), d, np.float32)
    return np.stack([(u - np.float32(CX)) * z / np.float32(FX), (vv - np.float32(CY)) * z / np.float32(FX), z],
                    -1).astype(np.float32)


def plane_frame():
    """The frame that allocates the scenarios' blocks (a plane at 0.6 m)."""
    return _plane(0.6)


def _scene():
    """Blocks of z index 18 and 19 (voxels 144..159) around the axis, allocated by a
    plane at 0.6 m fused four times (later frames insert what collisions deferred)."""
    o = O.OracleScene(1 << 16, 1 << 14, float(V))
    pts = plane_frame()
    for _ in range(4):
        o.integrate(pts, np.full((H, W, 3), 90, np.uint8), 0.02, EYE, INTR)
    coords, ptrs, _ = o.export_blocks()
    have = {tuple(c) for c in coords}
    for bx in (-1, 0):
        for by in (-1, 0):
            for bz in (18, 19):
                assert (bx, by, bz) in have
    return o, coords, ptrs


def designed_store(coords, layers_mm, colors=None):
    """(sdf [n*512] metres, weight [n*512], rgb [n*512, 3]) of every allocated block:
    voxel z index k gets layers_mm[k] millimetres with weight 1 and grey level
    colors[k] (default 0); every other voxel weight 0."""
    lz = np.arange(512) // 64
    kz = coords[:, 2:3] * 8 + lz[None, :]
    sdf = np.zeros(kz.shape, np.float64)
    w = np.zeros(kz.shape, np.uint8)
    rgb = np.zeros(kz.shape + (3,), np.uint8)
    for k, s in layers_mm.items():
        m = kz == k
        sdf[m] = s * MM
        w[m] = 1
        if colors and k in colors:
            rgb[m] = colors[k]
    return sdf.ravel(), w.ravel(), rgb.reshape(-1, 3)


def _set(o, coords, ptrs, layers_mm, colors=None):
    sdf, w, rgb = designed_store(coords, layers_mm, colors)
    idx = (ptrs[:, None].astype(np.int64) * 512 + np.arange(512)[None, :]).ravel()
    o.set_voxels(idx, sdf, w, rgb)


def march_start(n0):
    """minDepth that puts the first sample at q = n0 + 1/4."""
    return float(np.float32(V * np.float32(n0 + 0.75)))


def sample_positions(min_depth, count):
    """The reference's sample positions: float32 sums of v * 0.5f from minDepth."""
    t = [np.float32(min_depth)]
    inc = np.float32(V * np.float32(0.5))
    for _ in range(count - 1):
        t.append(np.float32(t[-1] + inc))
    return np.array(t, np.float64)


def field(layers_mm, q):
    """The sampled distance (mm) at q on the axis ray, or None where no voxel is weighted."""
    n = math.floor(q)
    w = q - n
    k0 = n if w < 0.5 else n + 1
    num = den = 0.0
    for k, wk in ((k0, 1 - w), (k0 + 1, w)):
        if k in layers_mm and wk > 0:
            num += wk * layers_mm[k]
            den += wk
    return num / den if den > 1e-6 else None


def regula_falsi_3(layers_mm, qa, qb):
    """The reference's three steps on [qa, qb]: c = a + fa/(fa - fb) (b - a); the end
    whose value has the sign of f(c) moves to c.  Returns c3 as a fraction of [qa, qb]."""
    f = lambda u: field(layers_mm, qa + u * (qb - qa))  # noqa: E731
    a, fa, b, fb = 0.0, f(0.0), 1.0, f(1.0)
    c = 0.0
    for _ in range(3):
        c = a + fa / (fa - fb) * (b - a)
        fc = f(c)
        if fa * fc > 0:
            a, fa = c, fc
        else:
            b, fb = c, fc
    return c


def _sequence(layers_mm, n0, count):
    return [field(layers_mm, n0 + 0.25 + 0.5 * j) for j in range(count)]


def _render(o, min_depth, max_depth):
    r = o.render(W, H, [0, 0, 0], EYE, INTR, min_depth, max_depth, pixels=[AXIS])
    return r["depth"][AXIS], r["col"][AXIS], r["hits"], r["misses"]


# The scenarios (layers in mm, first sample q = 146.25, maxDepth 0.64 m), shared with
# the GPU test that runs them through the C ABI: name -> (layers, colours).
REFINE = {146: -3, 147: -3, 148: -1, 149: 3}
REFINE_COLORS = {147: 0, 148: 200, 149: 40}


def threshold_layers(s148):
    return {146: -12, 147: -12, 148: s148, 149: -12, 150: -12, 151: -1, 152: 3, 153: 3, 154: 3}


RESET = {146: -0.5, 149: 10, 150: 10, 151: 10, 152: -3, 153: -3, 154: 2, 155: 2, 156: 2}
POS_NEG = {**{k: 1 for k in range(144, 149)}, **{k: -5 for k in range(149, 153)}, **{k: 5 for k in range(153, 158)}}
SCENARIOS = {
    "refine": (REFINE, REFINE_COLORS),
    "accept_11.8": (threshold_layers(11.6), None),
    "reject_12.2": (threshold_layers(12.4), None),
    "reset": (RESET, None),
    "pos_neg": (POS_NEG, None),
}


def test_three_regula_falsi_steps_with_a_kink():
    # Samples q = 146.25, 146.75, 147.25, 147.75: -3, -1.5, -2.5, +2 mm; the first
    # negative-to-positive change is between t2 and t3.  In u = (t - t2) / (t3 - t2) the
    # bracket's field is -2.5 + u (mm) for u < 1/2 (voxels 147, 148) and 2u after (148, 149).
    #   c1 = 2.5 / 4.5                                  = 5/9,   f = 10/9   > 0: b = c1
    #   c2 = 2.5 / (2.5 + 10/9) * 5/9                   = 5/13,  f = -55/26 < 0: a = c2
    #   c3 = 5/13 + (55/26) / (55/26 + 10/9) (5/9 - 5/13) = 75/151
    # (a fourth step: 0.5346).  The colour is the last evaluation's, at c3:
    # w = 1/4 + c3/2 = 0.49834 on voxels 147 (grey 0) and 148 (grey 200): 99.67 -> 99.
    assert _sequence(REFINE, 146, 4) == pytest.approx([-3, -1.5, -2.5, 2])
    assert regula_falsi_3(REFINE, 147.25, 147.75) == pytest.approx(75 / 151)
    o, coords, ptrs = _scene()
    _set(o, coords, ptrs, REFINE, REFINE_COLORS)
    md = march_start(146)
    t = sample_positions(md, 4)
    depth, col, hits, misses = _render(o, md, 0.64)
    assert hits == 1 and misses == 0
    step = t[3] - t[2]
    iterates = [5 / 9, 5 / 13, 75 / 151, 0.5346]
    assert min(abs(a - b) for i, a in enumerate(iterates) for b in iterates[i + 1:]) * float(step) > 2e-5
    assert abs(depth - (t[2] + 75 / 151 * step)) < 1e-6
    assert np.array_equal(np.round(col[:3] * 255), [99, 99, 99]) and col[3] == 1.0


@pytest.mark.parametrize("s148,accept", [(11.6, True), (12.4, False)])
def test_acceptance_threshold_three_voxels(s148, accept):
    # v = 4 mm: a change is accepted when |lastSdf - dist| < 3v = 12 mm (|dist| < 3v then
    # follows).  Samples (mm) from q = 146.25, s148 = 11.6 | 12.4:
    #   -12, 5.7 | 6.3   (jump 17.7 | 18.3: rejected, the march goes on)
    #   -6.1 | -5.9, -6.1 | -5.9, 5.7 | 6.3   (q = 147.75 -> 148.25: jump 11.8 | 12.2)
    #   -12, -12, -3.75, -9.25, 2.0   (q = 150.25 -> 150.75: jump 11.25)
    # 11.8 is accepted; with 12.2 the march goes on to the 11.25 jump.  A 2v threshold
    # rejects every change (a miss); 4v accepts 12.2.
    layers = threshold_layers(s148)
    seq = _sequence(layers, 146, 10)
    jump = s148 / 2 + 6
    assert seq == pytest.approx([-12, 3 * s148 / 4 - 3, s148 / 4 - 9, s148 / 4 - 9, 3 * s148 / 4 - 3, -12, -12,
                                 -3.75, -9.25, 2.0])
    assert (jump < 12) == accept and seq[1] - seq[0] > 16  # rejected by 3v and by 4v
    o, coords, ptrs = _scene()
    _set(o, coords, ptrs, layers)
    md = march_start(146)
    t = sample_positions(md, 12)
    depth, _, hits, _ = _render(o, md, 0.64)
    assert hits == 1
    if accept:  # bracket q = 147.75 -> 148.25 (samples 3, 4), a jump at q = 148 inside
        c = regula_falsi_3(layers, 147.75, 148.25)
        assert abs(depth - (t[3] + c * (t[4] - t[3]))) < 1e-6
    else:  # bracket q = 150.25 -> 150.75 (samples 8, 9)
        c = regula_falsi_3(layers, 150.25, 150.75)
        assert abs(depth - (t[8] + c * (t[9] - t[8]))) < 1e-6


def test_invalid_sample_resets_last_weight():
    # Voxels 147 and 148 have weight 0.  Samples (mm) from q = 146.25:
    #   -0.5 (voxel 146 only), invalid, invalid, 10 (voxel 149 only), 10, 10, 10, 10, 10,
    #   0.25, 6.75, -3, -3, 0.75
    # The invalid samples reset lastWeight, so -0.5 -> 10 is no crossing; the hit is the
    # change -3 -> 0.75 between q = 152.25 and 152.75 (samples 12, 13).  Without the
    # reset, -0.5 -> 10 would be refined: each regula-falsi step moves 0.5/10.5 of the way
    # and all three stay in q < 146.5 where the field is -0.5, so that change (jump 10.5 <
    # 12) would be accepted near t0.
    seq = _sequence(RESET, 146, 14)
    assert seq[1] is None and seq[2] is None
    assert [seq[0]] + seq[3:] == pytest.approx([-0.5, 10, 10, 10, 10, 10, 10, 0.25, 6.75, -3, -3, 0.75])
    o, coords, ptrs = _scene()
    _set(o, coords, ptrs, RESET)
    md = march_start(146)
    t = sample_positions(md, 14)
    depth, _, hits, _ = _render(o, md, 0.64)
    assert hits == 1
    c = regula_falsi_3(RESET, 152.25, 152.75)
    assert abs(depth - (t[12] + c * (t[13] - t[12]))) < 1e-6


def test_positive_to_negative_change_is_ignored():
    # s = +1 mm for k <= 148, -5 mm for 149..152, +5 mm from 153.  Samples (mm) from
    # q = 146.25: 1, 1, 1, -3.5 (positive to negative: ignored), -0.5, -5 (x6), 2.5.
    # The hit is the change at samples 10 -> 11 (q = 151.25 -> 151.75), where the bracket's
    # field is -5 for u < 1/2 and 5u - 2.5 after: regula falsi 2/3, 4/7, 8/15.
    seq = _sequence(POS_NEG, 146, 12)
    assert seq == pytest.approx([1, 1, 1, -3.5, -0.5, -5, -5, -5, -5, -5, -5, 2.5])
    assert regula_falsi_3(POS_NEG, 151.25, 151.75) == pytest.approx(8 / 15)
    o, coords, ptrs = _scene()
    _set(o, coords, ptrs, POS_NEG)
    md = march_start(146)
    t = sample_positions(md, 12)
    depth, _, hits, _ = _render(o, md, 0.64)
    assert hits == 1
    assert abs(depth - (t[10] + 8 / 15 * (t[11] - t[10]))) < 1e-6


def test_range_end_is_exclusive_and_counts_a_miss():
    # The refinement field with maxDepth at the crossing sample t3: the loop runs while
    # t < end, so t3 is never sampled (a miss); one float step further, the hit.
    o, coords, ptrs = _scene()
    _set(o, coords, ptrs, REFINE, REFINE_COLORS)
    md = march_start(146)
    t = sample_positions(md, 4)
    depth, _, hits, misses = _render(o, md, float(t[3]))
    assert hits == 0 and misses == 1 and depth == 1e10
    depth, _, hits, misses = _render(o, md, float(np.nextafter(np.float32(t[3]), np.float32(1))))
    assert hits == 1 and abs(depth - (t[2] + 75 / 151 * (t[3] - t[2]))) < 1e-6
