"""Pins the oracle's surface-point extraction (RayCastRender.cu:684-868) on stores whose
SDF is overwritten with a planar field f = z - z0 (weights from integrating a
fronto-parallel plane), where the points follow in closed form:

* a plane crossing inside blocks: the voxels of the two layers next to the plane whose
  x and y neighbours lie in the same block emit a point on z = z0 at their (x, y)
  centre, with normal (0, 0, 1) or 0 (at a block's x or y border the neighbour is read
  through getVoxel at the voxel centre +- v, which rounds the centre's z half away to
  the next layer, so those points carry a crossing along x or y too);
* a plane crossing a block border: the neighbour outside the block is read through
  getVoxel at the voxel centre +- v, whose half-away rounding can land one voxel
  further (the crossing then lies halfway between the centre and the plane): checked
  against a float32 numpy evaluation of that rounding;
* a voxel with |sdf| < 0.001 and no crossing emits its centre;
* max_points keeps the first points of the order and still counts them all.
"""
import numpy as np

import oracle as O

V = np.float32(0.004)
W, H, FX, CX, CY = 24, 18, 30.0, 11.5, 8.5
EYE = np.eye(4)


def _plane_scene(depth=0.6, trunc=0.05):
    u, vv = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
    z = np.full((H, W), depth, np.float32)
    pts = np.stack([(u - np.float32(CX)) * z / np.float32(FX), (vv - np.float32(CY)) * z / np.float32(FX), z], -1)
    o = O.OracleScene(1 << 14, 1 << 12, float(V))
    o.integrate(pts.astype(np.float32), np.full((H, W, 3), 80, np.uint8), trunc, EYE, (FX, FX, CX, CY))
    return o


def _voxel_table(o):
    """Per allocated voxel: flat index, integer voxel coordinates, centre (float32), weight."""
    coords, ptrs, _ = o.export_blocks()
    n = o.num_allocated()
    _, w, _, _ = o.export_voxels(n)
    loc = np.arange(512)
    lx, ly, lz = loc % 8, (loc // 8) % 8, loc // 64
    vid = (ptrs[:, None] * 512 + loc[None, :]).reshape(-1)
    vx = (coords[:, 0:1] * 8 + lx[None, :]).reshape(-1)
    vy = (coords[:, 1:2] * 8 + ly[None, :]).reshape(-1)
    vz = (coords[:, 2:3] * 8 + lz[None, :]).reshape(-1)

    def centre(vc):
        # (float)(block * 8) * v + ((float)l + 0.5f) * v, as the extraction computes it
        b, l = np.floor_divide(vc, 8), np.mod(vc, 8)
        return (b * 8).astype(np.float32) * V + (l.astype(np.float32) + np.float32(0.5)) * V

    return vid, vx, vy, vz, centre(vx), centre(vy), centre(vz), w[vid]


def _set_field(o, z0):
    vid, vx, vy, vz, cx, cy, cz, w = _voxel_table(o)
    f = (cz - np.float32(z0)).astype(np.float32)
    for i in np.flatnonzero(w > 0):
        o.set_sdf(int(vid[i]), float(f[i]))
    return vid, vx, vy, vz, cx, cy, cz, w, f


def test_plane_inside_blocks_gives_points_on_the_plane():
    o = _plane_scene()
    kz = int(np.floor(0.6 / float(V) / 8))
    z0 = np.float32((8 * kz + 3.7) * float(V))  # between local z 3 and 4 of block kz
    vid, vx, vy, vz, cx, cy, cz, w, f = _set_field(o, z0)
    r = o.extract_surface_points(-2 * float(V), 2 * float(V), 1)
    # interior emitters: weighted voxels of layers 8kz+3 and 8kz+4, x and y local index in
    # 1..6, whose z neighbour across the plane is weighted
    wmap = {(int(a), int(b), int(c)): int(ww) for a, b, c, ww in zip(vx, vy, vz, w)}
    exp = set()
    for a, b, c, ww, px, py in zip(vx, vy, vz, w, cx, cy):
        if ww == 0 or c not in (8 * kz + 3, 8 * kz + 4) or not (1 <= a % 8 <= 6 and 1 <= b % 8 <= 6):
            continue
        nb = (int(a), int(b), int(8 * kz + 4) if c == 8 * kz + 3 else int(8 * kz + 3))
        if wmap.get(nb, 0) > 0:
            exp.add((float(px), float(py), int(c)))
    inner = {(float(px), float(py)) for (px, py, c) in exp}
    sel = np.array([(float(np.float32(p[0])), float(np.float32(p[1]))) in inner for p in r["pos"]])
    assert len(exp) > 100 and sel.sum() == len(exp)
    assert np.abs(r["pos"][sel, 2] - float(z0)).max() < 1e-6
    # normal: central differences of getVoxel at z0 -+ v/2, which reads the column of
    # voxel (x + 1, y + 1) (half-away rounding of the centre's x and y): (0, 0, 1), or 0
    # where that column is not weighted in both layers (the weighted region's edge)
    up = np.abs(r["nrm"][sel] - np.array([0.0, 0.0, 1.0])).max(axis=1) < 1e-6
    zero = np.abs(r["nrm"][sel]).max(axis=1) == 0.0
    assert np.all(up | zero) and up.mean() > 0.9
    assert r["total"] >= len(exp)


def _centre_z(idx):
    return np.float32(np.float32(np.floor_divide(idx, 8) * 8) * V + (np.float32(idx % 8) + np.float32(0.5)) * V)


def _getvoxel_index(p):
    """getVoxel's voxel index of a coordinate p > 0: floor(p / v + 0.5), in float32."""
    q = np.float32(np.float32(p) / V)
    return int(np.floor(np.float32(q + np.float32(0.5))))


def test_plane_at_a_block_border_reads_the_outside_neighbour_through_getvoxel():
    o = _plane_scene()
    kz = int(np.floor(0.6 / float(V) / 8))
    z0 = np.float32((8 * kz + 7.7) * float(V))  # between local z 7 of block kz and 0 of block kz + 1
    vid, vx, vy, vz, cx, cy, cz, w, f = _set_field(o, z0)
    r = o.extract_surface_points(-2 * float(V), 2 * float(V), 1)
    i = 8 * kz + 7
    fz = {k: np.float32(_centre_z(k) - z0) for k in range(i - 3, i + 5)}
    expect = set()
    # last layer of block kz (f < 0): its +z neighbour is read at the centre + v
    j = _getvoxel_index(np.float32(_centre_z(i) + V))
    assert j in (i + 1, i + 2)
    if fz[j] > 0:
        t = np.float32(fz[i] / np.float32(fz[i] - fz[j]))
        expect.add(float(np.float32(_centre_z(i) + np.float32(V * t))))
    # first layer of block kz + 1 (f > 0): its -z neighbour is read at the centre - v
    m = _getvoxel_index(np.float32(_centre_z(i + 1) - V))
    assert m in (i, i + 1)
    if fz[m] < 0:
        t = np.float32(fz[i + 1] / np.float32(fz[i + 1] - fz[m]))
        expect.add(float(np.float32(_centre_z(i + 1) + np.float32(np.float32(-V) * t))))
    # a voxel whose neighbour read is unweighted (it lands in column (x + 1, y + 1): at the
    # weighted region's edge) finds no crossing and emits its centre when |sdf| < 0.001
    for k in (i, i + 1):
        if abs(float(fz[k])) < 0.001:
            expect.add(float(_centre_z(k)))
    assert expect
    # points of voxels with x and y local index in 1..6 (their x and y neighbours in the block)
    inner = {(float(a), float(b)) for a, b, ww, xa, ya in zip(cx, cy, w, vx, vy)
             if ww > 0 and 1 <= xa % 8 <= 6 and 1 <= ya % 8 <= 6}
    sel = np.array([(float(np.float32(p[0])), float(np.float32(p[1]))) in inner for p in r["pos"]])
    assert sel.sum() > 50
    zs = np.float32(r["pos"][sel, 2])
    assert set(np.unique(zs).astype(float)) <= expect
    if j == i + 2:  # read one voxel too far: the crossing lies halfway between centre and plane
        assert any(abs(e - float((_centre_z(i) + z0) / 2)) < 1e-6 for e in expect)


def test_centre_fallback_and_the_point_cap():
    o = _plane_scene()
    kz = int(np.floor(0.6 / float(V) / 8))
    z0 = np.float32((8 * kz + 3.7) * float(V))
    vid, vx, vy, vz, cx, cy, cz, w, f = _set_field(o, z0)
    # a weighted voxel far above the plane: sdf 0.0005 keeps its sign (no crossing)
    far = np.flatnonzero((w > 0) & (vz == 8 * kz + 9))
    k = far[len(far) // 2]
    o.set_sdf(int(vid[k]), 0.0005)
    r = o.extract_surface_points(-2 * float(V), 0.001, 1)
    hit = np.all(np.isclose(r["pos"], [cx[k], cy[k], cz[k]], atol=0, rtol=0), axis=1)
    assert hit.sum() == 1
    full = o.extract_surface_points(-2 * float(V), 2 * float(V), 1)
    capped = o.extract_surface_points(-2 * float(V), 2 * float(V), 1, max_points=7)
    assert capped["total"] == full["total"] and len(capped["pos"]) == 7
    for key in ("pos", "nrm", "col"):
        assert np.array_equal(capped[key], full[key][:7])
    # min_weight above every weight: nothing
    assert o.extract_surface_points(-2 * float(V), 2 * float(V), 256)["total"] == 0
