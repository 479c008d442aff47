"""Pins the oracle's directional TSDF extension (the build's definition of the design
the reference leaves in comments, CUDAHashRef.h:34-39, globalParamasConfig.cpp:101-111;
no reference behaviour pins it) on planes with prescribed normal fields:

* bin = signed axis of the largest world-normal component (+x, -x, +y, -y, +z, -z ->
  ids 1..6), computed from the camera-frame normal rotated by the pose;
* no update below the cosine 0.8, for a zero normal, or when both slots hold other
  directions;
* dtsdf is the running mean of the main update's truncated SDF, weight capped at 100.
"""
import numpy as np

import oracle as O

EYE = np.eye(4)
V, TAU, D = 0.004, 0.02, 0.6
W, H, FX, CX, CY = 24, 18, 30.0, 11.5, 8.5


def _plane(d=D):
    u, vv = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
    z = np.full((H, W), d, np.float32)
    return np.stack([(u - np.float32(CX)) * z / np.float32(FX), (vv - np.float32(CY)) * z / np.float32(FX), z],
                    -1).astype(np.float32)


def _normals(n):
    return np.broadcast_to(np.asarray(n, np.float32), (H, W, 3)).copy()


def _scene():
    return O.OracleScene(1 << 14, 1 << 12, V, directional=True)


def _frame(o, normals, pose=EYE, pts=None):
    pts = _plane() if pts is None else pts
    return o.integrate(pts, np.full((H, W, 3), 80, np.uint8), TAU, pose, (FX, FX, CX, CY), normals=normals)


def test_single_view_fills_the_minus_z_slot_with_the_tsdf():
    o = _scene()
    _frame(o, _normals([0.0, 0.0, -2.0]))  # not unit length: normalised first
    sdf, w, _, _ = o.export_voxels()
    dr = o.export_directional(o.num_allocated())
    upd = w == 1
    assert upd.sum() > 500
    assert np.all(dr["id"][upd, 0] == 6) and np.all(dr["weight"][upd, 0] == 1)
    assert np.all(dr["id"][upd, 1] == 0) and np.all(dr["weight"][upd, 1] == 0)
    # one update: dtsdf = (0 * 0 + sdf) / 1, the same float as the main TSDF
    assert np.array_equal(dr["dtsdf"][upd, 0].astype(np.float32), sdf[upd])
    assert np.all(dr["weight"][~upd] == 0)


def test_normals_below_the_cosine_threshold_or_zero_are_not_binned():
    for n in ([0.7071068, 0.0, -0.7071068], [0.0, 0.0, 0.0], [0.6, 0.0, -0.79]):
        o = _scene()
        _frame(o, _normals(n))
        _, w, _, _ = o.export_voxels()
        dr = o.export_directional(o.num_allocated())
        assert (w == 1).sum() > 500  # the reference's TSDF is updated regardless
        assert np.all(dr["weight"] == 0) and np.all(dr["id"] == 0), n
    o = _scene()
    _frame(o, _normals([0.5, 0.0, -0.8660254]))  # cosine 0.866 to -z
    dr = o.export_directional(o.num_allocated())
    assert np.all(np.isin(dr["id"][:, 0], [0, 6])) and (dr["id"][:, 0] == 6).sum() > 500


def test_pose_rotation_selects_the_world_axis():
    # camera rotated +90 deg about y: camera -z maps to world -x; camera +y stays +y
    R = np.array([[0.0, 0.0, 1.0, 0.0], [0.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0]])
    for n, want in (([0.0, 0.0, -1.0], 2), ([0.0, 1.0, 0.0], 3), ([1.0, 0.0, 0.0], 6)):
        o = _scene()
        _frame(o, _normals(n), pose=R)
        _, w, _, _ = o.export_voxels()
        dr = o.export_directional(o.num_allocated())
        upd = w == 1
        assert upd.sum() > 500 and np.all(dr["id"][upd, 0] == want), (n, want)


def test_two_directions_fill_both_slots_and_a_third_is_dropped():
    o = _scene()
    _frame(o, _normals([0.0, 0.0, -1.0]))
    n1 = o.num_allocated()
    _, w1, _, _ = o.export_voxels(n1)
    m = w1 == 1
    a = o.export_directional(n1)
    _frame(o, _normals([1.0, 0.0, 0.0]))  # same points, +x normals
    b = o.export_directional(n1)
    assert np.all(b["id"][m, 0] == 6) and np.all(b["id"][m, 1] == 1) and np.all(b["weight"][m] == 1)
    assert np.array_equal(b["dtsdf"][m, 0], a["dtsdf"][m, 0])
    _frame(o, _normals([0.0, -1.0, 0.0]))  # -y: both slots taken by other bins
    c = o.export_directional(n1)
    for k in ("id", "weight", "dtsdf"):
        assert np.array_equal(c[k][m], b[k][m]), k
    _frame(o, _normals([0.1, 0.0, -0.99]))  # -z again (cosine 0.995): slot 0, weight 2
    d = o.export_directional(n1)
    assert np.all(d["weight"][m, 0] == 2) and np.all(d["weight"][m, 1] == 1)
    assert np.abs(d["dtsdf"][m, 0] - a["dtsdf"][m, 0]).max() < 1e-7  # same SDF twice: the mean stays


def test_weight_caps_at_100_and_the_mean_follows_new_values():
    o = _scene()
    for _ in range(105):
        _frame(o, _normals([0.0, 0.0, -1.0]))
    n1 = o.num_allocated()
    sdf, w, _, _ = o.export_voxels(n1)
    dr = o.export_directional(n1)
    m = w >= 100
    assert m.sum() > 500 and np.all(dr["weight"][m, 0] == 100)
    assert np.abs(dr["dtsdf"][m, 0] - sdf[m]).max() < 1e-6
    # a plane 4 mm further: each update moves the slot by (new - mean) / 101 at the cap
    before = dr["dtsdf"][:, 0].copy()
    _frame(o, _normals([0.0, 0.0, -1.0]), pts=_plane(D + 0.004))
    after = o.export_directional(n1)["dtsdf"][:, 0]
    moved = m & (after != before)
    assert moved.sum() > 100
    step = (after - before)[moved]
    assert np.all(np.abs(step) <= TAU * 2 / 101 + 1e-7)
