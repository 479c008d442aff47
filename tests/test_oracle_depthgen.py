"""Pins the oracle's depth map from a point cloud (generateFromPointCloud,
CustomDepthMapGenerator.cpp:81-222) and the scan-data file formats (scanData.cpp:223-459,
778-834) with cases whose answers follow without the oracle's arithmetic:

* a point placed at a pixel centre (u, v) + 0.5 lands on that pixel, whatever the
  rounding of the projection; points with z <= 0, a NaN z or outside the image land
  nowhere; of several points on one pixel the last in array order wins;
* normals / colours shorter than the points: a pixel keeps the normal / colour of the
  last point that has one;
* finite-difference normals of planes are the planes' unit normals, (0, 0, 1) for a
  fronto-parallel plane (right - centre along +x, down - centre along +y); border
  pixels and pixels without depth keep 0;
* the .bin layout is 8 + 27 n bytes; a pose file's camera->world is Transform times the
  closed-form inverse [R^T, -R^T t] of a rigid CameraRT.
"""
import numpy as np
import pytest

import oracle as O
from vsg_amd import scandata

W, H = 8, 6
FX, FY, CX, CY = 10.0, 12.0, 3.5, 2.5


def _at_pixel(u, v, z):
    """A camera-frame point that projects to the centre of pixel (u, v)."""
    return [(u + 0.5 - CX) * z / FX, (v + 0.5 - CY) * z / FY, z]


def test_points_land_on_their_pixels_and_the_last_one_wins():
    pts = np.array([_at_pixel(1, 1, 2.0), _at_pixel(4, 3, 1.5), _at_pixel(1, 1, 3.0),  # 2 overwrites 0
                    [0.1, 0.1, 0.0], [0.1, 0.1, -1.0], [0.1, 0.1, np.nan],          # z <= 0, NaN
                    _at_pixel(-1, 2, 1.0), _at_pixel(W, 2, 1.0), _at_pixel(2, H, 1.0),  # outside
                    _at_pixel(7, 5, 0.5)], np.float32)
    nrm = np.tile(np.arange(len(pts), dtype=np.float32)[:, None], (1, 3))
    col = np.tile(np.arange(len(pts), dtype=np.uint8)[:, None] * 10, (1, 3))
    r = O.depthmap_from_points(pts, nrm, col, W, H, (FX, FY, CX, CY))
    hit = {(1, 1): 2, (4, 3): 1, (7, 5): 9}
    assert r["valid"] == len(hit)
    for p in range(W * H):
        u, v = p % W, p // W
        if (u, v) in hit:
            i = hit[(u, v)]
            assert np.array_equal(r["pointmap"][p], pts[i]) and r["depth"][p] == pts[i, 2]
            assert np.all(r["normalmap"][p] == i) and np.all(r["colormap"][p] == 10 * i)
        else:
            assert not r["pointmap"][p].any() and r["depth"][p] == 0 and not r["colormap"][p].any()
            assert not r["normalmap"][p].any()


def test_short_normal_and_colour_arrays_keep_the_last_point_that_has_one():
    pts = np.array([_at_pixel(2, 2, 1.0), _at_pixel(2, 2, 1.1), _at_pixel(2, 2, 1.2), _at_pixel(5, 1, 1.0)],
                   np.float32)
    nrm = np.array([[1, 0, 0], [0, 1, 0]], np.float32)  # points 0 and 1 only
    col = np.array([[7, 7, 7]], np.uint8)                # point 0 only
    r = O.depthmap_from_points(pts, nrm, col, W, H, (FX, FY, CX, CY))
    p = 2 * W + 2
    assert r["depth"][p] == pts[2, 2]                   # point 2 wins the pixel ...
    assert np.array_equal(r["normalmap"][p], [0, 1, 0])  # ... point 1's normal stays
    assert np.array_equal(r["colormap"][p], [7, 7, 7])   # ... point 0's colour stays
    q = 1 * W + 5
    assert r["depth"][q] == 1.0 and not r["normalmap"][q].any() and not r["colormap"][q].any()


def test_orthographic_projection():
    pts = np.array([[2.0 - CX + 0.5, 4.0 - CY + 0.5, 9.0], [0.0, 0.0, -1.0]], np.float32)
    r = O.depthmap_from_points(pts, None, None, W, H, (FX, FY, CX, CY), orthographic=True)
    assert r["valid"] == 1 and r["depth"][4 * W + 2] == 9.0


def _plane_points(Wp, Hp, fx, fy, cx, cy, depth_fn):
    """One point per pixel at integer pixel coordinates on the surface z = depth_fn(u, v)."""
    u, v = np.meshgrid(np.arange(Wp, dtype=np.float64), np.arange(Hp, dtype=np.float64))
    z = depth_fn(u, v)
    # nudge inside the pixel so that floor(u) is the pixel whatever the rounding
    pts = np.stack([(u + 0.25 - cx) * z / fx, (v + 0.25 - cy) * z / fy, z], -1).reshape(-1, 3)
    return pts.astype(np.float32)


@pytest.mark.parametrize("slope", [0.0, 0.3])
def test_finite_difference_normals_of_planes(slope):
    Wp, Hp, fx, fy, cx, cy = 40, 30, 50.0, 50.0, 19.5, 14.5
    # plane z = 2 + slope * X in camera coordinates: z = 2 / (1 - slope (u - cx) / fx)
    pts = _plane_points(Wp, Hp, fx, fy, cx, cy, lambda u, v: 2.0 / (1.0 - slope * (u + 0.25 - cx) / fx))
    r = O.depthmap_from_points(pts, None, None, Wp, Hp, (fx, fy, cx, cy))
    n = r["normalmap"].reshape(Hp, Wp, 3)
    # the finite differences back-project the depths at integer pixel coordinates (not at
    # the points' own positions): the expected normal is that of the surface through
    # those back-projected points, computed below in float64 with numpy's cross product
    d = r["depth"].reshape(Hp, Wp)
    assert np.all(n[0] == 0) and np.all(n[-1] == 0) and np.all(n[:, 0] == 0) and np.all(n[:, -1] == 0)
    inner = n[1:-1, 1:-1]
    assert np.allclose(np.linalg.norm(inner, axis=-1), 1.0, atol=1e-6)
    if slope == 0.0:
        assert np.abs(inner - np.array([0.0, 0.0, 1.0])).max() < 1e-6
    else:
        # back-projected points at integer u: X = (u - cx) d / fx, Z = d; along a row the
        # surface's tangent is (dX, 0, dZ) and down a column (0, dY, 0)
        u = np.arange(1, Wp - 1, dtype=np.float64)
        X0 = (u - cx) * d[1, 1:-1] / fx
        X1 = (u + 1 - cx) * d[1, 2:] / fx
        tx = np.stack([X1 - X0, np.zeros_like(u), d[1, 2:] - d[1, 1:-1]], -1)
        ty = np.array([0.0, 1.0, 0.0])
        exp = np.cross(tx, ty)
        exp /= np.linalg.norm(exp, axis=-1, keepdims=True)
        assert np.abs(inner[0] - exp).max() < 1e-4


def test_finite_difference_normals_skip_pixels_without_depth():
    Wp, Hp, fx, fy, cx, cy = 12, 10, 20.0, 20.0, 5.5, 4.5
    pts = _plane_points(Wp, Hp, fx, fy, cx, cy, lambda u, v: np.full_like(u, 1.5))
    keep = np.ones(len(pts), bool)
    keep[4 * Wp + 6] = False  # a hole
    r = O.depthmap_from_points(pts[keep], None, None, Wp, Hp, (fx, fy, cx, cy))
    n = r["normalmap"].reshape(Hp, Wp, 3)
    assert not n[4, 6].any()
    # a pixel whose right neighbour is the hole still gets a unit normal (the reference
    # does not check its neighbours), no longer (0, 0, 1)
    assert abs(np.linalg.norm(n[4, 5]) - 1.0) < 1e-6 and np.abs(n[4, 5] - [0, 0, 1]).max() > 0.1
    assert np.abs(n[2, 2] - [0, 0, 1]).max() < 1e-6


def test_point_cloud_bin_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(57, 3)).astype(np.float32)
    nrm = rng.normal(size=(57, 3)).astype(np.float32)
    col = rng.integers(0, 256, size=(57, 3)).astype(np.uint8)
    f = str(tmp_path / "000000.bin")
    scandata.save_point_cloud_bin(f, pts, nrm, col)
    raw = open(f, "rb").read()
    assert len(raw) == 8 + 27 * 57 and int.from_bytes(raw[:8], "little") == 57
    # record 5: 3 floats, 3 floats, 3 bytes
    r5 = raw[8 + 27 * 5: 8 + 27 * 6]
    assert np.array_equal(np.frombuffer(r5[:12], "<f4"), pts[5]) and np.array_equal(np.frombuffer(r5[12:24], "<f4"),
                                                                                  nrm[5])
    assert list(r5[24:]) == list(col[5])
    p2, n2, c2 = scandata.load_point_cloud_bin(f)
    assert np.array_equal(p2, pts) and np.array_equal(n2, nrm) and np.array_equal(c2, col)


def _rigid(rx, t):
    c, s = np.cos(rx), np.sin(rx)
    M = np.eye(4)
    M[:3, :3] = [[1, 0, 0], [0, c, -s], [0, s, c]]
    M[:3, 3] = t
    return M


def test_pose_file_sections_and_camera_to_world():
    T0 = _rigid(0.3, [1.0, 2.0, 3.0])
    T45 = _rigid(-0.7, [0.5, -1.0, 0.25])
    RT = _rigid(0.2, [0.1, 0.2, -0.3])

    def rows(M, prefix):
        return "\n".join(f"{prefix}{i}: " + " ".join(f"{x:.9g}" for x in M[i]) for i in range(4))
    text = "\n".join(["# header", "Transform_0", rows(T0, "row"), "", "Transform_45 (local->world)",
                      "  ignored words", rows(T45, "r"), "CameraRT:", rows(RT, "")])
    m = scandata.parse_matrix_params(text)
    for key, M in (("transform_0", T0), ("transform_45", T45), ("localToCamera", RT)):
        assert np.allclose(m[key], M, atol=1e-6), key
    R, t = RT[:3, :3], RT[:3, 3]
    inv = np.eye(4)
    inv[:3, :3] = R.T
    inv[:3, 3] = -R.T @ t
    assert np.allclose(m["cameraToWorld0"], T0 @ inv, atol=1e-5)
    assert np.allclose(m["cameraToWorld45"], T45 @ inv, atol=1e-5)
    # identity CameraRT: camera->world is the transform itself
    m2 = scandata.parse_matrix_params("Transform_0\n" + rows(T0, "") + "\nTransform_45\n" + rows(T45, ""))
    assert np.array_equal(m2["cameraToWorld0"], m2["transform_0"])
