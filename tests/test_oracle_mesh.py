"""Pins the oracle's marching cubes (MarchingCubes.cu:313-467, with the library's own
triangle table) on stores whose SDF is overwritten with fields whose meshes are known:

* a plane z = z0 crossing inside blocks: each block meshes the 7 x 7 cubes of the layer
  holding z0 (the reference's cubes stay inside the block) as 2 triangles per cube,
  vertices on the plane, normals (0, 0, 1) towards the outside (sdf >= iso), total area
  7 x 7 v^2 per block; with cross-block cubes the 8 x 8 cubes of every block whose +x,
  +y neighbours exist;
* a sphere inside a fully allocated box: with cross-block cubes the mesh is closed and
  consistently oriented (every side of a triangle met once in each direction, vertices
  compared bit for bit), a topological sphere (V - E + F = 2), its vertices within the
  interpolation error of the sphere and its normals pointing outwards; inside blocks only
  it has holes along block borders;
* min_weight leaves out cubes with a lighter corner; max_triangles keeps the first
  triangles of the order and still counts them all.
"""
from collections import Counter

import numpy as np

import oracle as O
from vsg_amd import vsg

V = np.float32(0.004)
EYE = np.eye(4)


def _box_scene(half=0.07, zc=0.5):
    """Every block of the box [-half, half]^2 x [zc - half, zc + half] allocated (weights 0)."""
    g = np.arange(-half, half + 1e-9, float(V) / 2, dtype=np.float64)
    X, Y = np.meshgrid(g, g)
    pts = np.stack([X, Y, np.full_like(X, zc)], -1).astype(np.float32)
    o = O.OracleScene(1 << 14, 1 << 12, float(V))
    # a bucket takes one new block per pass (its lock), so repeat until nothing is added
    n = -1
    while n != o.num_allocated():
        n = o.num_allocated()
        o.alloc(pts, half, EYE)
    return o


def _centres(o):
    coords, ptrs, _ = o.export_blocks()
    loc = np.arange(512)
    l3 = np.stack([loc % 8, (loc // 8) % 8, loc // 64], -1)
    vid = (ptrs[:, None] * 512 + loc[None, :]).reshape(-1)
    c = []
    for d in range(3):
        b = np.repeat(coords[:, d:d + 1], 512, 1).reshape(-1)
        l = np.tile(l3[:, d], len(coords))
        c.append((b * 8).astype(np.float32) * V + (l.astype(np.float32) + np.float32(0.5)) * V)
    return vid, np.stack(c, -1), coords


def _set(o, vid, f):
    for i, x in zip(vid, f.astype(np.float32)):
        o.set_sdf(int(i), float(x))


def test_plane_inside_blocks():
    o = _box_scene(half=0.05)
    vid, c, coords = _centres(o)
    z0 = 0.5 + 0.3 * float(V)  # between the voxel centres of local z 3 and 4 of block 15
    _set(o, vid, c[:, 2] - np.float32(z0))
    T = vsg.mc_table()
    r = o.marching_cubes(T)
    nblk = int(np.sum(coords[:, 2] == int(np.floor(z0 / (8 * float(V))))))
    assert r["total"] == 2 * 49 * nblk and nblk > 0
    v = r["verts"]
    assert np.abs(v[..., 2] - z0).max() < 1e-6
    assert np.abs(r["nrms"] - np.array([0.0, 0.0, 1.0])).max() < 1e-4
    a = np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=-1) / 2
    assert abs(a.sum() - 49 * nblk * float(V) ** 2) < 1e-3 * 49 * nblk * float(V) ** 2
    # cross-block cubes: blocks with their +x and +y neighbours give 8 x 8 cubes, the
    # others 8 x 7 or 7 x 8 or 7 x 7
    rc = o.marching_cubes(T, cross_block=True)
    layer = {tuple(b[:2]) for b in coords if b[2] == int(np.floor(z0 / (8 * float(V))))}
    cubes = sum((8 if (x + 1, y) in layer else 7) * (8 if (x, y + 1) in layer else 7) -
                (1 if (x + 1, y) in layer and (x, y + 1) in layer and (x + 1, y + 1) not in layer else 0)
                for x, y in layer)
    assert rc["total"] == 2 * cubes


def _sphere(cross):
    o = _box_scene()
    vid, c, _ = _centres(o)
    ctr = np.array([0.0013, -0.0021, 0.5007])
    rad = 0.045
    _set(o, vid, np.linalg.norm(c.astype(np.float64) - ctr, axis=-1) - rad)
    return o, o.marching_cubes(vsg.mc_table(), cross_block=cross), ctr, rad


def _edges(verts):
    keys = [tuple(t) for t in np.ascontiguousarray(verts.astype(np.float32)).view(np.uint32).reshape(-1, 3, 3)
            .reshape(-1, 9)]
    directed = Counter()
    vset = set()
    for k in keys:
        p = (k[0:3], k[3:6], k[6:9])
        vset.update(p)
        for i in range(3):
            directed[(p[i], p[(i + 1) % 3])] += 1
    return directed, vset


def test_sphere_cross_block_mesh_is_closed_and_outward():
    o, r, ctr, rad = _sphere(True)
    n = r["total"]
    assert n > 1000 and len(r["verts"]) == n
    directed, vset = _edges(r["verts"])
    assert all(c == 1 for c in directed.values())
    assert all(directed.get((b, a), 0) == 1 for (a, b) in directed)
    E = len(directed) // 2
    assert len(vset) - E + n == 2
    v = r["verts"]
    d = np.linalg.norm(v - ctr, axis=-1) - rad
    assert np.abs(d).max() < 0.25 * float(V)
    # computeNormal zeroes a normal whose unnormalised cross product is at most 1e-6
    # long: with 4 mm voxels that is every triangle of less than 5e-7 m^2
    nrm = r["nrms"][:, 0]
    assert np.array_equal(r["nrms"][:, 1], nrm) and np.array_equal(r["nrms"][:, 2], nrm)
    cr = np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=-1)
    ok = np.linalg.norm(nrm, axis=-1) > 0.5
    assert np.all(cr[~ok] <= 1.0001e-6) and np.all(cr[ok] >= 0.9999e-6) and ok.mean() > 0.8
    assert np.abs(np.linalg.norm(nrm[ok], axis=-1) - 1).max() < 1e-6
    assert np.all(np.einsum("ij,ij->i", nrm[ok], v[ok].mean(1) - ctr) > 0)


def test_sphere_inside_blocks_has_holes_at_block_borders():
    o, r, _, _ = _sphere(False)
    _, rc, _, _ = _sphere(True)
    assert 0 < r["total"] < rc["total"]
    directed, _ = _edges(r["verts"])
    open_sides = [k for k in directed if (k[1], k[0]) not in directed]
    assert len(open_sides) > 100
    # the reference's triangles are a subset of the cross-block ones (same cubes first)
    full = {tuple(t) for t in rc["verts"].astype(np.float32).reshape(-1, 9).view(np.uint32)}
    assert all(tuple(t) in full for t in r["verts"].astype(np.float32).reshape(-1, 9).view(np.uint32))


def test_min_weight_and_triangle_cap():
    o, r, _, _ = _sphere(True)
    T = vsg.mc_table()
    assert o.marching_cubes(T, cross_block=True, min_weight=1)["total"] == 0  # all weights 0
    cap = 777
    rc = o.marching_cubes(T, cross_block=True, max_triangles=cap)
    assert rc["total"] == r["total"] and len(rc["verts"]) == cap
    assert np.array_equal(rc["verts"], r["verts"][:cap]) and np.array_equal(rc["nrms"], r["nrms"][:cap])


def test_weld_and_ply_round_trip(tmp_path):
    _, r, _, _ = _sphere(True)
    verts = r["verts"].astype(np.float32)
    v, f = vsg.weld_vertices(verts)
    assert np.array_equal(v[f], verts)
    assert len(v) - len({tuple(sorted(e)) for t in f for e in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}) + len(f) == 2
    path = str(tmp_path / "m.ply")
    vsg.write_mesh_ply(path, verts)
    v2, f2 = vsg.read_mesh_ply(path)
    assert np.array_equal(v2, v) and np.array_equal(f2, f)
