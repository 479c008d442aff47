"""The marching-cubes table (vsg_mc_table), checked by what a triangulation of the
iso-surface must satisfy, on the table alone (host code, no GPU):

* per case, the triangles use exactly the edges whose corners differ in sign, with at
  most 10 triangles;
* on random sign fields over a grid of cubes, with every shared edge's vertex one point,
  each triangle side inside the grid is shared by exactly two triangles, traversed in
  opposite directions (a closed, consistently oriented surface), and no side belongs
  to more than two;
* triangle normals point from inside corners (sdf < iso) to outside ones: for a sphere's
  signed distance sampled on a grid, every triangle's normal points away from the centre.
"""
from collections import Counter

import numpy as np

from vsg_amd import vsg

EDGE = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
CORNER = [(c & 1, (c >> 1) & 1, (c >> 2) & 1) for c in range(8)]


def _tris(T, k):
    row = [int(x) for x in T[k]]
    n = row.index(-1) if -1 in row else 32
    assert n % 3 == 0
    return [tuple(row[i:i + 3]) for i in range(0, n, 3)]


def test_cases_use_exactly_the_sign_changing_edges():
    T = vsg.mc_table()
    for k in range(256):
        inside = [(k >> c) & 1 for c in range(8)]
        changing = {e for e, (a, b) in enumerate(EDGE) if inside[a] != inside[b]}
        tris = _tris(T, k)
        used = {e for t in tris for e in t}
        assert used == changing, k
        assert len(tris) <= 10
        assert all(len(set(t)) == 3 for t in tris)
    assert _tris(T, 0) == [] and _tris(T, 255) == []


def _mesh(T, inside):
    """Triangles over a grid of cubes as triples of global edge keys (corner, axis)."""
    nx, ny, nz = (s - 1 for s in inside.shape)
    tris = []
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                k = 0
                for c, (dx, dy, dz) in enumerate(CORNER):
                    k |= int(inside[x + dx, y + dy, z + dz]) << c
                for t in _tris(T, k):
                    key = []
                    for e in t:
                        a, b = EDGE[e]
                        axis = [i for i in range(3) if CORNER[a][i] != CORNER[b][i]][0]
                        lo = CORNER[a]
                        key.append((x + lo[0], y + lo[1], z + lo[2], axis))
                    tris.append(tuple(key))
    return tris


def _on_boundary(key, shape):
    x, y, z, axis = key
    p = [x, y, z]
    for i in range(3):
        lim = shape[i] - 1
        if i == axis:
            continue
        if p[i] == 0 or p[i] == lim:
            return True
    return False


def test_random_fields_give_closed_consistently_oriented_surfaces():
    T = vsg.mc_table()
    rng = np.random.default_rng(0)
    for trial in range(30):
        inside = rng.random((7, 6, 5)) < rng.uniform(0.2, 0.8)
        tris = _mesh(T, inside)
        directed = Counter()
        for a, b, c in tris:
            for u, v in ((a, b), (b, c), (c, a)):
                directed[(u, v)] += 1
        for (u, v), n in directed.items():
            assert n == 1, "a side traversed twice in one direction"
            if _on_boundary(u, inside.shape) and _on_boundary(v, inside.shape):
                continue
            assert directed.get((v, u), 0) == 1, "an interior side without its opposite"


def test_normals_point_to_the_outside():
    T = vsg.mc_table()
    n = 12
    g = np.arange(n + 1, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    c = np.array([6.1, 5.9, 6.3])
    sdf = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - 4.2
    tris = _mesh(T, sdf < 0)
    assert len(tris) > 200

    def vert(key):
        x, y, z, axis = key
        p0 = np.array([x, y, z], float)
        p1 = p0.copy()
        p1[axis] += 1
        s0, s1 = sdf[x, y, z], sdf[tuple(p1.astype(int))]
        return p0 + (0 - s0) / (s1 - s0) * (p1 - p0)

    for t in tris:
        v0, v1, v2 = (vert(k) for k in t)
        nrm = np.cross(v1 - v0, v2 - v0)
        if np.linalg.norm(nrm) < 1e-9:
            continue
        assert np.dot(nrm, (v0 + v1 + v2) / 3 - c) > 0
