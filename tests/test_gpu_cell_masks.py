"""The raycaster's positive-cell summary (per block: which 4^3-voxel cells hold a
voxel with weight > 0 and SDF > 0, and how many) is exact after every integration,
for both stores, and after voxels are overwritten from the host: recomputed here in
numpy from the exported voxels.  The march skips samples with it, so a stale or
missing bit would change renders; the parity tests check the renders themselves."""
import numpy as np
import pytest

from fusion_util import GpuRun
from vsg_amd import synth

pytestmark = pytest.mark.gpu


def _expected(sdf, w, nb):
    pos = ((w > 0) & (sdf > 0)).reshape(nb, 2, 4, 2, 4, 2, 4)  # [block, cz, z, cy, y, cx, x]
    counts = pos.sum(axis=(2, 4, 6))                              # [block, cz, cy, cx]
    # cell index c = cx | cy << 1 | cz << 2
    counts = counts.transpose(0, 3, 2, 1).reshape(nb, 2, 2, 2)    # [block, cx, cy, cz]
    flat = np.zeros((nb, 8), np.int64)
    for c in range(8):
        flat[:, c] = counts[:, c & 1, (c >> 1) & 1, (c >> 2) & 1]
    masks = ((flat > 0) * (1 << np.arange(8))).sum(axis=1)
    return masks.astype(np.uint8), flat


def _check(g):
    n = g.scene.counters()["num_allocated"]
    sdf, w, _ = g.scene.export_voxels(0, n)
    masks, counts = g.scene.export_cell_masks(0, n)
    em, ec = _expected(sdf, w, n)
    assert np.array_equal(counts.astype(np.int64), ec)
    assert np.array_equal(masks, em)
    return n, int((em != 0).sum())


@pytest.mark.parametrize("store", ["fp32", "fp16"])
def test_cell_masks_track_integration_and_import(torch_cuda, store):
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    g = GpuRun(torch_cuda, wl, store=store)
    for i in range(0, 40, 4):
        g.integrate(sc.render(i), sc.intrinsics)
        n, with_pos = _check(g)
    assert n > 1000 and 0 < with_pos < n
    # overwrite blocks from the host: flip the sign of half of them, zero weights of others
    rng = np.random.default_rng(3)
    for blk in rng.choice(n, 40, replace=False):
        sdf, w, rgbf = g.scene.export_voxels(int(blk), 1)
        if blk % 2:
            sdf = -sdf
        else:
            w = np.where(rng.random(w.shape) < 0.5, 0, w).astype(np.uint16)
        g.scene.import_voxels(int(blk), sdf, w, rgbf)
    _check(g)
    # integration continues from the imported state
    g.integrate(sc.render(41), sc.intrinsics)
    _check(g)
