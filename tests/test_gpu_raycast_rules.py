"""The designed fields of test_oracle_raycast_rules.py (refinement with a kink, the 3v
acceptance threshold on both sides, the lastWeight reset, an ignored positive-to-
negative change) imported into the HIP scene through the C ABI and rendered over the
whole image: depth, positions, colours and hit/miss counts bit-identical to the
oracle, at the default step budget, with the tail pass taking every ray after 4
steps, and with the tail pass off."""
import numpy as np
import pytest

import oracle as O
import test_oracle_raycast_rules as R
from vsg_amd import vsg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(R.SCENARIOS))
def test_designed_fields_match_oracle(torch_cuda, name):
    layers, colors = R.SCENARIOS[name]
    t = torch_cuda
    pts = R.plane_frame()
    rgb = np.full((R.H, R.W, 3), 90, np.uint8)
    o = O.OracleScene(1 << 16, 1 << 14, float(R.V))
    scene = vsg.VoxelScene(vsg.Params(1 << 16, 1 << 14, float(R.V), 0.02), "fp32")
    scene.set_stream(t.cuda.current_stream().cuda_stream)
    dp, dc = t.from_numpy(pts).cuda(), t.from_numpy(rgb).cuda()
    dn = t.zeros_like(dp)
    for _ in range(4):
        o.integrate(pts, rgb, 0.02, R.EYE, R.INTR)
        scene.integrateFromScanData(dp, dc, dn, R.W, R.H, 0.02, R.EYE, *R.INTR)
    t.cuda.synchronize()
    ocoords, optrs, _ = o.export_blocks()
    gcoords, _ = scene.export_blocks()
    assert np.array_equal(gcoords, ocoords) and np.array_equal(optrs, np.arange(len(optrs)))

    sdf, w, rgb3 = R.designed_store(ocoords, layers, colors)
    R._set(o, ocoords, optrs, layers, colors)
    _, _, rgbf = scene.export_voxels(0, len(gcoords))
    rgbf[:, :3] = rgb3
    scene.import_voxels(0, sdf.astype(np.float32), w.astype(np.uint16), rgbf)

    md = R.march_start(146)
    orr = o.render(R.W, R.H, [0, 0, 0], R.EYE, R.INTR, md, 0.64)
    assert orr["hits"] > 0
    rr = vsg.RayCastRender()
    assert rr.initialize(R.W, R.H)
    for budget in (None, 4, 0):
        if budget is not None:
            rr.set_step_budget(budget)
        rr.render(scene, [0, 0, 0], R.EYE, *R.INTR, md, 0.64)
        g = rr.downloadAll()
        assert (rr.counters()) == (orr["hits"], orr["misses"]), budget
        assert np.array_equal(g["depth"].view(np.uint32), orr["depth"].view(np.uint32)), budget
        assert np.array_equal(g["pos"].view(np.uint32), orr["pos"].view(np.uint32)), budget
        assert np.array_equal(g["col"].view(np.uint32), orr["col"].view(np.uint32)), budget
    print(f"{name}: {orr['hits']} hits, {orr['misses']} misses, axis depth {orr['depth'][R.AXIS]:.7f}")
