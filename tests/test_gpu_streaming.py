"""Streaming on the GPU (vsg_stream_out / vsg_stream_in through the C ABI) against the
oracle's restatement of VoxelStreamingManager (VoxelStreamingManager.cu:36-91,258-358,
443-719): the same blocks leave the scene with the same voxels, the scene that remains
equals the oracle's block by block (block coordinates and voxels; pointers differ by
design: the oracle reuses freed blocks from its heap, the GPU keeps pointers dense), later
integration and raycasts still match, stream-in gives the same status per block, and a
stream-out / stream-in round trip through the chunk grid of VoxelStreamingManager
restores the scene and its renders exactly."""
import numpy as np
import pytest

import oracle as O
from fusion_util import GpuRun, OracleRun, compare_raycast
from test_gpu_cell_masks import _check as check_cell_masks
from vsg_amd import synth, vsg

pytestmark = pytest.mark.gpu


def _gpu_by_coord(g: GpuRun, directional: bool = False):
    coords, _ = g.scene.export_blocks()
    n = len(coords)
    sdf, w, rgbf = g.scene.export_voxels(0, n)
    d = g.scene.export_directional(0, n) if directional else None
    out = {}
    for k, c in enumerate(map(tuple, coords)):
        sl = slice(k * 512, (k + 1) * 512)
        v = (sdf[sl], w[sl], rgbf[sl])
        if d is not None:
            v = v + (d["id"][sl], d["weight"][sl], d["dtsdf"][sl])
        out[c] = v
    return out


def _oracle_by_coord(o: O.OracleScene):
    coords, ptrs, _ = o.export_blocks()
    nb = int(ptrs.max()) + 1 if len(ptrs) else 0
    sdf, w, col, fl = o.export_voxels(nb)
    out = {}
    for c, p in zip(map(tuple, coords), ptrs):
        sl = slice(p * 512, (p + 1) * 512)
        out[c] = (sdf[sl], w[sl], col[sl], fl[sl])
    return out


def _same_state(g: GpuRun, o: OracleRun):
    gs, os_ = _gpu_by_coord(g), _oracle_by_coord(o.o)
    assert set(gs) == set(os_), f"block sets differ: {len(set(gs) ^ set(os_))} blocks"
    for c, (sdf, w, rgbf) in gs.items():
        osdf, ow, ocol, ofl = os_[c]
        assert np.array_equal(w.astype(np.int64), ow.astype(np.int64)), f"weights of block {c}"
        assert np.array_equal(sdf, osdf.astype(np.float32)), f"sdf of block {c}"
        assert np.array_equal(rgbf[:, :3], ocol) and np.array_equal(rgbf[:, 3], ofl), f"colours of block {c}"
    return len(gs)


def _radius_in_gap(coords, center, voxel, target):
    d = np.linalg.norm(np.asarray(coords, np.float64) * 8 * voxel - np.asarray(center, np.float64), axis=1)
    for r in np.arange(target, target + 0.05, 1e-4):
        if np.abs(d - r).min() > 1e-4:
            return float(r)
    raise AssertionError("no radius in a gap")


@pytest.mark.parametrize("store", ["fp32", "fp16"])
def test_stream_out_integrate_stream_in_matches_oracle(torch_cuda, store):
    wl = synth.WORKLOADS[0]
    sc = synth.SynthScene(0)
    intr = sc.intrinsics
    g = GpuRun(torch_cuda, wl, store=store)
    o = OracleRun(wl, store=store)
    for i in range(6):
        f = sc.render(i)
        g.integrate(f, intr)
        o.integrate(f, intr)
    n0 = _same_state(g, o)
    coords0, _ = g.scene.export_blocks()
    center = (np.asarray(coords0, np.float64) * 8 * wl.voxel).mean(axis=0) + np.array([0.03, -0.01, 0.02])
    d = np.linalg.norm(np.asarray(coords0, np.float64) * 8 * wl.voxel - center, axis=1)
    r = _radius_in_gap(coords0, center, wl.voxel, float(np.median(d)))
    print(f"[stream] {store}: {n0} blocks, radius {r:.4f}", flush=True)

    gc, gim = g.scene.stream_out(center.astype(np.float32), r, capacity=n0)
    oo = o.o.stream_out(center.astype(np.float32).astype(np.float64), r)
    assert {tuple(c) for c in gc} == {tuple(c) for c in oo["coords"]}
    assert 0 < len(gc) < n0
    dec = vsg.decode_block_images(gim, store)
    opos = {tuple(c): k for k, c in enumerate(oo["coords"])}
    for k, c in enumerate(map(tuple, gc)):
        j = opos[c]
        assert np.array_equal(dec["sdf"][k], oo["sdf"][j].astype(np.float32))
        assert np.array_equal(dec["weight"][k], oo["weight"][j])
        assert np.array_equal(dec["rgbf"][k], oo["rgbf"][j])
    assert _same_state(g, o) == n0 - len(gc)
    check_cell_masks(g)

    # integration re-allocates in the freed region; the raycast reads the rebuilt index
    for i in range(6, 12):
        f = sc.render(i)
        g.integrate(f, intr)
        o.integrate(f, intr)
    _same_state(g, o)
    check_cell_masks(g)
    f = sc.render(11)
    H, W = f["depth"].shape
    pix = np.arange(0, W * H, 7, dtype=np.int32)
    compare_raycast(g.raycast(f, intr), o.raycast(f, intr, pixels=pix), pixels=pix)

    # stream the removed blocks back in (GPU order) on both: same status per block
    gst = g.scene.stream_in(gc, gim)
    blocks = {"coords": gc, "sdf": dec["sdf"].astype(np.float64), "weight": dec["weight"], "rgbf": dec["rgbf"]}
    ost = o.o.stream_in(blocks)
    assert np.array_equal(gst, ost)
    assert (gst == 0).any() and (gst == 1).any()  # some were re-allocated by the later frames
    _same_state(g, o)
    check_cell_masks(g)
    compare_raycast(g.raycast(f, intr), o.raycast(f, intr, pixels=pix), pixels=pix)


def test_streaming_manager_round_trip_restores_scene_and_render(torch_cuda):
    wl = synth.WORKLOADS[0]
    sc = synth.SynthScene(0)
    intr = sc.intrinsics
    g = GpuRun(torch_cuda, wl)
    for i in range(5):
        g.integrate(sc.render(i), intr)
    f = sc.render(4)
    r0 = g.raycast(f, intr)
    before = _gpu_by_coord(g)
    mgr = vsg.VoxelStreamingManager(g.scene, g.params, voxelExtents=(0.25, 0.25, 0.25), gridDimensions=(40, 40, 40),
                                    minGridPos=(-20, -20, -20))
    mgr.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = 500  # several passes per stream-out
    cam = f["pose"][:3, 3]
    n_out = mgr.streamOut(g.scene, cam, 0.9)
    assert n_out > 500 and mgr.numStreamedBlocks() == n_out and mgr.droppedOutsideGrid == 0
    assert len(_gpu_by_coord(g)) == len(before) - n_out
    r1 = g.raycast(f, intr)
    assert not np.array_equal(r0["depth"], r1["depth"])
    mgr.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = 100000
    st = mgr.streamIn(g.scene, cam, 100.0)
    assert np.all(st == 0) and mgr.numStreamedBlocks() == 0
    after = _gpu_by_coord(g)
    assert set(after) == set(before)
    for c in before:
        for a, b in zip(before[c], after[c]):
            assert np.array_equal(a, b)
    check_cell_masks(g)
    r2 = g.raycast(f, intr)
    for k in ("depth", "pos", "col"):
        assert np.array_equal(r0[k], r2[k]), k


def test_stream_in_status_codes_match_oracle(torch_cuda):
    wl = synth.WORKLOADS[0]
    nb = 12
    coords = np.array([[i, 0, 0] for i in range(11)] + [[3, 0, 0]], np.int32)
    g = vsg.VoxelScene(vsg.Params(1, 64, wl.voxel, wl.trunc), "fp32")
    img = np.zeros((nb, g.block_image_bytes()), np.uint8)
    img[:, 4::8] = 1  # weight 1 in every voxel word
    st = g.stream_in(coords, img)
    o = O.OracleScene(1, 64, wl.voxel)
    blk = {"coords": coords, "sdf": np.zeros((nb, 512)), "weight": np.ones((nb, 512), np.uint16),
           "rgbf": np.zeros((nb, 512, 4), np.uint8)}
    assert np.array_equal(st, o.stream_in(blk))
    assert list(st) == [0] * 10 + [2, 1]
    assert list(g.stream_in(np.array([[1 << 20, 0, 0]], np.int32), img[:1])) == [4]
    g2 = vsg.VoxelScene(vsg.Params(97, 3, wl.voxel, wl.trunc), "fp32")
    c2 = np.array([[i, 2 * i, 0] for i in range(5)], np.int32)
    assert list(g2.stream_in(c2, img[:5])) == [0, 0, 0, 3, 3]
    oc, oim = g2.stream_out(np.zeros(3, np.float32), 0.0, capacity=1)
    assert len(oc) == 1
    assert list(g2.stream_in(c2[3:4], img[3:4])) == [0]
    _, w, _ = g2.export_voxels(0, 3)
    assert np.all(w == 1)


def test_directional_slots_survive_a_round_trip(torch_cuda):
    wl = synth.WORKLOADS[0]
    sc = synth.SynthScene(0)
    g = GpuRun(torch_cuda, wl, directional=True)
    for i in range(3):
        g.integrate(sc.render(i), sc.intrinsics)
    before = _gpu_by_coord(g, directional=True)
    gc, gim = g.scene.stream_out(np.zeros(3, np.float32), 0.0, capacity=len(before))
    assert len(gc) == len(before) and len(_gpu_by_coord(g)) == 0
    dec = vsg.decode_block_images(gim, "fp32", directional=True)
    k0 = tuple(gc[0])
    assert np.array_equal(dec["dir_id"][0].reshape(-1, 2), before[k0][3])
    assert np.array_equal(dec["dtsdf"][0].reshape(-1, 2), before[k0][5])
    st = g.scene.stream_in(gc, gim)
    assert np.all(st == 0)
    after = _gpu_by_coord(g, directional=True)
    assert set(after) == set(before)
    for c in before:
        for a, b in zip(before[c], after[c]):
            assert np.array_equal(a, b)


def test_room_full_size_round_trip_in_parts(torch_cuda):
    """640x480 room after 40 frames: the blocks beyond the median distance from the last
    camera leave in capacity-limited parts (every part a prefix of what remains, the parts
    together exactly the selected set) and come back in two batches; the scene and the
    last frame's render are then bit-identical to before."""
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    intr = sc.intrinsics
    g = GpuRun(torch_cuda, wl)
    for i in range(40):
        f = sc.render(i)
        g.integrate(f, intr)
    r0 = g.raycast(f, intr)
    before = _gpu_by_coord(g)
    coords = np.array(list(before.keys()), np.int32)
    cam = f["pose"][:3, 3].astype(np.float32)
    r = _radius_in_gap(coords, cam, wl.voxel, float(np.median(
        np.linalg.norm(coords.astype(np.float64) * 8 * wl.voxel - cam, axis=1))))
    d = np.linalg.norm(coords.astype(np.float64) * 8 * wl.voxel - cam.astype(np.float64), axis=1)
    want = {tuple(c) for c, dd in zip(coords, d) if dd >= r}
    print(f"[stream] room: {len(before)} blocks, {len(want)} beyond {r:.4f} m", flush=True)
    cap = len(want) // 3 + 1
    parts = []
    while True:
        pc, pim = g.scene.stream_out(cam, r, capacity=cap)
        parts.append((pc.copy(), pim.copy()))
        if len(pc) < cap:
            break
    got = [tuple(c) for pc, _ in parts for c in pc]
    assert len(got) == len(set(got)) and set(got) == want
    assert len(parts) == 3
    assert len(_gpu_by_coord(g)) == len(before) - len(want)
    allc = np.concatenate([p[0] for p in parts])
    alli = np.concatenate([p[1] for p in parts])
    h = len(allc) // 2
    assert np.all(g.scene.stream_in(allc[:h], alli[:h]) == 0)
    assert np.all(g.scene.stream_in(allc[h:], alli[h:]) == 0)
    after = _gpu_by_coord(g)
    assert set(after) == set(before)
    for c in before:
        for a, b in zip(before[c], after[c]):
            assert np.array_equal(a, b)
    check_cell_masks(g)
    r2 = g.raycast(f, intr)
    for k in ("depth", "pos", "col"):
        assert np.array_equal(r0[k], r2[k]), k
    assert (r0["hits"], r0["misses"]) == (r2["hits"], r2["misses"])


@pytest.mark.parametrize("cap", [100000, 700])
def test_chunk_grid_keeps_newest_image_after_reintegration(torch_cuda, cap):
    """Stream out beyond 1.2 m, integrate ten more frames (which re-create some of the
    streamed-out blocks with new data), stream out again from the same centre, stream
    everything in: every block comes back with its newest voxels (the re-created
    blocks' second image, the others' first), and the chunk grid held one image per
    distinct block coordinate in between.  With cap 700 the stream-in collects chunks
    in parts over several passes."""
    wl = synth.WORKLOADS[1]
    sc = synth.SynthScene(1)
    intr = sc.intrinsics
    g = GpuRun(torch_cuda, wl)
    for i in range(10):
        g.integrate(sc.render(i), intr)
    cam = sc.render(9)["pose"][:3, 3].astype(np.float32)
    state0 = _gpu_by_coord(g)
    mgr = vsg.VoxelStreamingManager(g.scene, g.params, (1.0, 1.0, 1.0), (64, 64, 64), (-32, -32, -32))
    mgr.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = cap
    n1 = mgr.streamOut(g.scene, cam, 1.2)
    out1 = set(state0) - set(_gpu_by_coord(g))
    assert n1 == len(out1) > 1000 and mgr.numStreamedBlocks() == n1
    for i in range(10, 20):
        g.integrate(sc.render(i), intr)
    state1 = _gpu_by_coord(g)
    again = out1 & set(state1)
    assert len(again) > 100  # blocks streamed out and re-created by integration
    n2 = mgr.streamOut(g.scene, cam, 1.2)
    out2 = set(state1) - set(_gpu_by_coord(g))
    assert n2 == len(out2) and again <= out2
    assert mgr.numStreamedBlocks() == len(out1 | out2)  # one image per distinct block
    st = mgr.streamIn(g.scene, cam, 100.0)
    assert np.all(st == 0) and len(st) == len(out1 | out2) and mgr.numStreamedBlocks() == 0
    final = _gpu_by_coord(g)
    expect = {**{c: state0[c] for c in out1}, **state1}
    assert set(final) == set(expect)
    for c, v in expect.items():
        for a, b in zip(v, final[c]):
            assert np.array_equal(a, b), c
    # the re-created blocks really changed, so the check above tells newest from oldest
    assert sum(not np.array_equal(state0[c][1], state1[c][1]) for c in again) > len(again) // 2
    check_cell_masks(g)
