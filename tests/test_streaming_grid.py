"""The host chunk grid of VoxelStreamingManager (VoxelStreamingManager.cpp:856-897,
993-1130) on the CPU, with a stand-in scene that holds blocks in a dict:

- one image per block coordinate per chunk, the newer replacing the older and keeping
  its place (the reference's chunkData, VoxelStreamingManager.h:78-118, appends a copy);
- a chunk larger than what one stream-in may still collect is collected in parts, and
  the chunks after it still come back (the reference returns at such a chunk,
  .cpp:1080-1085, and never collects it);
- streamIn runs collect + insert passes until the sphere's chunks are empty or nothing
  more fits; blocks the scene refuses for want of space go back to the grid;
- useParts takes one chunk per call; blocks outside the grid are dropped and counted.
"""
import numpy as np
import pytest

from vsg_amd import vsg

NB = 16  # bytes of a stand-in block image


class FakeScene:
    """Holds blocks by coordinate; `capacity` free blocks; stream_out hands out what the
    test queued."""

    def __init__(self, capacity=1 << 30):
        self.blocks = {}
        self.capacity = capacity
        self.queue = []

    def block_image_bytes(self):
        return NB

    def stream_out(self, center, radius, cap, coords_out=None, images_out=None):
        c, im = self.queue.pop(0) if self.queue else (np.zeros((0, 3), np.int32), np.zeros((0, NB), np.uint8))
        assert len(c) <= cap
        if coords_out is None:
            return c, im
        coords_out[:len(c)] = c  # like VoxelScene.stream_out: views of the caller's buffers
        images_out[:len(c)] = im
        return coords_out[:len(c)], images_out[:len(c)]

    def stream_in(self, coords, images):
        st = np.zeros(len(coords), np.uint8)
        for i, (c, im) in enumerate(zip(map(tuple, coords), images)):
            if c in self.blocks:
                st[i] = 1
            elif len(self.blocks) >= self.capacity:
                st[i] = 3
            else:
                self.blocks[c] = im.copy()
        return st


class P:
    voxelSize = 0.01  # blocks of 8 cm


def _mgr(scene, cap=100000):
    m = vsg.VoxelStreamingManager(scene, P(), voxelExtents=(1.0, 1.0, 1.0), gridDimensions=(8, 8, 8),
                                  minGridPos=(-4, -4, -4))
    m.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = cap
    return m


def _img(tag, n):
    return np.repeat(np.asarray(tag, np.uint8).reshape(-1, 1), NB, axis=1)[:n]


def _chunk_blocks(chunk_x, n, y=0):
    """n block coordinates inside chunk (chunk_x, 0, 0): 1 m chunks of 12.5 blocks, the
    chunk centred on chunk_x metres."""
    base = int(round(chunk_x / 0.08))
    return np.array([[base - 3 + (i % 6), y + i // 6 - 2, 0] for i in range(n)], np.int32)


def test_newer_image_replaces_older_in_place():
    sc = FakeScene()
    m = _mgr(sc)
    a = _chunk_blocks(0, 3)
    m.updateChunkGrid(a, _img([1, 2, 3], 3))
    m.updateChunkGrid(np.concatenate([a[1:2], _chunk_blocks(0, 4)[3:]]), _img([20, 4], 2))
    assert m.numStreamedBlocks() == 4
    assert m.streamInFindChunk((0.0, 0.0, 0.0), 5.0) == 4
    coords, images = m._in
    assert np.array_equal(coords, _chunk_blocks(0, 4))  # first-arrival order
    assert list(images[:, 0]) == [1, 20, 3, 4]
    assert m.numStreamedBlocks() == 0


def test_repeated_coordinate_within_one_batch_keeps_first_place_last_image():
    m = _mgr(FakeScene())
    a = _chunk_blocks(0, 2)
    m.updateChunkGrid(np.stack([a[0], a[1], a[0]]), _img([7, 8, 9], 3))
    assert m.numStreamedBlocks() == 2
    m.streamInFindChunk((0.0, 0.0, 0.0), 5.0)
    assert np.array_equal(m._in[0], a) and list(m._in[1][:, 0]) == [9, 8]


def test_stream_out_reintegrate_stream_out_stream_in_restores_newest():
    """Stream out, the block is re-created (integration) with new data, stream out again,
    stream in: the scene gets the newer image, and the grid held one image per block."""
    sc = FakeScene()
    m = _mgr(sc)
    a = _chunk_blocks(2, 5)
    sc.queue.append((a, _img([1] * 5, 5)))
    assert m.streamOut(sc, (0.0, 0.0, 0.0), 1.0) == 5
    sc.queue.append((a[:3], _img([2] * 3, 3)))
    assert m.streamOut(sc, (0.0, 0.0, 0.0), 1.0) == 3
    assert m.numStreamedBlocks() == 5
    st = m.streamIn(sc, (2.0, 0.0, 0.0), 1.0)
    assert np.all(st == 0) and len(sc.blocks) == 5
    assert [int(sc.blocks[tuple(c)][0]) for c in a] == [2, 2, 2, 1, 1]


def test_over_cap_chunk_is_collected_in_parts_and_later_chunks_still_come_back():
    sc = FakeScene()
    m = _mgr(sc, cap=3)
    big, small = _chunk_blocks(0, 7), _chunk_blocks(1, 2)
    m.updateChunkGrid(np.concatenate([big, small]), _img(np.arange(9), 9))
    seen = []
    for want in (3, 3, 3, 0):
        assert m.streamInFindChunk((0.5, 0.0, 0.0), 2.0) == want
        seen += [tuple(c) for c in m._in[0]]
    assert seen == [tuple(c) for c in np.concatenate([big, small])]
    assert m.numStreamedBlocks() == 0


def test_stream_in_runs_passes_until_empty():
    sc = FakeScene()
    m = _mgr(sc, cap=3)
    blocks = np.concatenate([_chunk_blocks(0, 7), _chunk_blocks(1, 2)])
    m.updateChunkGrid(blocks, _img(np.arange(9), 9))
    st = m.streamIn(sc, (0.5, 0.0, 0.0), 2.0)
    assert len(st) == 9 and np.all(st == 0) and m.numStreamedBlocks() == 0
    assert {tuple(c): int(sc.blocks[tuple(c)][0]) for c in blocks} == {tuple(c): i for i, c in enumerate(blocks)}


def test_refused_blocks_return_to_grid_and_stream_in_stops():
    sc = FakeScene(capacity=4)
    m = _mgr(sc, cap=3)
    blocks = np.concatenate([_chunk_blocks(0, 7), _chunk_blocks(1, 2)])
    m.updateChunkGrid(blocks, _img(np.arange(9), 9))
    st = m.streamIn(sc, (0.5, 0.0, 0.0), 2.0)
    assert int((st == 0).sum()) == 4 and int((st == 3).sum()) == len(st) - 4
    assert m.numStreamedBlocks() == 5 and len(sc.blocks) == 4
    sc.capacity = 100  # room again: the refused blocks come back from the grid
    left = m.streamIn(sc, (0.5, 0.0, 0.0), 2.0)
    assert np.all(left == 0) and len(left) == 5 and m.numStreamedBlocks() == 0 and len(sc.blocks) == 9


def test_use_parts_takes_one_chunk_per_call():
    m = _mgr(FakeScene())
    m.updateChunkGrid(np.concatenate([_chunk_blocks(0, 4), _chunk_blocks(1, 2)]), _img(np.arange(6), 6))
    assert m.streamInFindChunk((0.5, 0.0, 0.0), 2.0, useParts=True) == 4
    assert m.streamInFindChunk((0.5, 0.0, 0.0), 2.0, useParts=True) == 2
    assert m.streamInFindChunk((0.5, 0.0, 0.0), 2.0, useParts=True) == 0


def test_blocks_outside_grid_are_dropped_and_counted():
    m = _mgr(FakeScene())
    far = np.array([[200, 0, 0], [0, -200, 0]], np.int32)  # 16 m: beyond the 8-chunk grid
    m.updateChunkGrid(np.concatenate([far, _chunk_blocks(0, 1)]), _img([1, 2, 3], 3))
    assert m.droppedOutsideGrid == 2 and m.numStreamedBlocks() == 1


@pytest.mark.parametrize("seed", [0, 1])
def test_random_traffic_holds_no_more_images_than_distinct_blocks(seed):
    """Random stream-out batches with re-appearing coordinates: the grid never holds more
    images than distinct coordinates, and a final stream-in returns each coordinate's
    newest image."""
    rng = np.random.default_rng(seed)
    sc = FakeScene()
    m = _mgr(sc, cap=37)
    newest = {}
    for t in range(40):
        n = int(rng.integers(1, 60))
        c = rng.integers(-40, 40, size=(n, 3)).astype(np.int32)
        c = np.unique(c, axis=0)
        rng.shuffle(c)
        tag = np.full(len(c), t + 1, np.uint8)
        m.updateChunkGrid(c, _img(tag, len(c)))
        for x in map(tuple, c):
            newest[x] = t + 1
        assert m.numStreamedBlocks() == len(newest)
    st = m.streamIn(sc, (0.0, 0.0, 0.0), 10.0)
    assert np.all(st == 0) and m.numStreamedBlocks() == 0
    assert {k: int(v[0]) for k, v in sc.blocks.items()} == newest


def test_partial_take_then_put_keeps_order_and_newest():
    """A chunk taken in part keeps its remaining blocks in order; a later put replaces a
    remaining block's image in place and appends new ones after the rest (buffers grown
    by doubling, taken rows skipped, not moved)."""
    m = _mgr(FakeScene(), cap=4)
    a = _chunk_blocks(0, 10)
    m.updateChunkGrid(a, _img(np.arange(10), 10))
    assert m.streamInFindChunk((0.0, 0.0, 0.0), 5.0) == 4
    assert [int(v) for v in m._in[1][:, 0]] == [0, 1, 2, 3]
    extra = np.array([[40, 40, 0], [41, 40, 0]], np.int32)  # 3.2 m: another chunk
    m.updateChunkGrid(np.concatenate([a[5:6], extra]), _img([55, 60, 61], 3))
    m.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = 100
    assert m.streamInFindChunk((0.0, 0.0, 0.0), 5.0) == 8
    coords, images = m._in
    assert [int(v) for v in images[:6, 0]] == [4, 55, 6, 7, 8, 9]
    assert np.array_equal(coords[:6], a[4:])
    assert sorted(int(v) for v in images[6:, 0]) == [60, 61]
