"""Batched (segmented) LBVH build: every segment equals the standalone build
of its mesh (bit-exact against the oracle), with and without hipGraph replay,
for every tile size the autotuner can pick."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_segments(bvh, orc, b, v, f, fc, vc, which):
    assert b.status() == 0, "a look-back wait of the batched sort gave up"
    nodes = b.nodes()
    for i in which:
        fo, vo, no = b.offsets(i)
        n = int(fc[i])
        mv = v[vo:vo + int(vc[i])]
        mf = f[fo:fo + n]
        _, _, expect, _ = orc.lbvh(mv, mf, 30)
        got = nodes[no:no + 2 * n - 1]
        assert got.tobytes() == expect.tobytes(), f"segment {i} (n={n}) differs"


@pytest.mark.parametrize("bits", [8, 10])
@pytest.mark.parametrize("tile", [8, 9, 10, 11])
def test_batch_segments_bit_exact(bvh, orc, tile, bits):
    v, f, fc, vc = bvh.scene_batch(24, 77, 1, 20000)
    print(f"tile={tile} bits={bits} meshes={len(fc)} faces={len(f)} min={fc.min()} max={fc.max()}", flush=True)
    b = bvh.BatchBuilder(v, f, fc, vc, tile_log2=tile, radix_bits=bits)
    assert b.status() == 0  # before the first build
    b.build(use_graph=False)
    check_segments(bvh, orc, b, v, f, fc, vc, range(len(fc)))


@pytest.mark.parametrize("bits", [8, 10])
def test_batch_graph_replays(bvh, orc, bits):
    v, f, fc, vc = bvh.scene_batch(40, 5, 100, 60000)
    b = bvh.BatchBuilder(v, f, fc, vc, radix_bits=bits)
    b.build(use_graph=False)
    eager = b.nodes()
    for rep in range(4):
        b.nb.upload(np.zeros(len(eager), bvh.NODE_DTYPE))  # eager work between replays
        b.build(use_graph=True)
        got = b.nodes()
        assert got.tobytes() == eager.tobytes(), f"replay {rep} differs from the eager build"
    check_segments(bvh, orc, b, v, f, fc, vc, [0, 1, 2, 17, 39])


def test_batch_autotune_and_512_meshes(bvh, orc):
    v, f, fc, vc = bvh.scene_batch(512, 1000, 1000, 500000)
    print(f"512 meshes, {len(f)} faces", flush=True)
    b = bvh.BatchBuilder(v, f, fc, vc)
    best, ms = b.autotune()
    print("autotune", best, b.mem.radixBits, ms, flush=True)
    assert 8 <= best <= 11 and b.mem.radixBits in (8, 10) and len(ms) == 8 and all(x > 0 for x in ms)
    b.build(use_graph=True)
    check_segments(bvh, orc, b, v, f, fc, vc, [0, 1, 2, 100, 255, 511])


def test_batch_single_large_mesh_matches_standalone(bvh, orc):
    """One mesh of 300,000 soup triangles plus two tiny ones: sort tiles of a
    mesh span many look-back steps; a 1-triangle mesh has one partial tile."""
    v1, f1 = bvh.scene_soup(300000, 5)
    v2, f2 = bvh.scene_soup(1, 6)
    v3, f3 = bvh.scene_soup(3841, 7)  # one sort tile plus one pair
    v = np.concatenate([v1, v2, v3]).astype(np.float32)
    f = np.concatenate([f1, f2, f3]).astype(np.uint32)
    fc = np.array([len(f1), len(f2), len(f3)], np.uint32)
    vc = np.array([len(v1), len(v2), len(v3)], np.uint32)
    for bits in (8, 10):
        b = bvh.BatchBuilder(v, f, fc, vc, radix_bits=bits)
        b.build(use_graph=False)
        check_segments(bvh, orc, b, v, f, fc, vc, [0, 1, 2])


def test_batch_wide_sort_tile_edges(bvh, orc):
    """Meshes at the edges of the 8-bit pass's sort tiles (11,264 keys): one
    tile exactly, one key more, 16 tiles (every predecessor in the first
    look-back round) and one key more (a second round), with small meshes
    between them; every segment against the oracle."""
    tile = 1024 * 11
    sizes = [tile, 1, tile + 1, 2, 16 * tile, 16 * tile + 1, 3 * tile - 1, 5]
    parts = [bvh.scene_soup(n, 40 + i) for i, n in enumerate(sizes)]
    v = np.concatenate([p[0] for p in parts]).astype(np.float32)
    f = np.concatenate([p[1] for p in parts]).astype(np.uint32)
    fc = np.array([len(p[1]) for p in parts], np.uint32)
    vc = np.array([len(p[0]) for p in parts], np.uint32)
    assert list(fc) == sizes
    b = bvh.BatchBuilder(v, f, fc, vc, radix_bits=8)
    b.build(use_graph=False)
    check_segments(bvh, orc, b, v, f, fc, vc, range(len(fc)))


def test_batch_512_meshes_same_for_both_digit_widths(bvh):
    """The 38.4M-triangle batch: the 8-bit passes (wide tiles) and the 10-bit
    passes sort every mesh to the same order, so all 77M nodes agree."""
    v, f, fc, vc = bvh.scene_batch(512, 1000, 1000, 500000)
    out = []
    for bits in (8, 10):
        b = bvh.BatchBuilder(v, f, fc, vc, radix_bits=bits)
        b.build(use_graph=False)
        assert b.status() == 0
        out.append(b.nodes())
        b.free()
    assert out[0].tobytes() == out[1].tobytes()


@pytest.mark.parametrize("bits", [8, 10])
def test_batch_vertex_box_check_and_repair(bvh, orc, bits):
    """Each mesh's box comes from all its vertices and is checked against
    the box of the vertices its triangles reference (the reference's scene
    box).  Meshes with an unreferenced vertex outside (past the max, below
    the min, many of them) take the repair; one inside, or none, does not.
    Every segment against the oracle, twice in a row (a repair does not
    leak into the next build), with and without the graph."""
    rng = np.random.default_rng(11)
    parts = []
    for i in range(8):
        v, f = bvh.scene_heightfield(48 + 8 * i, 24, 20 + i, 0.2)
        lo, hi = v.min(axis=0), v.max(axis=0)
        if i == 1:
            v = np.concatenate([v, [[5.0, 3.0, 7.0]]])
        elif i == 3:
            v, f = np.concatenate([[[-2.0, -1.0, -3.0]], v]), f + 1
        elif i == 5:
            v = np.concatenate([v, [(lo + hi) / 2]])
        elif i == 6:
            v = np.concatenate([v, rng.random((1001, 3), dtype=np.float32) * 9 - 4])
        parts.append((v.astype(np.float32), f.astype(np.uint32)))
    v = np.concatenate([p[0] for p in parts]).astype(np.float32)
    f = np.concatenate([p[1] for p in parts]).astype(np.uint32)
    fc = np.array([len(p[1]) for p in parts], np.uint32)
    vc = np.array([len(p[0]) for p in parts], np.uint32)
    b = bvh.BatchBuilder(v, f, fc, vc, radix_bits=bits)
    for graph in (False, False, True, True):
        b.nb.upload(np.zeros(int(b.mem.totalNodes), bvh.NODE_DTYPE))
        b.build(use_graph=graph)
        check_segments(bvh, orc, b, v, f, fc, vc, range(len(fc)))
