"""Batched (segmented) LBVH build: every segment equals the standalone build
of its mesh (bit-exact against the oracle), with and without hipGraph replay,
for every tile size the autotuner can pick."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_segments(bvh, orc, b, v, f, fc, vc, which):
    nodes = b.nodes()
    for i in which:
        fo, vo, no = b.offsets(i)
        n = int(fc[i])
        mv = v[vo:vo + int(vc[i])]
        mf = f[fo:fo + n]
        _, _, expect, _ = orc.lbvh(mv, mf, 30)
        got = nodes[no:no + 2 * n - 1]
        assert got.tobytes() == expect.tobytes(), f"segment {i} (n={n}) differs"


@pytest.mark.parametrize("tile", [8, 9, 10, 11])
def test_batch_segments_bit_exact(bvh, orc, tile):
    v, f, fc, vc = bvh.scene_batch(24, 77, 1, 20000)
    print(f"tile={tile} meshes={len(fc)} faces={len(f)} min={fc.min()} max={fc.max()}", flush=True)
    b = bvh.BatchBuilder(v, f, fc, vc, tile_log2=tile)
    b.build(use_graph=False)
    check_segments(bvh, orc, b, v, f, fc, vc, range(len(fc)))


def test_batch_graph_replays(bvh, orc):
    v, f, fc, vc = bvh.scene_batch(40, 5, 100, 60000)
    b = bvh.BatchBuilder(v, f, fc, vc)
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
    print("autotune", best, ms, flush=True)
    assert 8 <= best <= 11 and all(x > 0 for x in ms)
    b.build(use_graph=True)
    check_segments(bvh, orc, b, v, f, fc, vc, [0, 1, 2, 100, 255, 511])
