"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

CPU: the oracle still produces exactly the frozen outputs.
GPU: the device LBVH (30- and 63-bit) and the primary-ray hits of the 65,536-
triangle heightfield equal the
fixtures byte for byte, without the oracle in the loop.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def tiny():
    with np.load(os.path.join(GOLDEN, "tiny.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def c0():
    with open(os.path.join(GOLDEN, "c0.json")) as fh:
        return json.load(fh)


def test_oracle_reproduces_tiny_fixtures(orc, tiny):
    got = make_golden.tiny_fixtures(orc)
    assert sorted(got) == sorted(tiny)
    for k, v in got.items():
        assert np.array_equal(v, tiny[k]), k


def test_oracle_reproduces_c0_fixture(orc, c0):
    import bvhtest_amd  # host-only scene generator (no device call)
    assert make_golden.c0_fixture(bvhtest_amd, orc) == c0


def test_c0_fixture_invariants(c0):
    b = c0["bittner_seq"]
    assert c0["mesh"]["faces"] == 65536 and c0["mesh"]["vertices"] == 257 * 129
    assert len(b["sah_1_1"]) == b["steps"] + 1
    assert b["sah_1_1"][0] == pytest.approx(c0["lbvh30"]["sah_1_1"], rel=1e-12)
    assert b["final_sah_1.2_1"] < 0.5 * c0["lbvh30"]["sah_1.2_1"]


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [30, 63])
def test_device_lbvh_equals_tiny_fixtures(bvh, tiny, bits):
    for n in make_golden.TINY_SIZES:
        v, f = tiny[f"v{n}"], tiny[f"f{n}"]
        mesh = bvh.DeviceMesh(v, f)
        b = bvh.LBVHBuilder(mesh, bits)
        tree = b.allocate_bvh()
        b.build(tree)
        got = tree.download()
        assert got.view(np.uint8).reshape(len(got), 48).tobytes() == tiny[f"nodes{n}_{bits}"].tobytes(), n
        b.free()


@pytest.mark.gpu
def test_device_c0_equals_fixture(bvh, c0):
    v, f = bvh.scene_heightfield(256, 128, 1, 0.2)
    assert sha(v) == c0["mesh"]["sha256_vertices"] and sha(f) == c0["mesh"]["sha256_faces"]
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, 30)
    tree = b.allocate_bvh()
    keys = []
    b.build_staged(tree, keys_out=keys)
    perm = b.sorted_faces()
    assert sha(perm) == c0["lbvh30"]["sha256_perm"]
    assert sha(keys[0][perm].astype(np.uint64)) == c0["lbvh30"]["sha256_keys"]
    assert sha(tree.download()) == c0["lbvh30"]["sha256_nodes"]
    cam = c0["trace1024"]["camera"]
    rays = bvh.generate_rays(1024, 1024, cam["pos"], cam["look_at"], cam["up"], cam["fov"])
    _, hits = bvh.trace_rays(tree, mesh, rays)
    assert int((hits["triId"] != 0xFFFFFFFF).sum()) == c0["trace1024"]["hits"]
    assert sha(hits) == c0["trace1024"]["sha256_hits"]
