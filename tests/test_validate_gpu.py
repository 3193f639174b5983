"""Device-side Validate (bvht_validate) against the oracle's Validate
(oracle/oracle.cpp orc_validate, lib/misc/Validate.cpp:42-263).

Classes 1-5 (bbox, strict bbox, SA, triangles, orphans) have the same
per-node definitions on both sides and must match exactly; the tree class is
compared as zero / non-zero (the oracle walks the tree and also checks levels,
which the device does not maintain).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_errors(orc, nodes, v, f):
    n = nodes.copy()
    ml = orc.fix_levels(n, 0) if orc.tree_is_valid(n, 0) else 0
    _, err = orc.validate(n, 0, ml, v, f)
    return [int(x) for x in err]


def compare(bvh, orc, nodes, mesh, v, f):
    tree = bvh.DeviceBVH.from_host(nodes)
    total, dev = bvh.validate_device(tree, mesh)
    ref = oracle_errors(orc, nodes, v, f)
    assert dev[1:] == ref[1:], (dev, ref)
    assert (dev[0] == 0) == (ref[0] == 0), (dev, ref)
    return dev


@pytest.fixture(scope="module")
def spheres(bvh):
    v, f = bvh.scene_spheres_soup(24, 3000, 17)
    mesh = bvh.DeviceMesh(v, f)
    return v, f, mesh


def test_valid_trees_have_no_errors(bvh, orc, spheres):
    v, f, mesh = spheres
    tree = bvh.lbvh_build(mesh)
    assert bvh.validate_device(tree, mesh) == (0, [0] * 6)
    bvh.Bittner13GPU(maxNumStepps=3, numChunks=8, batchPercent=5.0).run(tree, len(f))
    assert bvh.validate_device(tree, mesh) == (0, [0] * 6)
    assert compare(bvh, orc, tree.download(), mesh, v, f) == [0] * 6


def test_planted_errors_are_found(bvh, orc, spheres):
    v, f, mesh = spheres
    good = bvh.lbvh_build(mesh).download()
    n = len(f)
    inner = 1234

    bad = good.copy()
    bad["bbox"][inner][0] += np.float32(0.25)  # no longer contains a child
    d = compare(bvh, orc, bad, mesh, v, f)
    assert d[1] >= 1 and d[2] >= 1

    bad = good.copy()
    bad["surfaceArea"][n + 7] *= np.float32(1.5)
    assert compare(bvh, orc, bad, mesh, v, f)[3] == 1

    bad = good.copy()
    bad["left"][n - 1 + 20] = bad["left"][n - 1 + 21]  # two leaves share a face: one orphan
    d = compare(bvh, orc, bad, mesh, v, f)
    assert d[5] == 1

    bad = good.copy()
    bad["parent"][inner] = 3  # parent link that does not point back
    assert compare(bvh, orc, bad, mesh, v, f)[0] >= 1

    bad = good.copy()
    bad["left"][inner] = inner  # self child
    assert compare(bvh, orc, bad, mesh, v, f)[0] >= 1


def test_detached_cycle_is_a_tree_error(bvh, orc, spheres):
    """Inner node a and its inner child b point at each other: the subtree
    below them is detached from the root and every node in it is reported."""
    v, f, mesh = spheres
    nodes = bvh.lbvh_build(mesh).download()
    inner = nodes["isLeafFlag"] == 0
    a = next(i for i in range(1, len(nodes)) if inner[i] and inner[nodes["left"][i]])
    b = int(nodes["left"][a])
    nodes["parent"][a] = b
    tree = bvh.DeviceBVH.from_host(nodes)
    total, d = bvh.validate_device(tree, mesh)
    assert d[0] >= 2
    assert not orc.tree_is_valid(nodes, 0)
