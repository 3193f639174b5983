"""Dynamic refit path (lib/tracer/CUDAKernels.cu:456-530 and the viewer's use
of it, lib/viewer/LiveTracer.cpp:143-149, :284-294): select the leaves of the
dynamic triangles, move their vertices, refit those leaves and their
ancestors.

Oracle: kRefitLeafBVHTris restated (oracle/oracle.cpp, orc_refit_leaf_tris:
leaf box from the moved triangle widened by FLT_EPSILON, surfaceArea kept as
the reference keeps it), then kFixTree1 (every inner node from its children).
The device's targeted refit must give the same node array byte for byte."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(bvh, seed_static, seed_dyn, n_dyn):
    vs, fs = bvh.scene_heightfield(64, 48, seed_static, 0.2)
    vd, fd = bvh.scene_soup(n_dyn, seed_dyn)
    offset = len(fs)
    v = np.concatenate([vs, vd]).astype(np.float32)
    f = np.concatenate([fs, fd + len(vs)]).astype(np.uint32)
    mesh = bvh.DeviceMesh(v, f)
    tree = bvh.lbvh_build(mesh)
    opt = bvh.Bittner13GPU(maxNumStepps=2, numChunks=4, batchPercent=5.0)
    opt.run(tree, len(f))  # an optimised tree: leaves are no longer in build order
    return v, f, len(vs), offset, mesh, tree, opt


def _select(bvh, opt, tree, nodes, offset):
    L = bvh.lib()
    leaf_ids = np.nonzero(nodes["isLeafFlag"] != 0)[0]
    ids = bvh.DeviceBuffer(4 * len(leaf_ids))
    num = ctypes.c_uint32()
    L.selectDynamicLeafNodes(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), ids.ptr, ctypes.byref(num), offset)
    assert bvh.last_error() == ""
    return ids, num.value, leaf_ids


def test_dynamic_refit(bvh, orc):
    v, f, nvs, offset, mesh, tree, opt = _setup(bvh, 4, 17, 3000)
    L = bvh.lib()
    nodes = tree.download()
    ids, num, leaf_ids = _select(bvh, opt, tree, nodes, offset)
    # selectDynamicLeafNodes: the leaves whose face index is >= offset, in leaf-list order
    expect = leaf_ids[nodes["left"][leaf_ids] >= offset].astype(np.uint32)
    assert num == len(f) - offset == len(expect)
    sel = ids.download(np.uint32, num)
    assert np.array_equal(sel, expect)
    # move the dynamic part and refit only it
    v2 = v.copy()
    v2[nvs:] += np.array([0.15, 0.05, -0.1], np.float32)
    assert L.runMemcpy(mesh.raw.vert, v2.ctypes.data_as(ctypes.c_void_p), v2.nbytes, 1)
    L.refitDynamicTris(ctypes.byref(tree.ptr), mesh.raw, ids.ptr, num, 64)
    assert L.bvht_fixTreeFrom(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), ids.ptr, num)
    L.doCudaDevSync()
    assert bvh.last_error() == ""
    after = tree.download()
    want = orc.fix_tree1(orc.refit_leaf_tris(nodes, v2, f, sel))
    assert after.tobytes() == want.tobytes()
    assert np.array_equal(after["surfaceArea"][sel], nodes["surfaceArea"][sel])
    # every Validate check but the surface-area one (which only the moved
    # leaves, keeping their old areas as in the reference, may fail) passes
    total, err = bvh.validate_device(tree, mesh)
    assert err[:3] == [0, 0, 0] and err[4:] == [0, 0]
    assert total == err[3] <= num


def test_fix_tree_from_duplicate_starts(bvh, orc):
    """A start list that names leaves more than once: each leaf's parent is
    walked to once (by the last occurrence), the counters return to zero, so
    the result equals the oracle and the optimizer's next chunks still refit
    exactly (a later full refit changes nothing)."""
    v, f, nvs, offset, mesh, tree, opt = _setup(bvh, 6, 19, 2000)
    L = bvh.lib()
    nodes = tree.download()
    ids, num, _ = _select(bvh, opt, tree, nodes, offset)
    sel = ids.download(np.uint32, num)
    rng = np.random.default_rng(3)
    starts = np.concatenate([sel, rng.choice(sel, num // 2), sel[:7]]).astype(np.uint32)
    rng.shuffle(starts)
    dstarts = bvh.DeviceBuffer(starts.nbytes).upload(starts)
    v2 = v.copy()
    v2[nvs:] -= np.array([0.05, 0.1, 0.02], np.float32)
    assert L.runMemcpy(mesh.raw.vert, v2.ctypes.data_as(ctypes.c_void_p), v2.nbytes, 1)
    L.refitDynamicTris(ctypes.byref(tree.ptr), mesh.raw, ids.ptr, num, 64)
    assert L.bvht_fixTreeFrom(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), dstarts.ptr, len(starts))
    L.doCudaDevSync()
    assert bvh.last_error() == ""
    want = orc.fix_tree1(orc.refit_leaf_tris(nodes, v2, f, sel))
    assert tree.download().tobytes() == want.tobytes()
    # the shared arrival counters are balanced: optimizer chunks refit exactly
    opt.step(tree)
    L.doCudaDevSync()
    before = tree.download()
    L.fixTree1(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), 128)
    L.doCudaDevSync()
    assert tree.download().tobytes() == before.tobytes()
    st = opt.stats()
    assert st["applied"] > 0 and st["runaway"] == 0 and st["unresolved"] == 0
