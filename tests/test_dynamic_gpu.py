"""Dynamic refit path (lib/tracer/CUDAKernels.cu:456-530 and the viewer's use
of it, lib/viewer/LiveTracer.cpp:143-149, :284-294): select the leaves of the
dynamic triangles, move their vertices, refit those leaves and their
ancestors.  The result must be a valid tree whose boxes equal a full refit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dynamic_refit(bvh, orc):
    vs, fs = bvh.scene_heightfield(64, 48, 4, 0.2)
    vd, fd = bvh.scene_soup(3000, 17)
    offset = len(fs)
    v = np.concatenate([vs, vd]).astype(np.float32)
    f = np.concatenate([fs, fd + len(vs)]).astype(np.uint32)
    mesh = bvh.DeviceMesh(v, f)
    tree = bvh.lbvh_build(mesh)
    opt = bvh.Bittner13GPU(maxNumStepps=2, numChunks=4, batchPercent=5.0)
    opt.run(tree, len(f))  # an optimised tree: leaves are no longer in build order
    L = bvh.lib()
    nodes = tree.download()
    leaf_ids = np.nonzero(nodes["isLeafFlag"] != 0)[0]
    expect = leaf_ids[nodes["left"][leaf_ids] >= offset].astype(np.uint32)
    ids = bvh.DeviceBuffer(4 * len(leaf_ids))
    num = ctypes.c_uint32()
    L.selectDynamicLeafNodes(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), ids.ptr, ctypes.byref(num), offset)
    assert bvh.last_error() == ""
    assert num.value == len(fd) == len(expect)
    assert np.array_equal(ids.download(np.uint32, num.value), expect)
    # move the dynamic part and refit only it
    v2 = v.copy()
    v2[len(vs):] += np.array([0.15, 0.05, -0.1], np.float32)
    assert L.runMemcpy(mesh.raw.vert, v2.ctypes.data_as(ctypes.c_void_p), v2.nbytes, 1)
    L.refitDynamicTris(ctypes.byref(tree.ptr), mesh.raw, ids.ptr, num.value, 64)
    assert L.bvht_fixTreeFrom(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), ids.ptr, num.value)
    L.doCudaDevSync()
    assert bvh.last_error() == ""
    assert bvh.validate_device(tree, mesh) == (0, [0] * 6)
    after = tree.download()
    # a full refit from every leaf changes nothing more
    L.fixTree1(ctypes.byref(opt.mem), ctypes.byref(tree.ptr), 128)
    L.doCudaDevSync()
    assert tree.download().tobytes() == after.tobytes()
    # static leaves untouched, dynamic leaves hold the moved triangles
    static = leaf_ids[nodes["left"][leaf_ids] < offset]
    assert after[static].tobytes() == nodes[static].tobytes()
    moved = after[expect]
    tv = v2[f[moved["left"]]]
    assert np.all(moved["bbox"][:, :3] < tv.min(axis=1)) and np.all(moved["bbox"][:, 3:] > tv.max(axis=1))
