"""The LBVH build that allocates nothing (bvht_LBVH_buildNoAlloc).

Through the C ABI, on two calls in a row (the second with a smaller mesh, so
the shared device array holds the first call's contents), its node array must
equal the one bvht_LBVH_build writes for the same mesh, byte for byte, for
30- and 63-bit keys; a mesh above the maximum is refused with an error.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_FACES = 1 << 20


def _reference_build(bvh, mesh, bits):
    builder = bvh.LBVHBuilder(mesh, bits)
    tree = builder.allocate_bvh()
    builder.build(tree)
    bvh.lib().LBVH_doCUDASyc()
    return tree.download()


def _output_bvh(bvh, mesh):
    tree = bvh.DeviceBVH()
    assert bvh.lib().LBVH_allocateBVH(ctypes.byref(tree.ptr), ctypes.byref(mesh.raw))
    return tree


def _no_alloc_build(bvh, mesh, bits):
    L = bvh.lib()
    tree = _output_bvh(bvh, mesh)
    ok = L.bvht_LBVH_buildNoAlloc(ctypes.byref(mesh.raw), ctypes.byref(tree.ptr), bits)
    L.LBVH_doCUDASyc()
    assert ok, bvh.last_error()
    assert bvh.last_error() == ""
    return tree.download()


@pytest.mark.parametrize("bits", [30, 63])
def test_no_alloc_build_equals_the_allocating_build(bvh, orc, bits):
    scenes = [bvh.scene_spheres_soup(300, 60000, 7), bvh.scene_soup(777, 8)]  # second call smaller
    for v, f in scenes:
        print(f"bits={bits} n={len(f)}", flush=True)
        mesh = bvh.DeviceMesh(v, f)
        want = _reference_build(bvh, mesh, bits)
        got = _no_alloc_build(bvh, mesh, bits)
        assert got.tobytes() == want.tobytes()
        assert got.tobytes() == orc.lbvh(v, f, bits)[2].tobytes()


def test_no_alloc_build_at_the_maximum_and_beyond(bvh, orc):
    L = bvh.lib()
    v, f = bvh.scene_soup(MAX_FACES, 9)
    mesh = bvh.DeviceMesh(v, f)
    assert _no_alloc_build(bvh, mesh, 63).tobytes() == _reference_build(bvh, mesh, 63).tobytes()
    v2, f2 = bvh.scene_soup(MAX_FACES + 1, 10)
    mesh2 = bvh.DeviceMesh(v2, f2)
    tree2 = _output_bvh(bvh, mesh2)
    L.bvht_clearError()
    assert not L.bvht_LBVH_buildNoAlloc(ctypes.byref(mesh2.raw), ctypes.byref(tree2.ptr), 30)
    assert "exceed the maximum" in bvh.last_error()


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    for name, args in (("hipStreamCreate", [ctypes.POINTER(vp)]), ("hipStreamBeginCapture", [vp, ctypes.c_int]),
                       ("hipStreamEndCapture", [vp, ctypes.POINTER(vp)]),
                       ("hipGraphInstantiate", [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]),
                       ("hipGraphLaunch", [vp, vp]), ("hipStreamSynchronize", [vp]), ("hipGraphExecDestroy", [vp]),
                       ("hipGraphDestroy", [vp]), ("hipStreamDestroy", [vp])):
        fn = getattr(hip, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    return hip


def test_no_alloc_build_under_graph_capture(bvh, orc):
    """The build captured into a HIP graph on a stream of the caller's: every
    replay, with eager builds and synchronisations in between, writes the
    allocating build's nodes (the first call on the device happens before
    the capture, as the library looks its array up there once)."""
    L, hip = bvh.lib(), _hip()
    v, f = bvh.scene_spheres_soup(64, 20000, 12)
    mesh = bvh.DeviceMesh(v, f)
    want = _reference_build(bvh, mesh, 63)
    tree = _output_bvh(bvh, mesh)
    assert L.bvht_LBVH_buildNoAlloc(ctypes.byref(mesh.raw), ctypes.byref(tree.ptr), 63)
    L.LBVH_doCUDASyc()
    stream, graph, exe = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    old = L.bvht_getStream()
    try:
        L.bvht_setStream(stream)
        assert hip.hipStreamBeginCapture(stream, 2) == 0  # hipStreamCaptureModeRelaxed
        ok = L.bvht_LBVH_buildNoAlloc(ctypes.byref(mesh.raw), ctypes.byref(tree.ptr), 63)
        assert hip.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
        assert ok, bvh.last_error()
        assert hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        L.bvht_setStream(old)
        for rep in range(3):
            zero = np.zeros(len(want), want.dtype)
            assert L.runMemcpy(tree.ptr.nodes, zero.ctypes.data_as(ctypes.c_void_p), zero.nbytes, 1)
            assert hip.hipGraphLaunch(exe, stream) == 0
            assert hip.hipStreamSynchronize(stream) == 0
            assert tree.download().tobytes() == want.tobytes(), rep
            _reference_build(bvh, mesh, 30)  # eager work between the replays
    finally:
        L.bvht_setStream(old)
        if exe.value:
            hip.hipGraphExecDestroy(exe)
        if graph.value:
            hip.hipGraphDestroy(graph)
        hip.hipStreamDestroy(stream)
