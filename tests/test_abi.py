"""The C ABI library loads without a GPU and exports exactly what
include/bvhtest_amd.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bvhtest_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for ref in ("LBVH_initTriData", "LBVH_calcMortonCodes", "LBVH_sortMortonCodes", "LBVH_buildBVHTree",
                "LBVH_fixAABB", "LBVH_allocateBVH", "LBVH_allocateWorkingMemory", "LBVH_freeWorkingMemory",
                "LBVH_doCUDASyc", "allocateMemory", "freeMemory", "initData", "fixTree1", "fixTree3",
                "bn13_selectNodes", "bn13_rmAndReinsChunk", "bn13_doAlgorithmStep", "CUDAcalcSAH", "doCudaDevSync",
                "calcNumSkipped", "allocateRays", "allocateImage", "freeRays", "freeImage", "generateRays",
                "tracerImage", "copyImageToHost", "tracerDoCudaSync", "copyBVHToGPU", "copyMeshToGPU",
                "copyBVHToHost", "copyMeshToHost", "resetCUDA", "runMalloc", "runMemcpy", "runFree"):
        assert ref in names, ref


def test_library_exports_every_declared_symbol():
    import bvhtest_amd
    from bvhtest_amd._abi import SIGNATURES

    lib = bvhtest_amd.lib()
    assert lib.missing_symbols == []
    names = declared_functions()
    assert sorted(SIGNATURES) == names
    out = subprocess.run(["nm", "-D", "--defined-only", bvhtest_amd.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported
    # nothing else leaks out of the library (internal C++ symbols are hidden)
    assert exported - set(names) <= {"_init", "_fini"}


def test_struct_layouts_match_the_reference():
    from bvhtest_amd import _abi

    assert ctypes.sizeof(_abi.AABB) == 24
    assert ctypes.sizeof(_abi.MeshRaw) == 40
    assert _abi.NODE_DTYPE.itemsize == 48 and _abi.RAY_DTYPE.itemsize == 48
    off = {n: _abi.NODE_DTYPE.fields[n][1] for n in _abi.NODE_DTYPE.names}
    assert off == {"bbox": 0, "parent": 24, "left": 28, "right": 32, "isLeafFlag": 36, "isLeft": 38, "level": 40,
                   "pad": 42, "surfaceArea": 44}


def test_host_only_calls_work_without_gpu():
    import bvhtest_amd as bt

    L = bt.lib()
    assert L.bvht_version().startswith(b"bvhtest_amd")
    v, f = bt.scene_heightfield(256, 128, 1, 0.2)
    assert v.shape == (257 * 129, 3) and f.shape == (65536, 3)
    assert f.max() < len(v)
    # +y normals
    tv = v[f].astype(np.float64)
    n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    assert np.all(n[:, 1] > 0)
    assert v[:, 1].min() >= 0 and v[:, 1].max() <= 0.2
    v, f = bt.scene_spheres_soup(2048, 500000, 2)
    assert 3.9e6 < len(f) < 4.5e6 and f.max() < len(v)
    v2, f2 = bt.scene_spheres_soup(2048, 500000, 2)
    assert np.array_equal(v, v2) and np.array_equal(f, f2)
