"""ctypes view of the C ABI declared in include/bvhtest_amd.h.

Struct layouts mirror the header (and so the reference's lib/base/BVH.hpp
and lib/base/Ray.hpp layouts); numpy dtypes with the same itemsize are
provided for host-side arrays of nodes and rays.
"""
import ctypes
import os

import numpy as np

c_f = ctypes.c_float
c_u8 = ctypes.c_uint8
c_u16 = ctypes.c_uint16
c_u32 = ctypes.c_uint32
c_bool = ctypes.c_bool
c_vp = ctypes.c_void_p
c_size = ctypes.c_size_t


class Vec3(ctypes.Structure):
    _fields_ = [("x", c_f), ("y", c_f), ("z", c_f)]


class AABB(ctypes.Structure):
    _fields_ = [("minMax", Vec3 * 2)]


class MeshRaw(ctypes.Structure):
    _fields_ = [("vert", c_vp), ("norm", c_vp), ("faces", c_vp),
                ("numVert", c_u32), ("numNorm", c_u32), ("numFaces", c_u32)]


class CUDAMemoryBVHPointer(ctypes.Structure):
    _fields_ = [("bvh", c_vp), ("nodes", c_vp), ("numNodes", c_u32)]


class LBVHOptions(ctypes.Structure):
    _fields_ = [("mortonBits", c_u32), ("topology", c_u32), ("tileLog2", c_u32), ("digitBits", c_u32)]


class LBVHWorkingMemory(ctypes.Structure):
    _fields_ = [("lRes", c_bool), ("opts", LBVHOptions), ("numFaces", c_u32),
                ("keys", c_vp * 2), ("values", c_vp * 2), ("sortedValues", c_vp), ("sortedKeys", c_vp),
                ("sortScratch", c_vp), ("sortScratchBytes", c_size), ("params", c_vp),
                ("atomicLocks", c_vp), ("numLocks", c_u32), ("internal", c_vp)]


class LBVHBatchWorkingMemory(ctypes.Structure):
    _fields_ = [("lRes", c_bool), ("numMeshes", c_u32), ("totalFaces", c_u32), ("totalVert", c_u32),
                ("totalNodes", c_u32), ("tileLog2", c_u32), ("radixBits", c_u32), ("internal", c_vp)]


class AlgoCFG(ctypes.Structure):
    _fields_ = [("numChunks", c_u32), ("chunkSize", c_u32), ("blockSize", c_u32),
                ("offsetAccess", c_bool), ("altFindNode", c_bool), ("altFixTree", c_bool),
                ("altSort", c_bool), ("sort", c_bool), ("localPatchCPY", c_bool),
                ("priority", c_u32), ("lanesPerSearch", c_u32)]


class GPUWorkingMemory(ctypes.Structure):
    _fields_ = [("result", c_bool), ("numLeafNodes", c_u32), ("numInnerNodes", c_u32),
                ("numPatches", c_u32), ("numSkipped", c_u32), ("numNodesToFix", c_u32),
                ("numFlags", c_u32), ("internal", c_vp)]


class CUDAPixel(ctypes.Structure):
    _fields_ = [("hit", c_u8), ("diffuse", c_u8), ("intCount", c_u16)]


NODE_DTYPE = np.dtype([("bbox", "<f4", (6,)), ("parent", "<u4"), ("left", "<u4"), ("right", "<u4"),
                       ("isLeafFlag", "<u2"), ("isLeft", "<u2"), ("level", "<u2"), ("pad", "<u2"),
                       ("surfaceArea", "<f4")])
RAY_DTYPE = np.dtype([("pos", "<f4", (3,)), ("dir", "<f4", (3,)), ("invDir", "<f4", (3,)),
                      ("sign", "u1", (3,)), ("pad", "u1"), ("px", "<u4"), ("py", "<u4")])
PIXEL_DTYPE = np.dtype([("hit", "u1"), ("diffuse", "u1"), ("intCount", "<u2")])
HIT_DTYPE = np.dtype([("triId", "<u4"), ("t", "<f4")])
assert NODE_DTYPE.itemsize == 48 and RAY_DTYPE.itemsize == 48 and PIXEL_DTYPE.itemsize == 4

P = ctypes.POINTER

# name -> (restype, argtypes); every function declared in include/bvhtest_amd.h
SIGNATURES = {
    "bvht_setStream": (None, [c_vp]),
    "bvht_getStream": (c_vp, []),
    "bvht_getLastError": (ctypes.c_char_p, []),
    "bvht_clearError": (None, []),
    "bvht_setAllocFailAfter": (None, [ctypes.c_int]),
    "bvht_version": (ctypes.c_char_p, []),
    "bvht_deviceCount": (ctypes.c_int, []),
    "runMalloc": (c_bool, [P(c_vp), c_size]),
    "runMemcpy": (c_bool, [c_vp, c_vp, c_size, ctypes.c_int]),
    "runFree": (None, [c_vp]),
    "copyBVHToGPU": (c_bool, [c_vp, c_u32, c_u32, c_u16, P(CUDAMemoryBVHPointer)]),
    "copyMeshToGPU": (c_bool, [c_vp, c_u32, c_vp, c_u32, c_vp, c_u32, P(MeshRaw)]),
    "copyBVHToHost": (c_bool, [P(CUDAMemoryBVHPointer), c_vp, c_u32, P(c_u32), P(c_u32), P(c_u16)]),
    "copyMeshToHost": (c_bool, [P(MeshRaw), c_vp, c_vp, c_vp]),
    "resetCUDA": (None, []),
    "bvht_freeBVH": (None, [P(CUDAMemoryBVHPointer)]),
    "bvht_freeMesh": (None, [P(MeshRaw)]),
    "bvht_LBVH_defaultOptions": (LBVHOptions, []),
    "LBVH_allocateBVH": (c_bool, [P(CUDAMemoryBVHPointer), P(MeshRaw)]),
    "LBVH_allocateWorkingMemory": (LBVHWorkingMemory, [P(MeshRaw)]),
    "bvht_LBVH_allocateWorkingMemoryEx": (LBVHWorkingMemory, [P(MeshRaw), P(LBVHOptions)]),
    "LBVH_initTriData": (AABB, [P(LBVHWorkingMemory), P(MeshRaw)]),
    "LBVH_calcMortonCodes": (None, [P(LBVHWorkingMemory), AABB]),
    "LBVH_sortMortonCodes": (None, [P(LBVHWorkingMemory)]),
    "LBVH_buildBVHTree": (None, [P(LBVHWorkingMemory), P(CUDAMemoryBVHPointer)]),
    "LBVH_fixAABB": (None, [P(LBVHWorkingMemory), P(CUDAMemoryBVHPointer)]),
    "LBVH_freeWorkingMemory": (None, [P(LBVHWorkingMemory)]),
    "LBVH_doCUDASyc": (None, []),
    "bvht_LBVH_build": (c_bool, [P(LBVHWorkingMemory), P(MeshRaw), P(CUDAMemoryBVHPointer)]),
    "bvht_LBVH_sortStatus": (c_u32, [P(LBVHWorkingMemory)]),
    "bvht_sortKeyIndexPairs": (c_bool, [c_vp, c_u32, c_u32, c_u32, c_vp, P(c_u32)]),
    "bvht_LBVH_setTiming": (None, [c_bool]),
    "bvht_LBVH_setTimingPassGroups": (None, [c_bool]),
    "bvht_LBVH_buildNoAlloc": (c_bool, [P(MeshRaw), P(CUDAMemoryBVHPointer), c_u32]),
    "bvht_LBVH_stageTimes": (ctypes.c_int, [P(c_f), ctypes.c_int]),
    "bvht_LBVH_stageName": (ctypes.c_char_p, [ctypes.c_int]),
    "bvht_LBVH_autotune": (ctypes.c_int, [P(LBVHWorkingMemory), P(MeshRaw), P(CUDAMemoryBVHPointer), P(c_f)]),
    "bvht_LBVH_allocateBatch": (LBVHBatchWorkingMemory, [c_u32, P(c_u32), P(c_u32), c_u32]),
    "bvht_LBVH_buildBatch": (c_bool, [P(LBVHBatchWorkingMemory), c_vp, c_vp, c_vp, c_bool]),
    "bvht_LBVH_freeBatch": (None, [P(LBVHBatchWorkingMemory)]),
    "bvht_LBVH_batchOffsets": (c_bool, [P(LBVHBatchWorkingMemory), c_u32, P(c_u32), P(c_u32), P(c_u32)]),
    "bvht_LBVH_batchStatus": (c_u32, [P(LBVHBatchWorkingMemory)]),
    "bvht_LBVH_autotuneBatch": (ctypes.c_int, [P(LBVHBatchWorkingMemory), c_vp, c_vp, c_vp, P(c_f)]),
    "allocateMemory": (GPUWorkingMemory, [P(CUDAMemoryBVHPointer), c_u32, c_u32]),
    "freeMemory": (None, [P(GPUWorkingMemory)]),
    "initData": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), c_u32]),
    "fixTree1": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), c_u32]),
    "fixTree3": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), c_u32]),
    "bn13_selectNodes": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), AlgoCFG]),
    "bn13_rmAndReinsChunk": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), AlgoCFG, c_u32]),
    "bn13_doAlgorithmStep": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), AlgoCFG]),
    "CUDAcalcSAH": (c_f, [P(CUDAMemoryBVHPointer)]),
    "doCudaDevSync": (None, []),
    "calcNumSkipped": (c_u32, [P(GPUWorkingMemory)]),
    "bvht_calcSAHWeighted": (ctypes.c_double, [P(CUDAMemoryBVHPointer), c_f, c_f]),
    "bvht_bn13_stats": (ctypes.c_int, [P(GPUWorkingMemory), P(c_u32), ctypes.c_int]),
    "bvht_validate": (c_u32, [P(CUDAMemoryBVHPointer), c_u32, MeshRaw, P(c_u32)]),
    "selectDynamicLeafNodes": (None, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), ctypes.c_void_p, P(c_u32), c_u32]),
    "refitDynamicTris": (None, [P(CUDAMemoryBVHPointer), MeshRaw, ctypes.c_void_p, c_u32, c_u32]),
    "bvht_fixTreeFrom": (ctypes.c_bool, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), ctypes.c_void_p, c_u32]),
    "bvht_mlpNumParams": (c_u32, []),
    "bvht_mlpCreate": (ctypes.c_void_p, [c_u32]),
    "bvht_mlpFree": (None, [ctypes.c_void_p]),
    "bvht_mlpSetParams": (ctypes.c_bool, [ctypes.c_void_p, P(c_f)]),
    "bvht_mlpGetParams": (ctypes.c_bool, [ctypes.c_void_p, P(c_f)]),
    "bvht_mlpPredict": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p, c_u32, ctypes.c_void_p]),
    "bvht_mlpTrain": (c_f, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_u32, c_u32, c_u32, c_f, c_u32]),
    "bvht_bn13_setModel": (ctypes.c_bool, [P(GPUWorkingMemory), ctypes.c_void_p]),
    "bvht_bn13_features": (ctypes.c_bool, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), ctypes.c_void_p]),
    "bvht_bn13_setRecording": (ctypes.c_bool, [P(GPUWorkingMemory), c_u32]),
    "bvht_bn13_records": (c_u32, [P(GPUWorkingMemory), P(ctypes.c_void_p), P(ctypes.c_void_p)]),
    "bvht_bn13_probe": (ctypes.c_bool, [P(GPUWorkingMemory), P(CUDAMemoryBVHPointer), AlgoCFG, ctypes.c_void_p, c_u32]),
    "allocateRays": (c_bool, [P(c_vp), c_u32]),
    "allocateImage": (c_bool, [P(c_vp), c_u32, c_u32]),
    "freeRays": (None, [P(c_vp)]),
    "freeImage": (None, [c_vp]),
    "generateRays": (None, [c_vp, c_u32, c_u32, Vec3, Vec3, Vec3, c_f]),
    "tracerImage": (None, [c_vp, c_vp, c_vp, c_u32, MeshRaw, Vec3, c_u32, c_u32, c_bool]),
    "copyImageToHost": (None, [c_vp, c_vp, c_u32, c_u32]),
    "tracerDoCudaSync": (None, []),
    "bvht_traceRays": (c_bool, [c_vp, c_u32, c_vp, c_u32, MeshRaw, Vec3, ctypes.c_int, c_f, c_vp, c_vp, c_vp]),
    "bvht_setRayOrdering": (None, [c_bool]),
    "bvht_makeBounceRays": (c_bool, [c_vp, c_vp, c_u32, MeshRaw, c_u32, c_vp, c_vp]),
    "bvht_sceneHeightfield": (None, [c_u32, c_u32, c_u32, c_f, c_vp, c_vp, P(c_u32), P(c_u32)]),
    "bvht_sceneSpheresSoup": (None, [c_u32, c_u32, c_u32, c_vp, c_vp, P(c_u32), P(c_u32)]),
    "bvht_sceneSoup": (None, [c_u32, c_u32, c_f, c_vp, c_vp, P(c_u32), P(c_u32)]),
    "bvht_sceneBatch": (None, [c_u32, c_u32, c_u32, c_u32, c_vp, c_vp, c_vp, c_vp, P(ctypes.c_uint64),
                               P(ctypes.c_uint64)]),
}

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libbvhtest_amd.so")


def load_library(path=LIB_PATH):
    """Loads the gfx950 library and binds every C ABI signature.

    Raises RuntimeError when the library has not been built: there is no
    CPU fallback for any entry point.
    """
    if not os.path.exists(path):
        raise RuntimeError(f"bvhtest_amd HIP library not built: {path} is missing "
                           f"(build it with make -C bvhtest_amd/csrc)")
    lib = ctypes.CDLL(path)
    missing = []
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    lib.missing_symbols = missing
    return lib
