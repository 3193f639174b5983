"""Host-side Python mirror of BVHTest's GPU pipeline commands.

Thin wrappers over the C ABI (include/bvhtest_amd.h): every call goes to the
gfx950 library bvhtest_amd/lib/libbvhtest_amd.so; there is no CPU fallback.
The class and method names follow the reference's command layer:

  copyToGPU   -> Mesh.to_gpu / DeviceBVH.from_host   (lib/cuda/CopyToGPU.cpp:30-34)
  LBVH        -> lbvh_build                            (lib/builder/LBVH.cpp:28-55)
  bittner13GPU-> Bittner13GPU                          (lib/builder/Bittner13GPU.cpp:671-726)
  CUDATracer  -> generate_rays / trace_image           (lib/tracer/CUDATracer.cpp:75-150)
  copyToHost  -> DeviceBVH.to_host                     (lib/cuda/CopyToHost.cpp:30-36)
"""
import ctypes

import numpy as np

from ._abi import (AABB, HIT_DTYPE, LIB_PATH, NODE_DTYPE, PIXEL_DTYPE, RAY_DTYPE, SIGNATURES, AlgoCFG,
                   CUDAMemoryBVHPointer, GPUWorkingMemory, LBVHOptions, LBVHWorkingMemory, MeshRaw, Vec3,
                   load_library)

__all__ = ["lib", "Mesh", "DeviceMesh", "DeviceBVH", "lbvh_build", "NODE_DTYPE", "RAY_DTYPE", "PIXEL_DTYPE",
           "HIT_DTYPE", "BVHTestError", "scene_heightfield", "scene_spheres_soup", "scene_soup"]

_LIB = None


class BVHTestError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def last_error():
    return lib().bvht_getLastError().decode()


def _fail(what):
    msg = last_error()
    lib().bvht_clearError()
    raise BVHTestError(f"{what} failed: {msg}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------------------
# synthetic scenes (host, library generators)
# ---------------------------------------------------------------------------

def _scene(fn, *args):
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    fn(*args, None, None, ctypes.byref(nv), ctypes.byref(nf))
    v = np.empty((nv.value, 3), np.float32)
    f = np.empty((nf.value, 3), np.uint32)
    fn(*args, _ptr(v), _ptr(f), ctypes.byref(nv), ctypes.byref(nf))
    return v, f


def scene_heightfield(gx=256, gz=128, seed=1, amplitude=0.2):
    """Heightfield of gx x gz cells, 2 triangles per cell (the 65,536-triangle scene at 256 x 128)."""
    return _scene(lib().bvht_sceneHeightfield, gx, gz, seed, amplitude)


def scene_spheres_soup(num_spheres=2048, num_soup=500000, seed=2):
    """Icosphere clusters plus triangle soup (the ~4M-triangle scene at the defaults)."""
    return _scene(lib().bvht_sceneSpheresSoup, num_spheres, num_soup, seed)


def scene_soup(num_tris, seed=3, size=0.01):
    return _scene(lib().bvht_sceneSoup, num_tris, seed, size)


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------

class DeviceMesh:
    """MeshRaw on the device (copyMeshToGPU, lib/cuda/cudaFN.cu:76-109)."""

    def __init__(self, vert, faces, norm=None):
        vert = np.ascontiguousarray(vert, np.float32).reshape(-1, 3)
        faces = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        norm = np.zeros((0, 3), np.float32) if norm is None else np.ascontiguousarray(norm, np.float32)
        self.raw = MeshRaw()
        if not lib().copyMeshToGPU(_ptr(vert), len(vert), _ptr(norm), len(norm), _ptr(faces), len(faces),
                                   ctypes.byref(self.raw)):
            _fail("copyMeshToGPU")
        self.num_faces = len(faces)
        self.num_vert = len(vert)

    def free(self):
        if self.raw.vert:
            lib().bvht_freeMesh(ctypes.byref(self.raw))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBVH:
    """CUDAMemoryBVHPointer: device header + nodes."""

    def __init__(self):
        self.ptr = CUDAMemoryBVHPointer()

    @classmethod
    def from_host(cls, nodes, root=0, max_level=0):
        b = cls()
        nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
        if not lib().copyBVHToGPU(_ptr(nodes), len(nodes), root, max_level, ctypes.byref(b.ptr)):
            _fail("copyBVHToGPU")
        return b

    @property
    def num_nodes(self):
        return self.ptr.numNodes

    def download(self):
        """Copies the nodes to the host without freeing the device BVH."""
        out = np.empty(self.ptr.numNodes, NODE_DTYPE)
        if self.ptr.numNodes and not lib().runMemcpy(_ptr(out), self.ptr.nodes, out.nbytes, 2):
            _fail("runMemcpy")
        return out

    def to_host(self):
        """copyBVHToHost: nodes + (root, maxLevel) from the header; frees the device BVH."""
        out = np.empty(self.ptr.numNodes, NODE_DTYPE)
        n, root, ml = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint16()
        if not lib().copyBVHToHost(ctypes.byref(self.ptr), _ptr(out), len(out), ctypes.byref(n), ctypes.byref(root),
                                   ctypes.byref(ml)):
            _fail("copyBVHToHost")
        return out[:n.value], root.value, ml.value

    def free(self):
        if self.ptr.nodes:
            lib().bvht_freeBVH(ctypes.byref(self.ptr))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_u32(dev_ptr, count):
    out = np.empty(count, np.uint32)
    if count and not lib().runMemcpy(_ptr(out), dev_ptr, out.nbytes, 2):
        _fail("runMemcpy")
    return out


def download_u64(dev_ptr, count):
    out = np.empty(count, np.uint64)
    if count and not lib().runMemcpy(_ptr(out), dev_ptr, out.nbytes, 2):
        _fail("runMemcpy")
    return out


class LBVHBuilder:
    """Working memory of the LBVH command (lib/builder/LBVH.cpp:28-55)."""

    def __init__(self, mesh, morton_bits=30, topology=0, digit_bits=0, tile_log2=0):
        self.mesh = mesh
        opts = lib().bvht_LBVH_defaultOptions()
        opts.mortonBits = morton_bits
        opts.topology = topology
        opts.digitBits = digit_bits
        if tile_log2:
            opts.tileLog2 = tile_log2
        self.mem = lib().bvht_LBVH_allocateWorkingMemoryEx(ctypes.byref(mesh.raw), ctypes.byref(opts))
        if not self.mem.lRes:
            _fail("LBVH_allocateWorkingMemory")
        self.morton_bits = morton_bits

    def allocate_bvh(self):
        b = DeviceBVH()
        if not lib().LBVH_allocateBVH(ctypes.byref(b.ptr), ctypes.byref(self.mesh.raw)):
            _fail("LBVH_allocateBVH")
        return b

    def build(self, bvh):
        if not lib().bvht_LBVH_build(ctypes.byref(self.mem), ctypes.byref(self.mesh.raw), ctypes.byref(bvh.ptr)):
            _fail("bvht_LBVH_build")

    def build_staged(self, bvh, keys_out=None):
        """The five reference stages; keys_out (list) receives the unsorted Morton keys."""
        L = lib()
        box = L.LBVH_initTriData(ctypes.byref(self.mem), ctypes.byref(self.mesh.raw))
        self._check("LBVH_initTriData")
        L.LBVH_calcMortonCodes(ctypes.byref(self.mem), box)
        self._check("LBVH_calcMortonCodes")
        if keys_out is not None:
            L.LBVH_doCUDASyc()
            dl = download_u64 if self.morton_bits == 63 else download_u32
            keys_out.append(dl(self.mem.keys[0], self.mesh.num_faces))
        L.LBVH_sortMortonCodes(ctypes.byref(self.mem))
        self._check("LBVH_sortMortonCodes")
        L.LBVH_buildBVHTree(ctypes.byref(self.mem), ctypes.byref(bvh.ptr))
        self._check("LBVH_buildBVHTree")
        L.LBVH_fixAABB(ctypes.byref(self.mem), ctypes.byref(bvh.ptr))
        self._check("LBVH_fixAABB")
        L.LBVH_doCUDASyc()
        self._check("LBVH_doCUDASyc")
        return np.array([[box.minMax[0].x, box.minMax[0].y, box.minMax[0].z],
                         [box.minMax[1].x, box.minMax[1].y, box.minMax[1].z]], np.float32)

    def sort_status(self):
        """Error bits of the last sort (0: completed)."""
        return lib().bvht_LBVH_sortStatus(ctypes.byref(self.mem))

    def sorted_faces(self):
        lib().LBVH_doCUDASyc()
        return download_u32(self.mem.sortedValues, self.mesh.num_faces)

    def sorted_keys(self):
        if not self.mem.sortedKeys:
            return None
        lib().LBVH_doCUDASyc()
        dl = download_u64 if self.morton_bits == 63 else download_u32
        return dl(self.mem.sortedKeys, self.mesh.num_faces)

    @staticmethod
    def _check(what):
        if last_error():
            _fail(what)

    def free(self):
        if self.mem.lRes:
            lib().LBVH_freeWorkingMemory(ctypes.byref(self.mem))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def lbvh_build(mesh, morton_bits=30, topology=0):
    """LBVH command: allocate, build (fused, on-stream), free working memory."""
    b = LBVHBuilder(mesh, morton_bits, topology)
    bvh = b.allocate_bvh()
    b.build(bvh)
    lib().LBVH_doCUDASyc()
    b.free()
    return bvh


def vec3(x):
    return Vec3(float(x[0]), float(x[1]), float(x[2]))


# ---------------------------------------------------------------------------
# tracer (CUDATracer command, lib/tracer/CUDATracer.cpp:75-150)
# ---------------------------------------------------------------------------

def sort_key_index_pairs(keys, key_bits, digit_bits=0):
    """(indices, sorted keys) of the LBVH build's sort on host keys (uint32
    for 30 bits, uint64 for 63)."""
    keys = np.ascontiguousarray(keys, np.uint64 if key_bits == 63 else np.uint32)
    n = len(keys)
    kin = DeviceBuffer(keys.nbytes).upload(keys)
    kout = DeviceBuffer(keys.nbytes)
    vout = DeviceBuffer(4 * n)
    if not lib().bvht_sortKeyIndexPairs(kin.ptr, n, key_bits, digit_bits, kout.ptr,
                                        ctypes.cast(vout.ptr, ctypes.POINTER(ctypes.c_uint32))):
        _fail("bvht_sortKeyIndexPairs")
    return vout.download(np.uint32, n), kout.download(keys.dtype, n)


class DeviceBuffer:
    """Plain device allocation (runMalloc / runFree)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        if not lib().runMalloc(ctypes.byref(self.ptr), max(nbytes, 16)):
            _fail("runMalloc")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes and not lib().runMemcpy(self.ptr, _ptr(arr), arr.nbytes, 1):
            _fail("runMemcpy")
        return self

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes and not lib().runMemcpy(_ptr(out), self.ptr, out.nbytes, 2):
            _fail("runMemcpy")
        return out

    def free(self):
        if self.ptr:
            lib().runFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_rays_device(w, h, pos, look_at, up, fov):
    buf = DeviceBuffer(w * h * RAY_DTYPE.itemsize)
    lib().generateRays(buf.ptr, w, h, vec3(pos), vec3(look_at), vec3(up), float(fov))
    if last_error():
        _fail("generateRays")
    return buf


def generate_rays(w, h, pos, look_at, up, fov):
    buf = generate_rays_device(w, h, pos, look_at, up, fov)
    return buf.download(RAY_DTYPE, w * h)


def trace_rays(tree, mesh, rays, light=(2.0, 2.0, 2.0), mode=0, tmin=0.0, skip=None, root=0):
    """bvht_traceRays on host ray arrays; returns (pixels, hits)."""
    n = len(rays)
    rb = DeviceBuffer(n * RAY_DTYPE.itemsize).upload(np.ascontiguousarray(rays, RAY_DTYPE))
    hb = DeviceBuffer(n * HIT_DTYPE.itemsize)
    pb = DeviceBuffer(n * PIXEL_DTYPE.itemsize)
    sb = DeviceBuffer(n * 4).upload(np.ascontiguousarray(skip, np.uint32)) if skip is not None else None
    ok = lib().bvht_traceRays(rb.ptr, n, tree.ptr.nodes, root, mesh.raw, vec3(light), mode, tmin,
                              sb.ptr if sb else None, hb.ptr, pb.ptr)
    if not ok:
        _fail("bvht_traceRays")
    lib().tracerDoCudaSync()
    return pb.download(PIXEL_DTYPE, n), hb.download(HIT_DTYPE, n)


def trace_image(tree, mesh, w, h, pos, look_at, up, fov, light=(2.0, 2.0, 2.0), root=0, bundle=False):
    """generateRays + tracerImage + copyImageToHost: the CUDATracer command
    (bundle: the 8x8 ray-bundle traversal, option rayBundles)."""
    rays = generate_rays_device(w, h, pos, look_at, up, fov)
    img = DeviceBuffer(w * h * PIXEL_DTYPE.itemsize)
    lib().tracerImage(rays.ptr, img.ptr, tree.ptr.nodes, root, mesh.raw, vec3(light), w, h, bool(bundle))
    if last_error():
        _fail("tracerImage")
    out = np.empty(w * h, PIXEL_DTYPE)
    lib().copyImageToHost(_ptr(out), img.ptr, w, h)
    return out


def bounce_rays(mesh, rays, hits, seed=1):
    n = len(rays)
    rb = DeviceBuffer(n * RAY_DTYPE.itemsize).upload(np.ascontiguousarray(rays, RAY_DTYPE))
    hb = DeviceBuffer(n * HIT_DTYPE.itemsize).upload(np.ascontiguousarray(hits, HIT_DTYPE))
    ob = DeviceBuffer(n * RAY_DTYPE.itemsize)
    sb = DeviceBuffer(n * 4)
    if not lib().bvht_makeBounceRays(rb.ptr, hb.ptr, n, mesh.raw, seed, ob.ptr, sb.ptr):
        _fail("bvht_makeBounceRays")
    lib().tracerDoCudaSync()
    return ob.download(RAY_DTYPE, n), sb.download(np.uint32, n)


# ---------------------------------------------------------------------------
# Bittner13 GPU optimizer (lib/builder/Bittner13GPU.cpp:46-133)
# ---------------------------------------------------------------------------

class Bittner13GPU:
    """bittner13GPU command: options and defaults of lib/builder/Bittner13GPU.hpp:28-37."""

    def __init__(self, maxNumStepps=500, numChunks=32, CUDABlockSize=128, batchPercent=1.0, sort=True,
                 offsetAccess=False, altFindNode=False, altFixTree=False, altSort=False, localPatchCPY=False,
                 priority=0):
        batchPercent = min(max(batchPercent, 0.01), 75.0)  # Bittner13GPU.cpp:652-653
        self.opts = dict(maxNumStepps=maxNumStepps, numChunks=numChunks, CUDABlockSize=CUDABlockSize,
                         batchPercent=batchPercent, sort=sort, offsetAccess=offsetAccess, altFindNode=altFindNode,
                         altFixTree=altFixTree, altSort=altSort, localPatchCPY=localPatchCPY, priority=priority)
        self.mem = None
        self.model = None

    def _sizes(self, num_nodes):
        # (uint32_t)((batchPercent / 100.0f) * (float)numNodes), chunk = that / numChunks
        bp = np.float32(self.opts["batchPercent"])
        num = int(np.float32(bp / np.float32(100.0)) * np.float32(num_nodes))
        chunk = num // self.opts["numChunks"]
        return chunk

    def cfg(self, chunk):
        c = AlgoCFG()
        c.numChunks = self.opts["numChunks"]
        c.chunkSize = chunk
        c.blockSize = self.opts["CUDABlockSize"]
        for k in ("offsetAccess", "altFindNode", "altFixTree", "altSort", "sort", "localPatchCPY"):
            setattr(c, k, self.opts[k])
        c.priority = self.opts["priority"]
        return c

    def set_model(self, model):
        """Priority model for priority=1 steps (None clears)."""
        self.model = model
        if not lib().bvht_bn13_setModel(ctypes.byref(self.mem), model.handle if model else None):
            _fail("bvht_bn13_setModel")

    def set_priority(self, priority):
        self.opts["priority"] = priority

    def features(self, tree):
        """(numInner, 12) features of every inner node, inner-list order."""
        n = self.mem.numInnerNodes
        buf = DeviceBuffer(n * 48)
        if not lib().bvht_bn13_features(ctypes.byref(self.mem), ctypes.byref(tree.ptr), buf.ptr):
            _fail("bvht_bn13_features")
        return buf.download(np.float32, n * 12).reshape(n, 12)

    def record(self, max_records):
        """Start (or clear) SAH-gain recording."""
        if not lib().bvht_bn13_setRecording(ctypes.byref(self.mem), max_records):
            _fail("bvht_bn13_setRecording")

    def records(self):
        """(features (n, 12), gains (n,)) recorded so far."""
        pf, pg = ctypes.c_void_p(), ctypes.c_void_p()
        n = lib().bvht_bn13_records(ctypes.byref(self.mem), ctypes.byref(pf), ctypes.byref(pg))
        feat = np.empty((n, 12), np.float32)
        gain = np.empty(n, np.float32)
        if n:
            if not lib().runMemcpy(_ptr(feat), pf, feat.nbytes, 2) or not lib().runMemcpy(_ptr(gain), pg, gain.nbytes, 2):
                _fail("runMemcpy")
        return feat, gain

    def probe(self, tree, positions):
        """Record the gains of the inner nodes at `positions` (inner-list order) without changing the tree."""
        pos = np.ascontiguousarray(positions, np.uint32)
        buf = DeviceBuffer(max(pos.nbytes, 4)).upload(pos)
        if not lib().bvht_bn13_probe(ctypes.byref(self.mem), ctypes.byref(tree.ptr), self.cfg(self.chunk), buf.ptr,
                                     len(pos)):
            _fail("bvht_bn13_probe")
        lib().doCudaDevSync()

    def setup(self, tree, num_faces):
        chunk = self._sizes(tree.num_nodes)
        if chunk == 0:
            raise BVHTestError("bittner13GPU: batch too small for the number of chunks")
        self.chunk = chunk
        self.mem = lib().allocateMemory(ctypes.byref(tree.ptr), chunk, num_faces)
        if not self.mem.result:
            _fail("allocateMemory")
        lib().initData(ctypes.byref(self.mem), ctypes.byref(tree.ptr), self.opts["CUDABlockSize"])
        if last_error():
            _fail("initData")

    def step(self, tree):
        lib().bn13_doAlgorithmStep(ctypes.byref(self.mem), ctypes.byref(tree.ptr), self.cfg(self.chunk))
        if last_error():
            _fail("bn13_doAlgorithmStep")

    def run(self, tree, num_faces, steps=None, sah_every=0, callback=None):
        """Runs the optimizer; returns the list of per-step unit-cost SAH values
        (CUDAcalcSAH) when sah_every > 0."""
        self.setup(tree, num_faces)
        sahs = []
        if sah_every:
            sahs.append(lib().CUDAcalcSAH(ctypes.byref(tree.ptr)))
        n = self.opts["maxNumStepps"] if steps is None else steps
        for i in range(n):
            self.step(tree)
            if sah_every and (i + 1) % sah_every == 0:
                sahs.append(lib().CUDAcalcSAH(ctypes.byref(tree.ptr)))
            if callback:
                callback(i)
        lib().doCudaDevSync()
        return sahs

    def stats(self):
        out = (ctypes.c_uint32 * 8)()
        lib().bvht_bn13_stats(ctypes.byref(self.mem), out, 8)
        return dict(conflicts=out[0], failed=out[1], contained=out[2], overflows=out[3], applied=out[4],
                    unresolved=out[5], runaway=out[6], sort_errors=out[7],
                    skipped=lib().calcNumSkipped(ctypes.byref(self.mem)))

    def free(self):
        if self.mem is not None and self.mem.result:
            lib().freeMemory(ctypes.byref(self.mem))
        self.mem = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def calc_sah(tree, c_inner=1.0, c_leaf=1.0):
    return lib().bvht_calcSAHWeighted(ctypes.byref(tree.ptr), c_inner, c_leaf)


def validate_device(tree, mesh, root=0):
    """Device-side Validate: (total, [tree, bbox, strict, SA, triangle, orphan] counts)."""
    err = (ctypes.c_uint32 * 6)()
    total = lib().bvht_validate(ctypes.byref(tree.ptr), root, mesh.raw, err)
    if total == 0xFFFFFFFF:
        _fail("bvht_validate")
    return total, list(err)


# ---------------------------------------------------------------------------
# learned removal priority (12-64-64-1 MLP, BF16 MFMA)
# ---------------------------------------------------------------------------

class PriorityModel:
    """Device MLP ranking Bittner13 removal candidates by predicted SAH gain."""

    def __init__(self, seed=0, params=None):
        self.handle = lib().bvht_mlpCreate(seed)
        if not self.handle:
            _fail("bvht_mlpCreate")
        self.num_params = lib().bvht_mlpNumParams()
        if params is not None:
            self.set_params(params)

    def params(self):
        out = np.empty(self.num_params, np.float32)
        if not lib().bvht_mlpGetParams(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
            _fail("bvht_mlpGetParams")
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (self.num_params,)
        if not lib().bvht_mlpSetParams(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
            _fail("bvht_mlpSetParams")

    def predict(self, feat):
        feat = np.ascontiguousarray(feat, np.float32).reshape(-1, 12)
        n = len(feat)
        fb = DeviceBuffer(max(feat.nbytes, 16)).upload(feat)
        ob = DeviceBuffer(max(4 * n, 16))
        if not lib().bvht_mlpPredict(self.handle, fb.ptr, n, ob.ptr):
            _fail("bvht_mlpPredict")
        lib().doCudaDevSync()
        return ob.download(np.float32, n)

    def train(self, feat, target, epochs=10, batch=4096, lr=1e-3, seed=0):
        """Adam on the mean squared error; returns the last epoch's mean loss."""
        feat = np.ascontiguousarray(feat, np.float32).reshape(-1, 12)
        target = np.ascontiguousarray(target, np.float32)
        assert len(feat) == len(target) and len(feat) > 0
        fb = DeviceBuffer(feat.nbytes).upload(feat)
        tb = DeviceBuffer(target.nbytes).upload(target)
        loss = lib().bvht_mlpTrain(self.handle, fb.ptr, tb.ptr, len(feat), epochs, batch, lr, seed)
        if loss < 0:
            _fail("bvht_mlpTrain")
        return loss

    def save(self, path):
        np.save(path, self.params())

    @classmethod
    def load(cls, path):
        return cls(params=np.load(path, allow_pickle=False))

    def free(self):
        if self.handle:
            lib().bvht_mlpFree(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# batched (segmented) builds
# ---------------------------------------------------------------------------

def scene_batch(num_meshes=512, seed=1000, min_tris=1000, max_tris=500000):
    """Concatenated meshes; returns (vert, faces, face_counts, vert_counts)."""
    from ._abi import LBVHBatchWorkingMemory  # noqa: F401
    L = lib()
    fc = np.empty(num_meshes, np.uint32)
    vc = np.empty(num_meshes, np.uint32)
    tv, tf = ctypes.c_uint64(), ctypes.c_uint64()
    L.bvht_sceneBatch(num_meshes, seed, min_tris, max_tris, _ptr(fc), _ptr(vc), None, None, ctypes.byref(tv),
                      ctypes.byref(tf))
    v = np.empty((tv.value, 3), np.float32)
    f = np.empty((tf.value, 3), np.uint32)
    L.bvht_sceneBatch(num_meshes, seed, min_tris, max_tris, _ptr(fc), _ptr(vc), _ptr(v), _ptr(f), ctypes.byref(tv),
                      ctypes.byref(tf))
    return v, f, fc, vc


class BatchBuilder:
    """Segmented LBVH build of concatenated meshes (one launch sequence, optional hipGraph)."""

    def __init__(self, vert, faces, face_counts, vert_counts, tile_log2=10, radix_bits=8):
        L = lib()
        fc = np.ascontiguousarray(face_counts, np.uint32)
        vc = np.ascontiguousarray(vert_counts, np.uint32)
        self.mem = L.bvht_LBVH_allocateBatch(len(fc), fc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                             vc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), tile_log2)
        if not self.mem.lRes:
            _fail("bvht_LBVH_allocateBatch")
        self.mem.radixBits = radix_bits
        self.face_counts, self.vert_counts = fc, vc
        self.vb = DeviceBuffer(vert.nbytes).upload(np.ascontiguousarray(vert, np.float32))
        self.fb = DeviceBuffer(faces.nbytes).upload(np.ascontiguousarray(faces, np.uint32))
        self.nb = DeviceBuffer(self.mem.totalNodes * NODE_DTYPE.itemsize)

    def build(self, use_graph=False):
        if not lib().bvht_LBVH_buildBatch(ctypes.byref(self.mem), self.vb.ptr, self.fb.ptr, self.nb.ptr, use_graph):
            _fail("bvht_LBVH_buildBatch")

    def autotune(self):
        """(tile log2, per-choice mean ms: tiles 2^8..2^11 with 8-bit digits, then with 10-bit digits);
        the chosen digit width is in self.mem.radixBits."""
        ms = (ctypes.c_float * 8)()
        best = lib().bvht_LBVH_autotuneBatch(ctypes.byref(self.mem), self.vb.ptr, self.fb.ptr, self.nb.ptr, ms)
        return best, list(ms)

    def status(self):
        """Error bits of the last batched sort: 0 = complete, 1 = a look-back wait gave up."""
        return lib().bvht_LBVH_batchStatus(ctypes.byref(self.mem))

    def offsets(self, i):
        fo, vo, no = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        lib().bvht_LBVH_batchOffsets(ctypes.byref(self.mem), i, ctypes.byref(fo), ctypes.byref(vo), ctypes.byref(no))
        return fo.value, vo.value, no.value

    def nodes(self):
        lib().LBVH_doCUDASyc()
        return self.nb.download(NODE_DTYPE, self.mem.totalNodes)

    def free(self):
        if self.mem.lRes:
            lib().bvht_LBVH_freeBatch(ctypes.byref(self.mem))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
