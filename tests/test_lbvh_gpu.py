"""GPU parity of the LBVH builder against the oracle, through the C ABI.

Bar: bit-exact.  Morton keys, sorted permutation, and every byte of every
48-byte node (topology, boxes, SA, flags) must equal the oracle's
restatement of lib/builder/LBVH_CUDA.cu.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def soup(n, seed=0, spread=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    v = (c[:, None, :] + (rng.random((n, 3, 3), dtype=np.float32) - 0.5) * spread).reshape(-1, 3)
    return v.astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def meshes(bvh):
    out = {f"soup{n}": soup(n, seed=n) for n in (1, 2, 3, 17, 64, 1000, 1025, 4097, 100003)}
    out["heightfield65536"] = bvh.scene_heightfield(256, 128, 1, 0.2)
    out["spheres_small"] = bvh.scene_spheres_soup(64, 5000, 3)
    # shared vertices and degenerate / coincident triangles
    v, f = bvh.scene_heightfield(40, 20, 4, 0.0)
    out["flat_dupes"] = (v, np.concatenate([f, f[:100]]))
    return out


@pytest.fixture(scope="module")
def cases(bvh):
    return meshes(bvh)


def nodes_equal(a, b):
    if a.tobytes() == b.tobytes():
        return True
    for name in a.dtype.names:
        if not np.array_equal(a[name], b[name]):
            bad = np.nonzero((a[name] != b[name]).reshape(len(a), -1).any(axis=1))[0]
            print(f"field {name} differs at {len(bad)} nodes, first {bad[:5]}: {a[bad[:3]]} vs {b[bad[:3]]}")
    return False


@pytest.mark.parametrize("digits", [16, 8], ids=["msd", "onesweep"])
@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("name", ["soup1", "soup2", "soup3", "soup17", "soup64", "soup1000", "soup1025", "soup4097",
                                  "soup100003", "heightfield65536", "spheres_small", "flat_dupes"])
def test_fused_build_bit_exact(bvh, orc, cases, name, bits, digits):
    v, f = cases[name]
    print(f"case {name} n={len(f)} bits={bits} digits={digits}", flush=True)
    keys, sf, expect, scene = orc.lbvh(v, f, bits)
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, bits, digit_bits=digits)
    assert b.sort_status() == 0  # before the first build
    tree = b.allocate_bvh()
    b.build(tree)
    got_sf = b.sorted_faces()
    got = tree.download()
    assert b.sort_status() == 0
    assert np.array_equal(got_sf, sf)
    assert nodes_equal(got, expect)


@pytest.mark.parametrize("tile", [9, 10, 11])
@pytest.mark.parametrize("name", ["soup2", "soup17", "soup1000", "soup1025", "soup4097", "soup100003", "flat_dupes"])
def test_tile_sizes_bit_exact(bvh, orc, cases, name, tile):
    # tiles of 2^9..2^11 sorted positions (2^8 is the default every other test
    # builds with), each cut at n: the inner nodes of a tile come from the
    # lanes of their indices and from the upper levels
    v, f = cases[name]
    print(f"case {name} n={len(f)} tile=2^{tile}", flush=True)
    keys, sf, expect, scene = orc.lbvh(v, f, 30)
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, 30, tile_log2=tile)
    tree = b.allocate_bvh()
    b.build(tree)
    assert np.array_equal(b.sorted_faces(), sf)
    assert nodes_equal(tree.download(), expect)


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("name", ["soup3", "soup1025", "heightfield65536", "spheres_small"])
def test_staged_build_bit_exact(bvh, orc, cases, name, bits):
    v, f = cases[name]
    print(f"case {name} n={len(f)} bits={bits}", flush=True)
    keys, sf, expect, scene = orc.lbvh(v, f, bits)
    unsorted = np.empty_like(keys)
    unsorted[sf] = keys
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, bits)
    tree = b.allocate_bvh()
    got_keys = []
    box = b.build_staged(tree, keys_out=got_keys)
    assert np.array_equal(box.reshape(-1), scene[:6])
    assert np.array_equal(got_keys[0].astype(np.uint64), unsorted)
    assert np.array_equal(b.sorted_faces(), sf)
    assert nodes_equal(tree.download(), expect)


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("name", ["soup17", "soup4097", "heightfield65536", "spheres_small", "flat_dupes"])
@pytest.mark.parametrize("digits", [16, 8], ids=["msd", "onesweep"])
def test_karras_topology_extension(bvh, orc, cases, name, bits, digits):
    """topology 1: Karras 2012 over the full (morton, index) key."""
    v, f = cases[name]
    print(f"case {name} n={len(f)} bits={bits} digits={digits}", flush=True)
    keys, sf, expect, _ = orc.lbvh(v, f, bits, topology=1)
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, bits, topology=1, digit_bits=digits)
    tree = b.allocate_bvh()
    b.build(tree)
    assert np.array_equal(b.sorted_keys().astype(np.uint64), keys)
    got = tree.download()
    # parent of leaves is written by the topology kernel; the reference's
    # kGenLeafNodes UINT32_MAX is overwritten for n > 1
    assert nodes_equal(got, expect)


def test_copy_to_host_and_validate(bvh, orc):
    v, f = bvh.scene_spheres_soup(128, 20000, 9)
    mesh = bvh.DeviceMesh(v, f)
    tree = bvh.lbvh_build(mesh)
    nodes, root, ml = tree.to_host()
    assert root == 0 and len(nodes) == 2 * len(f) - 1
    max_level = orc.fix_levels(nodes, root)
    total, err = orc.validate(nodes, root, max_level, v, f)
    assert total == 0, err


def test_full_size_63bit_scene(bvh, orc):
    """The ~4M-triangle scene at full size: bit-exact against the oracle."""
    v, f = bvh.scene_spheres_soup(2048, 500000, 2)
    print(f"full scene n={len(f)}", flush=True)
    keys, sf, expect, _ = orc.lbvh(v, f, 63)
    mesh = bvh.DeviceMesh(v, f)
    for digits in (16, 8):
        b = bvh.LBVHBuilder(mesh, 63, digit_bits=digits)
        tree = b.allocate_bvh()
        for _ in range(2):  # a rebuild into the same memory gives the same tree
            b.build(tree)
            assert b.sort_status() == 0
            assert np.array_equal(b.sorted_faces(), sf)
            assert tree.download().tobytes() == expect.tobytes()
        tree.free()


@pytest.mark.timeout(300)
def test_9M_triangles_both_key_widths_bit_exact(bvh, orc):
    """4,096 spheres + 2,000,000 soup triangles (about 9.1M, past 2^23): more
    than twice the keys one round of wide sort tiles holds, so both key
    widths sort with the 256-thread passes over several rounds of tiles
    (30-bit keys past one round are tested nowhere else); permutation and
    every node byte against the oracle."""
    v, f = bvh.scene_spheres_soup(4096, 2000000, 5)
    print(f"scene n={len(f)}", flush=True)
    assert len(f) > (1 << 23)
    mesh = bvh.DeviceMesh(v, f)
    for bits in (30, 63):
        keys, sf, expect, _ = orc.lbvh(v, f, bits)
        b = bvh.LBVHBuilder(mesh, bits)
        tree = b.allocate_bvh()
        b.build(tree)
        assert b.sort_status() == 0
        assert np.array_equal(b.sorted_faces(), sf), bits
        assert tree.download().tobytes() == expect.tobytes(), bits
        tree.free()
        b.free()


def test_63bit_past_wide_tiles(bvh, orc):
    """2,048 spheres + 800,000 soup triangles (4.54M): more keys than one round
    of wide sort tiles holds (256 x 17,408 on MI355X), so the full build
    sorts full keys with values (256-thread tiles) instead of packed keys."""
    v, f = bvh.scene_spheres_soup(2048, 800000, 4)
    print(f"scene n={len(f)}", flush=True)
    assert len(f) > 256 * 17408
    keys, sf, expect, _ = orc.lbvh(v, f, 63)
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, 63)
    tree = b.allocate_bvh()
    b.build(tree)
    assert b.sort_status() == 0
    assert np.array_equal(b.sorted_faces(), sf)
    assert tree.download().tobytes() == expect.tobytes()


class ShiftedMesh:
    """A MeshRaw whose vertex array starts 4 bytes into its allocation (not
    16-byte aligned)."""

    def __init__(self, bvh, v, f):
        import ctypes
        v = np.ascontiguousarray(v, np.float32)
        f = np.ascontiguousarray(f, np.uint32)
        self.vbuf = bvh.DeviceBuffer(v.nbytes + 16)
        self.fbuf = bvh.DeviceBuffer(f.nbytes).upload(f)
        base = self.vbuf.ptr.value + 4
        assert bvh.lib().runMemcpy(ctypes.c_void_p(base), v.ctypes.data, v.nbytes, 1)
        self.raw = bvh.MeshRaw()
        self.raw.vert, self.raw.faces = base, self.fbuf.ptr.value
        self.raw.numVert, self.raw.numFaces = len(v), len(f)
        self.num_faces, self.num_vert = len(f), len(v)


def stray_vertex_cases(bvh):
    """The full build takes its scene box from every vertex and checks it
    against the triangles' box (the reference's), recomputing the keys when
    they differ.  Unreferenced vertices outside the triangles' box take that
    repair; inside it, or none at all, they do not."""
    v, f = bvh.scene_heightfield(64, 32, 5, 0.2)
    lo, hi = v.min(axis=0), v.max(axis=0)
    out = {
        "outside_max": (np.concatenate([v, [[5.0, 3.0, 7.0]]]).astype(np.float32), f),
        "outside_min": (np.concatenate([[[-2.0, -1.0, -3.0]], v]).astype(np.float32), f + 1),
        "inside": (np.concatenate([v, [(lo + hi) / 2]]).astype(np.float32), f),
        "many_unreferenced": (np.concatenate([v, np.random.default_rng(3).random((1001, 3), dtype=np.float32) * 9 - 4]
                                             ).astype(np.float32), f),
    }
    return out


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("name", ["outside_max", "outside_min", "inside", "many_unreferenced", "unaligned"])
def test_vertex_box_check_and_repair(bvh, orc, name, bits):
    cases = stray_vertex_cases(bvh)
    v, f = cases["outside_max" if name == "unaligned" else name]
    print(f"case {name} n={len(f)} V={len(v)} bits={bits}", flush=True)
    keys, sf, expect, _ = orc.lbvh(v, f, bits)
    mesh = ShiftedMesh(bvh, v, f) if name == "unaligned" else bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, bits, digit_bits=8)
    tree = b.allocate_bvh()
    for _ in range(2):  # the repair flag of one build does not leak into the next
        b.build(tree)
        assert b.sort_status() == 0
        assert np.array_equal(b.sorted_faces(), sf)
        assert nodes_equal(tree.download(), expect)
    tree.free()


def test_errors_are_reported(bvh):
    L = bvh.lib()
    v, f = soup(10)
    mesh = bvh.DeviceMesh(v, f)
    with pytest.raises(bvh.BVHTestError, match="mortonBits"):
        bvh.LBVHBuilder(mesh, 31)
    L.bvht_setAllocFailAfter(2)
    with pytest.raises(bvh.BVHTestError, match="injected allocation failure"):
        bvh.LBVHBuilder(mesh, 30)
    L.bvht_setAllocFailAfter(0)
    b = bvh.LBVHBuilder(mesh, 30)
    tree = b.allocate_bvh()
    with pytest.raises(bvh.BVHTestError, match="already allocated"):
        import ctypes
        if not L.LBVH_allocateBVH(ctypes.byref(tree.ptr), ctypes.byref(mesh.raw)):
            bvh._fail("LBVH_allocateBVH")


@pytest.mark.parametrize("bits", [30, 63])
def test_stage_timing_entries(bvh, bits):
    """Per-kernel timing of a build names every kernel in launch order; with
    pass groups the onesweep passes are one entry.  The timed builds give
    the same tree as an untimed one."""
    import ctypes

    L = bvh.lib()
    v, f = soup(50000, seed=7)
    mesh = bvh.DeviceMesh(v, f)
    b = bvh.LBVHBuilder(mesh, bits)
    tree = b.allocate_bvh()
    b.build(tree)
    ref = tree.download().tobytes()

    def timed(group):
        L.bvht_LBVH_setTiming(True)
        L.bvht_LBVH_setTimingPassGroups(group)
        try:
            b.build(tree)
            buf = (ctypes.c_float * 32)()
            k = L.bvht_LBVH_stageTimes(buf, 32)
            return [L.bvht_LBVH_stageName(i).decode() for i in range(k)], [buf[i] for i in range(k)]
        finally:
            L.bvht_LBVH_setTimingPassGroups(False)
            L.bvht_LBVH_setTiming(False)

    passes = 5 if bits == 63 else 4
    names, ms = timed(False)
    assert names[:3] == ["k_vertBox", "k_morton", "k_mortonFix"]
    assert names.count("k_onesweep") == passes and "k_buildLevel0" in names
    assert all(x > 0 for x in ms)
    assert tree.download().tobytes() == ref
    gnames, gms = timed(True)
    assert gnames.count("k_onesweep") == 1
    assert [x for x in gnames if x != "k_onesweep"] == [x for x in names if x != "k_onesweep"]
    assert all(x > 0 for x in gms)
    assert tree.download().tobytes() == ref
