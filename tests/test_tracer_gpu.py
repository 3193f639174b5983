"""GPU parity of ray generation and traversal against the oracle.

Bar: bit-exact.  Rays (f64 camera arithmetic rounded once), and per ray the
hit triangle id, hit distance, hit flag, diffuse value and inner-node count
of the reference's stackless traversal (lib/tracer/CUDAKernels.cu:114-223)
must equal the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C0_CAM = dict(pos=(0.5, 1.0, 1.25), look_at=(0.5, 0.1, 0.25), up=(0, 1, 0), fov=45.0)
BIG_CAM = dict(pos=(0.5, 0.5, 2.0), look_at=(0.5, 0.5, 0.5), up=(0, 1, 0), fov=45.0)


@pytest.fixture(scope="module")
def c0(bvh, orc):
    v, f = bvh.scene_heightfield(256, 128, 1, 0.2)
    mesh = bvh.DeviceMesh(v, f)
    tree = bvh.lbvh_build(mesh)
    nodes = tree.download()
    assert nodes.tobytes() == orc.lbvh(v, f)[2].tobytes()  # trace only a tree known to be valid
    return v, f, mesh, tree, nodes


@pytest.fixture(scope="module")
def spheres(bvh, orc):
    v, f = bvh.scene_spheres_soup(256, 40000, 5)
    mesh = bvh.DeviceMesh(v, f)
    tree = bvh.lbvh_build(mesh)
    nodes = tree.download()
    assert nodes.tobytes() == orc.lbvh(v, f)[2].tobytes()
    return v, f, mesh, tree, nodes


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (640, 360)])
def test_generate_rays_bit_exact(bvh, orc, w, h):
    got = bvh.generate_rays(w, h, **C0_CAM)
    expect = orc.gen_rays(w, h, C0_CAM["pos"], C0_CAM["look_at"], C0_CAM["up"], C0_CAM["fov"])
    assert got.tobytes() == expect.tobytes()


def test_trace_c0_bit_exact(bvh, orc, c0):
    v, f, mesh, tree, nodes = c0
    rays = bvh.generate_rays(512, 512, **C0_CAM)
    pix, hits = bvh.trace_rays(tree, mesh, rays)
    opix, ohits = orc.trace(rays, nodes, 0, v, f)
    assert (hits["triId"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(hits["triId"], ohits["triId"])
    assert np.array_equal(hits["t"], ohits["t"])
    assert pix.tobytes() == opix.tobytes()


def test_trace_image_command(bvh, orc, c0):
    """generateRays + tracerImage + copyImageToHost (the CUDATracer command)."""
    v, f, mesh, tree, nodes = c0
    img = bvh.trace_image(tree, mesh, 320, 200, **C0_CAM)
    rays = orc.gen_rays(320, 200, C0_CAM["pos"], C0_CAM["look_at"], C0_CAM["up"], C0_CAM["fov"])
    opix, _ = orc.trace(rays, nodes, 0, v, f)
    assert img.tobytes() == opix.tobytes()


@pytest.mark.parametrize("scene", ["c0", "spheres"])
def test_ray_bundles_bit_exact(bvh, orc, scene, request):
    """rayBundles: kTraceRayBundle (lib/tracer/CUDAKernels.cu:240-383), the
    rays of an 8x8 tile walking one node sequence chosen by majority vote.
    The image (hit, diffuse and the tile's inner-node count) equals the
    oracle's restatement of the bundle walk byte for byte, partial border
    tiles included (their missing lanes take no part in the vote)."""
    v, f, mesh, tree, nodes = request.getfixturevalue(scene)
    cam = C0_CAM if scene == "c0" else BIG_CAM
    w, h = 517, 301  # partial 8x8 tiles at the borders
    bundle = bvh.trace_image(tree, mesh, w, h, bundle=True, **cam)
    rays = orc.gen_rays(w, h, cam["pos"], cam["look_at"], cam["up"], cam["fov"])
    opix, ohits = orc.trace_bundle(rays, w, h, nodes, 0, v, f)
    assert bundle.tobytes() == opix.tobytes()
    assert bundle["hit"].sum() > 0.05 * w * h
    # and the per-ray contract: same hits as the single-ray walk
    single = bvh.trace_image(tree, mesh, w, h, **cam)
    assert np.array_equal(single["hit"], bundle["hit"])


def test_trace_spheres_bit_exact(bvh, orc, spheres):
    v, f, mesh, tree, nodes = spheres
    rays = bvh.generate_rays(384, 216, **BIG_CAM)
    pix, hits = bvh.trace_rays(tree, mesh, rays)
    opix, ohits = orc.trace(rays, nodes, 0, v, f)
    assert np.array_equal(hits, ohits) and pix.tobytes() == opix.tobytes()


def test_bounce_rays(bvh, orc, spheres):
    v, f, mesh, tree, nodes = spheres
    rays = bvh.generate_rays(256, 144, **BIG_CAM)
    _, hits = bvh.trace_rays(tree, mesh, rays)
    brays, skip = bvh.bounce_rays(mesh, rays, hits, seed=7)
    hit = hits["triId"] != 0xFFFFFFFF
    assert hit.sum() > 1000
    d = brays["dir"][hit].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    tv = v[f[hits["triId"][hit]]].astype(np.float64)
    n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n *= -np.sign((n * rays["dir"][hit]).sum(1))[:, None]
    cos = (n * d).sum(1)
    assert np.all(cos > -1e-4)
    assert cos.mean() == pytest.approx(2.0 / 3.0, abs=0.02)  # cosine-weighted: E[cos] = 2/3
    assert np.array_equal(skip[hit], hits["triId"][hit]) and np.all(skip[~hit] == 0xFFFFFFFF)
    pix, bh = bvh.trace_rays(tree, mesh, brays, mode=1, tmin=0.0, skip=skip)
    opix, obh = orc.trace(brays, nodes, 0, v, f, mode=1, tmin=0.0, skip=skip)
    assert np.array_equal(bh, obh) and pix.tobytes() == opix.tobytes()


def test_ray_ordering_does_not_change_hits(bvh, spheres):
    """Secondary rays traced after a sort by (direction cell, origin cell)
    (bvht_setRayOrdering(true)) must give the same hits and pixels byte for
    byte as the default slot order."""
    v, f, mesh, tree, nodes = spheres
    rays = bvh.generate_rays(320, 180, **BIG_CAM)
    _, hits = bvh.trace_rays(tree, mesh, rays)
    brays, skip = bvh.bounce_rays(mesh, rays, hits, seed=11)
    L = bvh.lib()
    pix0, bh0 = bvh.trace_rays(tree, mesh, brays, mode=1, tmin=0.0, skip=skip)
    try:
        L.bvht_setRayOrdering(True)
        pix1, bh1 = bvh.trace_rays(tree, mesh, brays, mode=1, tmin=0.0, skip=skip)
    finally:
        L.bvht_setRayOrdering(False)
    assert (bh0["triId"] != 0xFFFFFFFF).sum() > 1000
    assert bh0.tobytes() == bh1.tobytes() and pix0.tobytes() == pix1.tobytes()


def test_trace_full_scene_subsample(bvh, orc):
    """The ~4M-triangle scene at its full size: 1920x1080 primary rays and one
    diffuse bounce traced on the GPU; every 16th ray of each set is traced
    again by the oracle on the same tree and must match bit for bit."""
    v, f = bvh.scene_spheres_soup(2048, 500000, 2)
    print(f"full scene n={len(f)}", flush=True)
    mesh = bvh.DeviceMesh(v, f)
    tree = bvh.lbvh_build(mesh, 63)
    nodes = tree.download()
    rays = bvh.generate_rays(1920, 1080, **BIG_CAM)
    pix, hits = bvh.trace_rays(tree, mesh, rays)
    sub = slice(0, None, 16)
    opix, ohits = orc.trace(np.ascontiguousarray(rays[sub]), nodes, 0, v, f)
    assert (ohits["triId"] != 0xFFFFFFFF).mean() > 0.5
    assert np.array_equal(hits[sub], ohits) and pix[sub].tobytes() == opix.tobytes()
    brays, skip = bvh.bounce_rays(mesh, rays, hits, seed=3)
    bpix, bh = bvh.trace_rays(tree, mesh, brays, mode=1, tmin=0.0, skip=skip)
    obpix, obh = orc.trace(np.ascontiguousarray(brays[sub]), nodes, 0, v, f, mode=1, tmin=0.0,
                           skip=np.ascontiguousarray(skip[sub]))
    assert (obh["triId"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(bh[sub], obh) and bpix[sub].tobytes() == obpix.tobytes()


@pytest.mark.parametrize("count", [1, 15, 16, 17, 63, 64, 65, 1000, 4097])
def test_secondary_ray_counts(bvh, orc, spheres, count):
    """Secondary rays go through lane refills from a shared counter: any
    number of rays, including fewer than one refill, must match the oracle."""
    v, f, mesh, tree, nodes = spheres
    rays = bvh.generate_rays(128, 72, **BIG_CAM)
    _, hits = bvh.trace_rays(tree, mesh, rays)
    brays, skip = bvh.bounce_rays(mesh, rays, hits, seed=5)
    brays = np.ascontiguousarray(brays[:count])
    skip = np.ascontiguousarray(skip[:count])
    pix, bh = bvh.trace_rays(tree, mesh, brays, mode=1, tmin=0.0, skip=skip)
    opix, obh = orc.trace(brays, nodes, 0, v, f, mode=1, tmin=0.0, skip=skip)
    assert np.array_equal(bh, obh) and pix.tobytes() == opix.tobytes()


@pytest.mark.parametrize("w,h", [(61, 37), (64, 48), (1, 1)])
def test_image_rays_traced_in_tiles(bvh, orc, spheres, w, h):
    """Primary rays that generateRays wrote as an image are traced in 8x8
    pixel tiles (edge tiles partly empty); hits and pixels must equal the
    oracle's, ray by ray."""
    v, f, mesh, tree, nodes = spheres
    L = bvh.lib()
    rays_dev = bvh.generate_rays_device(w, h, **BIG_CAM)
    n = w * h
    hb = bvh.DeviceBuffer(n * bvh.HIT_DTYPE.itemsize)
    pb = bvh.DeviceBuffer(n * bvh.PIXEL_DTYPE.itemsize)
    assert L.bvht_traceRays(rays_dev.ptr, n, tree.ptr.nodes, 0, mesh.raw, bvh.vec3((2.0, 2.0, 2.0)), 0, 0.0, None,
                            hb.ptr, pb.ptr)
    L.tracerDoCudaSync()
    hits, pix = hb.download(bvh.HIT_DTYPE, n), pb.download(bvh.PIXEL_DTYPE, n)
    rays = rays_dev.download(bvh.RAY_DTYPE, n)
    opix, ohits = orc.trace(rays, nodes, 0, v, f)
    assert np.array_equal(hits, ohits) and pix.tobytes() == opix.tobytes()


@pytest.mark.parametrize("w,h", [(61, 37), (128, 72)])
def test_image_bounce_rays_on_device_buffers(bvh, orc, spheres, w, h):
    """A frame as the benchmark runs it, on device buffers: image rays traced
    in 8x8 tiles, bvht_makeBounceRays, secondary rays through lane refills;
    every secondary ray must match the oracle."""
    v, f, mesh, tree, nodes = spheres
    L = bvh.lib()
    n = w * h
    light = bvh.vec3((2.0, 2.0, 2.0))
    rays_dev = bvh.generate_rays_device(w, h, **BIG_CAM)
    hb = bvh.DeviceBuffer(n * bvh.HIT_DTYPE.itemsize)
    assert L.bvht_traceRays(rays_dev.ptr, n, tree.ptr.nodes, 0, mesh.raw, light, 0, 0.0, None, hb.ptr, None)
    br = bvh.DeviceBuffer(n * bvh.RAY_DTYPE.itemsize)
    sk = bvh.DeviceBuffer(n * 4)
    assert L.bvht_makeBounceRays(rays_dev.ptr, hb.ptr, n, mesh.raw, 9, br.ptr, sk.ptr)
    bh = bvh.DeviceBuffer(n * bvh.HIT_DTYPE.itemsize)
    pb = bvh.DeviceBuffer(n * bvh.PIXEL_DTYPE.itemsize)
    assert L.bvht_traceRays(br.ptr, n, tree.ptr.nodes, 0, mesh.raw, light, 1, 0.0, sk.ptr, bh.ptr, pb.ptr)
    L.tracerDoCudaSync()
    brays, skip = br.download(bvh.RAY_DTYPE, n), sk.download(np.uint32, n)
    hits, pix = bh.download(bvh.HIT_DTYPE, n), pb.download(bvh.PIXEL_DTYPE, n)
    opix, ohits = orc.trace(brays, nodes, 0, v, f, mode=1, tmin=0.0, skip=skip)
    assert (ohits["triId"] != 0xFFFFFFFF).sum() > 0
    assert np.array_equal(hits, ohits) and pix.tobytes() == opix.tobytes()
