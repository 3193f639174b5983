"""CPU tests that pin the oracle (oracle/) before it is trusted as the checker.

Each check is independent of the oracle's own formula: known answers derived
by hand, a different algorithm for the same object (top-down radix tree vs.
the reference's determineRange/findSplit, exhaustive search vs. branch and
bound, brute force vs. traversal), or an invariant the mathematics fixes.
"""
import numpy as np
import pytest

F32_EPS = np.float32(np.finfo(np.float32).eps)


def interleave(x, y, z, bits):
    """Morton code by string interleaving (x most significant in each triple)."""
    bx, by, bz = (format(v, f"0{bits}b") for v in (x, y, z))
    return int("".join(a + b + c for a, b, c in zip(bx, by, bz)), 2)


def radix_tree(n):
    """Binary radix tree over the keys 0..n-1, built top-down from its
    definition (split where the highest differing bit of the range's end
    keys flips), numbered the way Karras 2012 numbers it: a left child is
    stored at its last key, a right child at its first key, leaves at n-1+i.
    Returns left, right, parent, isLeft arrays of length 2n-1."""
    N = 2 * n - 1
    left = np.zeros(N, np.int64)
    right = np.zeros(N, np.int64)
    parent = np.zeros(N, np.int64)
    is_left = np.zeros(N, np.int64)
    if n == 1:
        return left, right, parent, is_left
    stack = [(0, n - 1, 0)]
    while stack:
        l, r, idx = stack.pop()
        h = (l ^ r).bit_length() - 1
        m = (r >> h) << h
        a = (n - 1 + l) if l == m - 1 else m - 1
        b = (n - 1 + r) if m == r else m
        left[idx], right[idx] = a, b
        parent[a], parent[b] = idx, idx
        is_left[a], is_left[b] = 1, 0
        if l != m - 1:
            stack.append((l, m - 1, m - 1))
        if m != r:
            stack.append((m, r, m))
    return left, right, parent, is_left


def tri_mesh_soup(n, seed=0, spread=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    v = (c[:, None, :] + (rng.random((n, 3, 3), dtype=np.float32) - 0.5) * spread).reshape(-1, 3)
    return v.astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def test_spread3_known_answers(orc):
    assert orc.lib().orc_spread3(0x3FF, 10) == sum(8 ** b for b in range(10)) == 0x9249249
    assert orc.lib().orc_spread3(1, 10) == 1 and orc.lib().orc_spread3(2, 10) == 8
    rng = np.random.default_rng(1)
    for v in rng.integers(0, 1 << 21, 200):
        expect = int("".join(c + "00" for c in format(int(v), "021b")), 2) >> 2
        assert orc.lib().orc_spread3(int(v), 21) == expect


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_known_answers(orc, bits):
    """Centroids at cell centres of the quantisation grid give known keys."""
    per = bits // 3
    Q = 1 << per
    rng = np.random.default_rng(7)
    cells = rng.integers(0, Q, size=(200, 3))
    tris = []
    for kx, ky, kz in cells:
        c = (np.array([kx, ky, kz], np.float64) + 0.5) / Q
        d = 0.1 / Q  # small triangle around the centroid
        tris.append([c + [d, 0, 0], c + [-d, d, 0], c + [0, -d, -d * 0]])
    # one large triangle pins the scene box to [-eps, 1 + eps]^3
    tris.append([[0, 0, 0], [1, 0, 1], [0, 1, 0]])
    tris.append([[1, 1, 1], [1, 1, 0.999], [0.999, 1, 1]])
    v = np.array(tris, np.float32).reshape(-1, 3)
    f = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    keys, sf, nodes, scene = orc.lbvh(v, f, bits)
    assert np.all(scene[:3] == -F32_EPS) and np.all(scene[3:6] == np.float32(1) + F32_EPS)
    unsorted = np.empty_like(keys)
    unsorted[sf] = keys
    # the centroid of each small triangle is its centre to within 1e-4 of a cell
    for i, (kx, ky, kz) in enumerate(cells):
        assert int(unsorted[i]) == interleave(int(kx), int(ky), int(kz), per), i
    # the corner triangle's centroid at (1, 1, 0.9997) clamps x, y to Q-1
    last = int(unsorted[-1])
    assert last >> (3 * per - 2) == 0b11


def test_scene_box_and_sort(orc):
    v, f = tri_mesh_soup(5000, seed=3)
    v[:, 1] -= 0.25  # some negative coordinates
    keys, sf, nodes, scene = orc.lbvh(v, f)
    tv = v[f]  # (n, 3, 3)
    mn = np.minimum(tv.min(axis=1) - F32_EPS, 0).min(axis=0)
    mx = np.maximum(tv.max(axis=1) + F32_EPS, 0).max(axis=0)
    assert np.array_equal(scene[:3], mn.astype(np.float32)) and np.array_equal(scene[3:6], mx.astype(np.float32))
    assert np.array_equal(scene[6:9], (np.float32(0) - mn).astype(np.float32))
    unsorted = np.empty_like(keys)
    unsorted[sf] = keys
    assert np.array_equal(sf, np.argsort(unsorted, kind="stable"))
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)


@pytest.mark.parametrize("n", list(range(1, 70)) + [255, 256, 257, 1000, 4097, 65536, 100003])
def test_topology_is_radix_tree_of_indices(orc, n):
    """determineRange/findSplit with delta = clz32(i ^ j) (the reference as
    compiled) gives the radix tree over the indices for any keys."""
    v, f = tri_mesh_soup(n, seed=n)
    _, _, nodes, _ = orc.lbvh(v, f)
    left, right, parent, is_left = radix_tree(n)
    inner = nodes["isLeafFlag"] == 0
    assert inner.sum() == n - 1
    assert np.array_equal(nodes["left"][inner], left[:n - 1]) and np.array_equal(nodes["right"][inner], right[:n - 1])
    if n > 1:
        assert np.array_equal(nodes["parent"][1:], parent[1:]) and np.array_equal(nodes["isLeft"][1:], is_left[1:])
    assert nodes["parent"][0] == 0


def test_65536_is_perfect_depth16(orc):
    v, f = tri_mesh_soup(65536, seed=5)
    _, _, nodes, _ = orc.lbvh(v, f)
    assert orc.fix_levels(nodes, 0) == 16
    assert np.all(nodes["level"][65535:] == 16)


@pytest.mark.parametrize("bits", [30, 63])
def test_refit_invariants(orc, bits):
    v, f = tri_mesh_soup(20000, seed=11)
    _, sf, nodes, _ = orc.lbvh(v, f, bits)
    ml = orc.fix_levels(nodes, 0)
    total, err = orc.validate(nodes, 0, ml, v, f)
    assert total == 0, err
    tv = v[f]
    assert np.array_equal(nodes["bbox"][0][:3], (tv.min(axis=1) - F32_EPS).min(axis=0))
    assert np.array_equal(nodes["bbox"][0][3:], (tv.max(axis=1) + F32_EPS).max(axis=0))
    leaves = nodes[len(f) - 1:]
    assert np.array_equal(leaves["left"], sf) and np.all(leaves["right"] == 1)


def test_validate_detects_corruption(orc):
    v, f = tri_mesh_soup(1000, seed=2)
    _, _, nodes, _ = orc.lbvh(v, f)
    ml = orc.fix_levels(nodes, 0)
    bad = nodes.copy()
    bad["bbox"][5][0] -= 0.5  # inner box no longer the strict union
    assert orc.validate(bad, 0, ml, v, f)[1][2] > 0
    bad = nodes.copy()
    bad["parent"][1500] = 3
    assert orc.validate(bad, 0, ml, v, f)[1][0] > 0


def test_calc_sah(orc):
    v, f = tri_mesh_soup(3000, seed=4)
    _, _, nodes, _ = orc.lbvh(v, f)
    sa = nodes["surfaceArea"].astype(np.float64)
    leaf = nodes["isLeafFlag"] != 0
    expect = (sa[1:] * np.where(leaf[1:], 1.0, 1.2)).sum() / sa[0]
    assert orc.calc_sah(nodes, 0, 1.2, 1.0) == pytest.approx(expect, rel=1e-6)
    # SA of the boxes is 2(dx dy + dx dz + dy dz)
    b = nodes["bbox"].astype(np.float64)
    d = b[:, 3:] - b[:, :3]
    assert np.allclose(sa, 2 * (d[:, 0] * d[:, 1] + d[:, 0] * d[:, 2] + d[:, 1] * d[:, 2]), rtol=1e-5)


@pytest.mark.parametrize("bits", [30, 63])
def test_omp_build_equals_serial(orc, bits):
    v, f = tri_mesh_soup(30000, seed=9)
    _, sf, nodes, _ = orc.lbvh(v, f, bits)
    _, nodes2, sf2 = orc.lbvh_omp(v, f, bits, threads=4)
    assert np.array_equal(sf, sf2)
    assert nodes.tobytes() == nodes2.tobytes()


def heightfield(gx=32, gz=16, amp=0.2, seed=1):
    """Deterministic small heightfield (numpy), faces wound to +y normals."""
    xs = np.arange(gx + 1) / gx
    zs = np.arange(gz + 1) / gx
    X, Z = np.meshgrid(xs, zs)
    Y = amp * (0.5 + 0.5 * np.sin(7 * X + seed) * np.cos(5 * Z))
    v = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    f = []
    for j in range(gz):
        for i in range(gx):
            v00 = j * (gx + 1) + i
            f += [[v00, v00 + gx + 1, v00 + 1], [v00 + 1, v00 + gx + 1, v00 + gx + 2]]
    return v, np.array(f, np.uint32)


def test_trace_matches_bruteforce(orc):
    v, f = heightfield(48, 24)
    _, _, nodes, _ = orc.lbvh(v, f)
    rays = orc.gen_rays(96, 96, (0.5, 1.0, 1.25), (0.5, 0.1, 0.25), (0, 1, 0), 45.0)
    pix, hits = orc.trace(rays, nodes, 0, v, f)
    bf = orc.trace_bruteforce(rays, v, f, tmin=0.0)
    hit = bf["triId"] != 0xFFFFFFFF
    assert hit.sum() > 0.2 * len(rays)
    assert np.array_equal(hits["triId"] != 0xFFFFFFFF, hit)
    assert np.array_equal(pix["hit"] == 1, hit)
    same = hits["triId"] == bf["triId"]
    # ids may differ only on exact distance ties (shared edges)
    assert np.all(same | (hits["t"] == bf["t"]))
    assert np.array_equal(hits["t"][hit], bf["t"][hit])


def test_trace_flat_plane_analytic(orc):
    v, f = heightfield(16, 16, amp=0.0)
    _, _, nodes, _ = orc.lbvh(v, f)
    pos = np.array([0.5, 1.0, 0.5], np.float32)
    rays = orc.gen_rays(64, 64, pos, (0.5, 0.0, 0.5), (0, 0, -1), 60.0)
    _, hits = orc.trace(rays, nodes, 0, v, f)
    d = rays["dir"].astype(np.float64)
    ok = hits["triId"] != 0xFFFFFFFF
    t = pos[1] / -d[:, 1]
    hp = pos + t[:, None] * d
    inside = (hp[:, 0] > 1e-3) & (hp[:, 0] < 1 - 1e-3) & (hp[:, 2] > 1e-3) & (hp[:, 2] < 1 - 1e-3)
    assert np.array_equal(ok[inside], np.ones(inside.sum(), bool))
    assert np.allclose(hits["t"][inside], t[inside], rtol=1e-5)


def test_gen_rays_geometry(orc):
    w, h = 65, 33
    pos, at = np.array([1.0, 2.0, 3.0]), np.array([0.2, 0.1, -0.5])
    rays = orc.gen_rays(w, h, pos, at, (0, 1, 0), 45.0)
    d = rays["dir"].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    centre = d[(h // 2) * w + w // 2]
    expect = (at - pos) / np.linalg.norm(at - pos)
    assert np.allclose(centre, expect, atol=1e-6)
    # angle between centre and the top-centre pixel = atan(tan(fov/2) * (1 - 1/h))
    top = d[w // 2]
    ang = np.arccos(np.clip(top @ centre, -1, 1))
    assert ang == pytest.approx(np.arctan(np.tan(np.radians(22.5)) * (1 - 1 / h)), rel=1e-5)
    assert np.allclose(rays["invDir"], 1.0 / rays["dir"])
    assert np.array_equal(rays["sign"], (rays["invDir"] < 0).astype(np.uint8))


def insertion_costs(nodes, root, box, parent_of):
    """Exhaustive Bittner cost of inserting `box` as sibling of every node."""
    def sa(b):
        d = b[3:] - b[:3]
        return 2.0 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])

    def merge(a, b):
        return np.concatenate([np.minimum(a[:3], b[:3]), np.maximum(a[3:], b[3:])])

    bb = nodes["bbox"].astype(np.float64)
    cost = np.empty(len(nodes))
    for x in range(len(nodes)):
        c = sa(merge(bb[x], box))
        a = x
        while a != root:
            a = parent_of[a]
            c += sa(merge(bb[a], box)) - sa(bb[a])
        cost[x] = c
    return cost


def test_find_node_is_exhaustive_optimum(orc):
    v, f = tri_mesh_soup(400, seed=21, spread=0.05)
    _, _, nodes, _ = orc.lbvh(v, f)
    rng = np.random.default_rng(5)
    for _ in range(40):
        c = rng.random(3).astype(np.float32)
        e = (rng.random(3) * 0.2).astype(np.float32)
        box = np.concatenate([c - e, c + e]).astype(np.float32)
        sa = float(orc.lib().orc_box_sa(orc.p(box)))
        got = orc.seq_find_node_box(nodes, 0, box, sa)
        cost = insertion_costs(nodes, 0, box.astype(np.float64), nodes["parent"])
        assert cost[got] <= cost.min() * (1 + 1e-5)


def test_bittner_seq_improves_valid_tree(orc):
    v, f = heightfield(64, 32)
    _, _, nodes, _ = orc.lbvh(v, f)
    opt, root, sah = orc.bittner_seq(nodes, 0, 10.0, 10)
    ml = orc.fix_levels(opt, root)
    total, err = orc.validate(opt, root, ml, v, f)
    assert total == 0, err
    assert sah[-1] < 0.8 * sah[0]
    assert orc.calc_sah(opt, root) == pytest.approx(sah[-1], rel=1e-9)


def test_bittner_gpuref_improves_valid_tree(orc):
    v, f = heightfield(64, 32)
    _, _, nodes, _ = orc.lbvh(v, f)
    rc, opt, sah, stats = orc.bittner_gpuref(nodes, 10.0, 8, 10)
    assert rc == 0
    assert orc.tree_is_valid(opt, 0)
    ml = orc.fix_levels(opt, 0)
    total, err = orc.validate(opt, 0, ml, v, f)
    assert total == 0, err
    assert sah[-1] < 0.9 * sah[0]


def test_bundle_of_one_ray_is_the_single_ray_walk(orc):
    """A tile with one valid lane votes with that ray's own preference (ties
    of the two entry distances go right, as in kTraceRay), so the bundle
    walk is the single-ray walk: pixel (including intCount) and hit equal."""
    v, f = heightfield(40, 20)
    _, _, nodes, _ = orc.lbvh(v, f)
    rays = orc.gen_rays(24, 24, (0.5, 1.0, 1.25), (0.5, 0.1, 0.25), (0, 1, 0), 45.0)
    pix, hits = orc.trace(rays, nodes, 0, v, f)
    for i in range(len(rays)):
        bp, bh = orc.trace_bundle(rays[i:i + 1], 1, 1, nodes, 0, v, f)
        assert bp.tobytes() == pix[i:i + 1].tobytes(), i
        assert bh.tobytes() == hits[i:i + 1].tobytes(), i


@pytest.mark.parametrize("scene", ["heightfield", "soup"])
def test_bundle_finds_every_rays_nearest_hit(orc, scene):
    """The shared walk enters every child some ray's box test accepts, which
    includes every node on the path to each ray's nearest triangle, so each
    ray ends with its own nearest distance: t equals the single-ray walk and
    brute force for every ray; the triangle differs only where two lie at
    exactly that distance; intCount is one value per 8x8 tile."""
    if scene == "heightfield":
        v, f = heightfield(64, 32)
        cam = ((0.5, 1.0, 1.25), (0.5, 0.1, 0.25))
    else:
        v, f = tri_mesh_soup(3000, seed=4, spread=0.08)
        cam = ((0.5, 0.5, 2.0), (0.5, 0.5, 0.5))
    _, _, nodes, _ = orc.lbvh(v, f)
    w, h = 77, 45  # partial tiles on both borders
    rays = orc.gen_rays(w, h, cam[0], cam[1], (0, 1, 0), 45.0)
    pix, hits = orc.trace(rays, nodes, 0, v, f)
    bpix, bhits = orc.trace_bundle(rays, w, h, nodes, 0, v, f)
    bf = orc.trace_bruteforce(rays, v, f, tmin=0.0)
    hit = hits["triId"] != 0xFFFFFFFF
    assert hit.sum() > 0.1 * len(rays)
    assert np.array_equal(bhits["triId"] != 0xFFFFFFFF, hit)
    assert np.array_equal(bpix["hit"], pix["hit"])
    assert np.array_equal(bhits["t"], hits["t"])
    assert np.array_equal(bhits["t"][hit], bf["t"][hit])
    diff = np.nonzero(bhits["triId"] != hits["triId"])[0]
    for i in diff:  # a tie: the bundle's triangle is hit at the same distance
        one = orc.trace_bruteforce(rays[i:i + 1], v, f[bhits["triId"][i]:bhits["triId"][i] + 1], tmin=0.0)
        assert one["t"][0] == hits["t"][i]
    ic = bpix["intCount"].reshape(h, w)
    for ty in range(0, h, 8):
        for tx in range(0, w, 8):
            assert len(np.unique(ic[ty:ty + 8, tx:tx + 8])) == 1


def test_leaf_refit_matches_triangle_boxes_and_keeps_area(orc):
    """kRefitLeafBVHTris: the leaf box equals the triangle's box widened by
    FLT_EPSILON (kInitTriData's box, pinned by numpy min/max) and the
    surface area field is left as it was; a full refit afterwards equals the
    union of the new leaf boxes."""
    v, f = tri_mesh_soup(500, seed=3)
    _, _, nodes, _ = orc.lbvh(v, f)
    v2 = v + np.float32(0.25)
    leaves = np.nonzero(nodes["isLeafFlag"])[0][::3].astype(np.uint32)
    out = orc.refit_leaf_tris(nodes, v2, f, leaves)
    tv = v2[f[nodes["left"][leaves]]]
    eps = F32_EPS
    assert np.array_equal(out["bbox"][leaves][:, :3], (tv.min(axis=1) - eps).astype(np.float32))
    assert np.array_equal(out["bbox"][leaves][:, 3:], (tv.max(axis=1) + eps).astype(np.float32))
    assert np.array_equal(out["surfaceArea"], nodes["surfaceArea"])
    others = np.setdiff1d(np.arange(len(nodes)), leaves)
    assert out[others].tobytes() == nodes[others].tobytes()
    full = orc.fix_tree1(out)
    assert orc.tree_is_valid(full, 0)
    inner = np.nonzero(full["isLeafFlag"] == 0)[0]
    for i in inner[:2000]:
        l, r = full["left"][i], full["right"][i]
        u = np.concatenate([np.minimum(full["bbox"][l][:3], full["bbox"][r][:3]),
                            np.maximum(full["bbox"][l][3:], full["bbox"][r][3:])])
        assert np.array_equal(full["bbox"][i], u)
    assert full["bbox"][0][0] == min(full["bbox"][nodes["isLeafFlag"] != 0][:, 0])
    # a built tree is already exact: the full refit changes nothing
    assert orc.fix_tree1(nodes).tobytes() == nodes.tobytes()
