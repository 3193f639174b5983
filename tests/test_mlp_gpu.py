"""Learned removal priority: features, BF16-MFMA MLP inference and training,
SAH-gain records and the learned selection path, against numpy references.

The references round exactly where the device rounds (features, weights and
every post-ReLU activation to bf16, round to nearest even; accumulation and
bias add in float), so a wrong MFMA operand layout shows up as O(1) errors.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    w1 = p[:768].reshape(12, 64)
    b1 = p[768:832]
    w2 = p[832:4928].reshape(64, 64)
    b2 = p[4928:4992]
    w3 = p[4992:5056]
    return w1, b1, w2, b2, w3, p[5056]


def forward_ref(p, x):
    w1, b1, w2, b2, w3, b3 = (a.astype(np.float64) if isinstance(a, np.ndarray) else a for a in unpack(p))
    xb = bf16(x).astype(np.float64)
    z1 = xb @ bf16(w1).astype(np.float64) + b1
    h1 = bf16(np.maximum(z1, 0).astype(np.float32)).astype(np.float64)
    z2 = h1 @ bf16(w2).astype(np.float64) + b2
    h2 = bf16(np.maximum(z2, 0).astype(np.float32)).astype(np.float64)
    y = h2 @ bf16(w3).astype(np.float64) + b3
    return y, (xb, z1, h1, z2, h2)


def grad_ref(p, x, t):
    """d(mean squared error)/dp of the bf16-rounded network (rounding passed straight through)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    y, (xb, z1, h1, z2, h2) = forward_ref(p, x)
    d = 2.0 * (y - t) / len(t)
    g = np.zeros(NP)
    g[4992:5056] = h2.T @ d
    g[5056] = d.sum()
    dz2 = np.outer(d, bf16(w3).astype(np.float64)) * (z2 > 0)
    g[4928:4992] = dz2.sum(0)
    g[832:4928] = (h1.T @ dz2).reshape(-1)
    dz1 = (dz2 @ bf16(w2).astype(np.float64).T) * (z1 > 0)
    g[768:832] = dz1.sum(0)
    g[:768] = (xb.T @ dz1).reshape(-1)
    return g


def random_features(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, 12)) * 0.7).astype(np.float32)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 70001])
def test_inference_matches_reference(bvh, n):
    m = bvh.PriorityModel(seed=3)
    p = m.params()
    p[768:832] = np.linspace(-0.2, 0.2, 64)  # nonzero biases
    p[4928:4992] = np.linspace(0.1, -0.1, 64)
    p[5056] = 0.25
    m.set_params(p)
    x = random_features(n, n)
    got = m.predict(x)
    ref, _ = forward_ref(p, x)
    scale = np.abs(ref).max() + 1e-6
    assert np.abs(got - ref).max() <= 1e-2 * scale
    assert np.median(np.abs(got - ref)) <= 1e-3 * scale


def test_first_adam_step_follows_the_gradient(bvh):
    """Adam's first step moves every parameter by -lr * g / (|g| + eps)."""
    m = bvh.PriorityModel(seed=5)
    p0 = m.params()
    x = random_features(4096, 1)
    t = (x[:, 0] * 0.5 - x[:, 3] ** 2 + 0.1 * x[:, 7]).astype(np.float32)
    lr = 1e-3
    m.train(x, t, epochs=1, batch=4096, lr=lr, seed=0)
    step = m.params() - p0
    g = grad_ref(p0, x, t)
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    assert big.sum() > 1000
    agree = np.sign(step[big]) == -np.sign(g[big])
    assert agree.mean() >= 0.99, agree.mean()
    assert np.allclose(np.abs(step[big]), lr, rtol=1e-3)


def test_training_fits_a_smooth_target(bvh):
    x = random_features(32768, 2)
    t = (np.tanh(x[:, 0] - 0.5 * x[:, 1]) + 0.3 * x[:, 2] * x[:, 5]).astype(np.float32)
    m = bvh.PriorityModel(seed=7)
    y0 = m.predict(x)
    base = float(np.mean((y0 - t) ** 2))
    loss = m.train(x, t, epochs=30, batch=2048, lr=2e-3, seed=1)
    y = m.predict(x)
    mse = float(np.mean((y - t) ** 2))
    assert loss < 0.05 * base and mse < 0.05 * base, (base, loss, mse)
    # a fixed seed trains the same weights
    m2 = bvh.PriorityModel(seed=7)
    m2.train(x, t, epochs=30, batch=2048, lr=2e-3, seed=1)
    assert np.array_equal(m2.params(), m.params())


def features_ref(nodes, root=0):
    """The 12 features of mlp.hip in float64, post-order over the tree."""
    sa = nodes["surfaceArea"].astype(np.float64)
    bb = nodes["bbox"].astype(np.float64)
    N = len(nodes)
    leaves = np.zeros(N)
    sub = np.zeros(N)
    order = []
    st = [root]
    while st:
        x = st.pop()
        order.append(x)
        if nodes["isLeafFlag"][x] == 0:
            st += [int(nodes["left"][x]), int(nodes["right"][x])]
    for x in reversed(order):
        if nodes["isLeafFlag"][x]:
            leaves[x], sub[x] = 1, sa[x]
        else:
            l, r = nodes["left"][x], nodes["right"][x]
            leaves[x] = leaves[l] + leaves[r]
            sub[x] = sub[l] + sub[r] + sa[x]
    depth = np.zeros(N)
    for x in order:
        if x != root:
            depth[x] = depth[nodes["parent"][x]] + 1

    def vol(b):
        return np.prod(np.maximum(b[3:] - b[:3], 0))

    def ov(a, b):
        d = np.minimum(a[3:], b[3:]) - np.maximum(a[:3], b[:3])
        return np.prod(d) if np.all(d > 0) else 0.0

    inner = np.nonzero(nodes["isLeafFlag"] == 0)[0]
    F = np.zeros((len(inner), 12))
    for i, x in enumerate(inner):
        S, L, R = sa[x], sa[nodes["left"][x]], sa[nodes["right"][x]]
        F[i, 0] = np.log2(S / sa[root]) / 16
        F[i, 1], F[i, 2] = L / S, R / S
        mn = min(L, R)
        F[i, 5] = mn / (L + R)
        F[i, 6] = np.log2(2 * S ** 3 / ((L + R) * mn) / sa[root]) / 16
        if x == root:
            F[i, 3], F[i, 4], F[i, 10] = 1, 0, 0
        else:
            p = nodes["parent"][x]
            s = nodes["right"][p] if nodes["isLeft"][x] else nodes["left"][p]
            F[i, 3], F[i, 4] = S / sa[p], sa[s] / sa[p]
            F[i, 10] = ov(bb[x], bb[s]) / vol(bb[p]) if vol(bb[p]) > 0 else 0
        F[i, 7] = depth[x] / 64
        F[i, 8] = np.log2(leaves[x]) / 24
        F[i, 9] = ov(bb[nodes["left"][x]], bb[nodes["right"][x]]) / vol(bb[x]) if vol(bb[x]) > 0 else 0
        F[i, 11] = np.log2(sub[x] / (S * leaves[x])) / 8
    return F


def _tree(bvh, v, f):
    mesh = bvh.DeviceMesh(v, f)
    return mesh, bvh.lbvh_build(mesh)


def test_features_match_reference(bvh):
    v, f = bvh.scene_spheres_soup(6, 600, 31)
    mesh, tree = _tree(bvh, v, f)
    opt = bvh.Bittner13GPU(maxNumStepps=2, numChunks=4, batchPercent=5.0)
    opt.run(tree, len(f))  # a tree that is no longer the LBVH layout
    got = opt.features(tree)
    ref = features_ref(tree.download())
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6), np.abs(got - ref).max(axis=0)


def test_probe_records_gains_without_changing_the_tree(bvh, orc):
    v, f = bvh.scene_heightfield(40, 32, 9, 0.2)
    mesh, tree = _tree(bvh, v, f)
    opt = bvh.Bittner13GPU(maxNumStepps=1, numChunks=4, batchPercent=5.0)
    opt.setup(tree, len(f))
    before = tree.download()
    n = opt.mem.numInnerNodes
    opt.record(n)
    opt.probe(tree, np.arange(n))
    feat, gain = opt.records()
    assert tree.download().tobytes() == before.tobytes()
    assert 0.5 * n < len(gain) <= n and np.all(np.isfinite(gain))
    rows = {r.tobytes() for r in opt.features(tree)}
    assert all(r.tobytes() in rows for r in feat)  # each record carries its node's features
    # a chunk of one patch: the recorded gain predicts the realized SAH drop
    pred, real = [], []
    one = bvh.Bittner13GPU(maxNumStepps=1, numChunks=64, batchPercent=100.0 * 64.5 / len(before))
    one.setup(tree, len(f))
    one.record(1024)
    cfg = one.cfg(1)
    cfg.chunkSize, cfg.numChunks = 1, 64
    L = bvh.lib()
    L.bn13_selectNodes(ctypes.byref(one.mem), ctypes.byref(tree.ptr), cfg)
    for c in range(64):
        s0 = L.CUDAcalcSAH(ctypes.byref(tree.ptr))
        k0 = one.records()[1].size
        L.bn13_rmAndReinsChunk(ctypes.byref(one.mem), ctypes.byref(tree.ptr), cfg, c)
        s1 = L.CUDAcalcSAH(ctypes.byref(tree.ptr))
        g = one.records()[1]
        if g.size > k0 and one.stats()["applied"] == len(real) + 1:
            pred.append(g[-1])
            real.append(s0 - s1)
    assert len(real) > 20
    assert np.corrcoef(pred, real)[0, 1] > 0.9, (pred[:10], real[:10])


def test_learned_selection_keeps_a_valid_tree(bvh, orc):
    v, f = bvh.scene_spheres_soup(8, 2000, 41)
    mesh, tree = _tree(bvh, v, f)
    opt = bvh.Bittner13GPU(maxNumStepps=4, numChunks=8, batchPercent=3.0, priority=1)
    opt.setup(tree, len(f))
    model = bvh.PriorityModel(seed=11)
    opt.set_model(model)
    s0 = bvh.calc_sah(tree)
    for _ in range(4):
        opt.step(tree)
    nodes = tree.download()
    assert orc.tree_is_valid(nodes, 0)
    ml = orc.fix_levels(nodes, 0)
    total, err = orc.validate(nodes, 0, ml, v, f)
    assert total == 0, err
    assert bvh.calc_sah(tree) < s0
    # the subtree leaf counts and SA sums were carried through the steps'
    # refits, not recomputed: the features still match a fresh evaluation
    got = opt.features(tree)
    ref = features_ref(nodes)
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6), np.abs(got - ref).max(axis=0)
    # selection without a model is refused
    bare = bvh.Bittner13GPU(maxNumStepps=1, numChunks=8, batchPercent=3.0, priority=1)
    bare.setup(tree, len(f))
    with pytest.raises(bvh.BVHTestError):
        bare.step(tree)
