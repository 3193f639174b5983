"""Learned Bittner13 removal priority: records, training and evaluation.

Records come from the plain GPU optimizer itself (1% of the nodes per step,
32 chunks, the reference defaults) on 32 seeded scenes of icosphere
clusters + triangle soup (~1M triangles each).  Every patch a step evaluates is recorded as the
features of its node at selection time and the SAH gain the searches found for
it.  Each step also probes uniformly sampled inner nodes, so the model sees
nodes the heuristic would not have picked.  The regression target is an
increasing map of the gain (z-score, asinh or rank; see transform_target):
the optimizer only uses the order of the model's outputs.

The model is trained with Adam (hand-written BF16-MFMA kernels, bvht_mlpTrain).
Evaluation is time to reach the plain optimizer's final SAH(1.2, 1) after 100
steps within 0.5%, on 8 held-out ~16.9M-triangle scenes (8,192 icospheres +
2M soup triangles). The learned run falls back to the plain heuristic for
one step whenever a learned step fails to lower the true SAH.

    python -m bvhtest_amd.priority_training --out weights.npy
"""
import argparse
import json
import sys
import time

import numpy as np

from . import (Bittner13GPU, BVHTestError, DeviceMesh, PriorityModel, calc_sah, lbvh_build, lib,
               scene_spheres_soup, validate_device)


def scene(seed, spheres, soup):
    return scene_spheres_soup(spheres, soup, seed)


def generate_records(seeds, spheres, soup, steps, probes_per_step, batch_percent=1.0, chunks=32, log=None):
    feats, gains = [], []
    for seed in seeds:
        v, f = scene(seed, spheres, soup)
        mesh = DeviceMesh(v, f)
        tree = lbvh_build(mesh, 63)
        opt = Bittner13GPU(numChunks=chunks, batchPercent=batch_percent)
        opt.setup(tree, len(f))
        n_inner = opt.mem.numInnerNodes
        per_step = opt.chunk * chunks + probes_per_step
        opt.record(per_step * steps)
        rng = np.random.default_rng(seed)
        for _ in range(steps):
            opt.probe(tree, rng.integers(0, n_inner, probes_per_step))
            opt.step(tree)
        lib().doCudaDevSync()
        x, g = opt.records()
        feats.append(x)
        gains.append(g)
        if log:
            log(f"records: seed {seed}, {len(f)} triangles, {len(g)} records")
        opt.free()
        tree.free()
    x = np.concatenate(feats)
    g = np.concatenate(gains)
    keep = np.isfinite(g) & np.all(np.isfinite(x), axis=1)
    return x[keep], g[keep]


def transform_target(g, kind):
    """Regression target from the SAH gains.  Only the order of the model's
    outputs matters to the optimizer, so any increasing map of the gain is a
    valid target: "z" z-scores the gains, "asinh" compresses their heavy tail
    first (scale: the median absolute gain), "rank" replaces them by their
    ranks spread over [-1, 1]."""
    g = np.asarray(g, np.float64)
    if kind == "asinh":
        scale = float(np.median(np.abs(g))) or 1.0
        g = np.arcsinh(g / scale)
    elif kind == "rank":
        r = np.argsort(np.argsort(g, kind="stable"), kind="stable")
        return (2.0 * r / max(len(g) - 1, 1) - 1.0).astype(np.float32), {"target": kind}
    elif kind != "z":
        raise ValueError(f"unknown target transform {kind!r}")
    mu, sd = float(g.mean()), float(g.std() + 1e-12)
    return ((g - mu) / sd).astype(np.float32), {"target": kind, "target_mean": mu, "target_std": sd}


def train(x, g, seed=0, epochs=(12, 6, 4), lrs=(2e-3, 5e-4, 1e-4), batch=4096, target="asinh", log=None):
    t, info = transform_target(g, target)
    model = PriorityModel(seed=seed)
    for e, lr in zip(epochs, lrs):
        loss = model.train(x, t, epochs=e, batch=batch, lr=lr, seed=seed + e)
        if log:
            log(f"train: {e} epochs at lr {lr}: loss {loss:.4f}")
    pred = model.predict(x)
    rank = np.corrcoef(np.argsort(np.argsort(pred)), np.argsort(np.argsort(t)))[0, 1]
    info.update(train_rank_corr=float(rank), records=int(len(t)))
    return model, info


def _timed_step(opt, tree):
    t0 = time.perf_counter()
    opt.step(tree)
    lib().doCudaDevSync()
    return time.perf_counter() - t0


def evaluate(model, seed, spheres, soup, plain_steps, batch_percent=1.0, chunks=32, max_factor=2.0, inspect=None):
    """Plain: `plain_steps` steps -> SAH S_p in time T_p.  Learned: steps until
    SAH <= 1.005 S_p (at most max_factor * plain_steps) -> T_l.  Step times
    include selection (features + inference for learned) and are measured
    with a device sync per step; the SAH evaluations that check each step
    are not timed.  inspect(tree, mesh, v, f), when given, sees the learned
    optimizer's final tree before it is freed."""
    v, f = scene(seed, spheres, soup)
    mesh = DeviceMesh(v, f)
    res = {"seed": seed, "triangles": int(len(f))}
    plain_tree = lbvh_build(mesh, 63)
    s0 = calc_sah(plain_tree, 1.2, 1.0)
    opt = Bittner13GPU(numChunks=chunks, batchPercent=batch_percent)
    opt.setup(plain_tree, len(f))
    tp = sum(_timed_step(opt, plain_tree) for _ in range(plain_steps))
    sp = calc_sah(plain_tree, 1.2, 1.0)
    opt.free()
    plain_tree.free()
    res.update(sah_initial=s0, plain_steps=plain_steps, plain_s=tp, plain_sah=sp)

    tree = lbvh_build(mesh, 63)
    opt = Bittner13GPU(numChunks=chunks, batchPercent=batch_percent, priority=1)
    opt.setup(tree, len(f))
    opt.set_model(model)
    target = 1.005 * sp
    tl, prev, steps, fallbacks, fallback_next = 0.0, s0, 0, 0, False
    sah = s0
    while steps < int(max_factor * plain_steps):
        opt.set_priority(0 if fallback_next else 1)
        tl += _timed_step(opt, tree)
        steps += 1
        sah = calc_sah(tree, 1.2, 1.0)
        if fallback_next:
            fallbacks += 1
            fallback_next = False
        elif sah >= prev:
            fallback_next = True  # the plain heuristic takes the next step
        prev = sah
        if sah <= target:
            break
    total, _ = validate_device(tree, mesh)
    res.update(learned_steps=steps, learned_s=tl, learned_sah=sah, fallback_steps=fallbacks,
               reached=bool(sah <= target), time_ratio=(tl / tp) if sah <= target else None,
               learned_tree_valid=total == 0)
    if inspect:
        inspect(tree, mesh, v, f)
    opt.free()
    tree.free()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="weights (.npy, fp32)")
    ap.add_argument("--train-seeds", default="100-131")
    ap.add_argument("--eval-seeds", default="200-207")
    ap.add_argument("--spheres", type=int, default=512)
    ap.add_argument("--soup", type=int, default=125000)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--probes", type=int, default=10000)
    ap.add_argument("--eval-spheres", type=int, default=8192)
    ap.add_argument("--eval-soup", type=int, default=2000000)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--target", default="asinh", choices=("z", "asinh", "rank"))
    ap.add_argument("--save-records", help="write a random subsample of the records (.npz: x, g)")
    ap.add_argument("--save-count", type=int, default=200000)
    args = ap.parse_args(argv)

    def seeds(spec):
        a, _, b = spec.partition("-")
        return list(range(int(a), int(b or a) + 1))

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    x, g = generate_records(seeds(args.train_seeds), args.spheres, args.soup, args.steps, args.probes, log=log)
    if args.save_records:
        keep = np.random.default_rng(0).permutation(len(g))[: args.save_count]
        np.savez_compressed(args.save_records, x=x[keep], g=g[keep])
    model, info = train(x, g, target=args.target, log=log)
    model.save(args.out)
    log(f"saved {args.out}: {json.dumps(info)}")
    for s in seeds(args.eval_seeds):
        r = evaluate(model, s, args.eval_spheres, args.eval_soup, args.eval_steps)
        log(f"eval: {json.dumps(r)}")
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    try:
        main()
    except BVHTestError as e:
        print(f"error: {e}", file=sys.stderr)
        sys.exit(1)
