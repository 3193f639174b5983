"""Writes the golden fixtures of this directory from the oracle (run from the
repository root: python tests/golden/make_golden.py).

The reference ships no golden vectors or known-answer tests for its GPU path,
so these fixtures are the oracle's outputs, frozen.  They pin the oracle
against later edits (tests/test_golden.py regenerates and compares) and let
the GPU tests compare with fixed data.

  tiny.npz   soup meshes of n = 1..24, 31..33, 63..65, 100 triangles: vertices,
             faces, and for 30- and 63-bit keys the sorted keys, the sorted
             face permutation and the full node array (48-byte BVHNode)
  c0.json    the 65,536-triangle heightfield (256x128 cells, seed 1, heights
             0..0.2): SHA-256 of the mesh, keys, permutation and nodes
             (30-bit LBVH); SAH(1,1) and SAH(1.2,1); the sequential Bittner13
             run (10% of the nodes per step, 50 steps) -- per-step SAH(1,1),
             final SAH(1.2,1) and tree hash; the 1024x1024 primary-ray hits
             (camera pos (0.5,1,1.25), lookAt (0.5,0.1,0.25), up +y, fov 45)
             as a hash of (triId, t) plus the hit count
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

TINY_SIZES = list(range(1, 25)) + [31, 32, 33, 63, 64, 65, 100]
C0_CAMERA = dict(pos=[0.5, 1.0, 1.25], look_at=[0.5, 0.1, 0.25], up=[0.0, 1.0, 0.0], fov=45.0)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tiny_mesh(n):
    """Seeded soup: triangle i around a random centre, edges up to 0.05."""
    rng = np.random.default_rng(1000 + n)
    c = rng.random((n, 3), dtype=np.float32)
    v = (c[:, None, :] + (rng.random((n, 3, 3), dtype=np.float32) - 0.5) * np.float32(0.05)).reshape(-1, 3)
    return v.astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def tiny_fixtures(orc):
    out = {}
    for n in TINY_SIZES:
        v, f = tiny_mesh(n)
        out[f"v{n}"], out[f"f{n}"] = v, f
        for bits in (30, 63):
            keys, sf, nodes, _ = orc.lbvh(v, f, bits)
            out[f"keys{n}_{bits}"] = keys
            out[f"perm{n}_{bits}"] = sf
            out[f"nodes{n}_{bits}"] = nodes.view(np.uint8).reshape(len(nodes), 48)
    return out


def c0_fixture(bt, orc):
    v, f = bt.scene_heightfield(256, 128, 1, 0.2)
    keys, sf, nodes, _ = orc.lbvh(v, f, 30)
    opt, root, sah = orc.bittner_seq(nodes, 0, 10.0, 50)
    rays = orc.gen_rays(1024, 1024, C0_CAMERA["pos"], C0_CAMERA["look_at"], C0_CAMERA["up"], C0_CAMERA["fov"])
    _, hits = orc.trace(rays, nodes, 0, v, f)
    return {
        "mesh": {"vertices": int(len(v)), "faces": int(len(f)), "sha256_vertices": sha(v), "sha256_faces": sha(f)},
        "lbvh30": {"sha256_keys": sha(keys), "sha256_perm": sha(sf), "sha256_nodes": sha(nodes),
                   "sah_1_1": orc.calc_sah(nodes, 0, 1.0, 1.0), "sah_1.2_1": orc.calc_sah(nodes, 0, 1.2, 1.0)},
        "bittner_seq": {"batch_percent": 10.0, "steps": 50, "sah_1_1": [float(x) for x in sah],
                        "final_sah_1.2_1": orc.calc_sah(opt, root, 1.2, 1.0), "root": int(root),
                        "sha256_nodes": sha(opt)},
        "trace1024": {"camera": C0_CAMERA, "hits": int((hits["triId"] != 0xFFFFFFFF).sum()),
                      "sha256_hits": sha(hits)},
    }


def main():
    import bvhtest_amd as bt
    import _oracle as orc

    np.savez_compressed(os.path.join(HERE, "tiny.npz"), **tiny_fixtures(orc))
    with open(os.path.join(HERE, "c0.json"), "w") as fh:
        json.dump(c0_fixture(bt, orc), fh, indent=1, sort_keys=True)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
