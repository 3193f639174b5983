"""Writes c2_scale.json: the reference GPU Bittner13 algorithm at its own
scale (run from the repository root: python tests/golden/make_c2_scale.py;
about ten minutes on 8 cores).

Workload: the ~4.24M-triangle scene of 2,048 seeded icospheres + 500,000
soup triangles (seed 2), its 63-bit LBVH (the oracle's build, which the
device build equals byte for byte), then 100 steps of the GPU optimizer
(lib/builder/Bittner13CUDA.cu, bn13_doAlgorithmStep, :768-772) with 1% of the
nodes per step in 32 chunks and the defaults of Bittner13GPU.hpp:28-37:
findNode1's heap search and kCheckConflicts serialised in thread order
(oracle/oracle_bittner.cpp, orc_bittner13_gpuref, search 0).  The fixture
holds the per-step SAH (CUDAcalcSAH's (sum of areas - root area) / root
area), the conflict statistics and a hash of the starting tree.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

SCENE = dict(spheres=2048, soup=500000, seed=2)
RUN = dict(batch_percent=1.0, num_chunks=32, steps=100)


def main():
    import _oracle as orc
    import bvhtest_amd as bt

    v, f = bt.scene_spheres_soup(SCENE["spheres"], SCENE["soup"], SCENE["seed"])
    _, _, nodes, _ = orc.lbvh(v, f, 63)
    t0 = time.time()
    rc, out, sah, stats = orc.bittner_gpuref(nodes, RUN["batch_percent"], RUN["num_chunks"], RUN["steps"], search=0)
    assert rc == 0
    res = dict(scene=SCENE, run=RUN, triangles=int(len(f)), nodes=int(len(nodes)),
               start_nodes_sha256=hashlib.sha256(nodes.tobytes()).hexdigest(),
               sah=[float(x) for x in sah], stats=dict(zip(["skipped", "dropped", "contained", "overflows", "runaway"],
                                                           [int(x) for x in stats])),
               final_tree_sha256=hashlib.sha256(out.tobytes()).hexdigest(), oracle_seconds=round(time.time() - t0, 1))
    with open(os.path.join(HERE, "c2_scale.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print(f"SAH {sah[0]:.4f} -> {sah[-1]:.4f} in {res['oracle_seconds']} s; stats {res['stats']}")


if __name__ == "__main__":
    main()
