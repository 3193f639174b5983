"""Benchmark of the MI355X BVH pipeline on the ~4M-triangle scene.

Workload (one GPU): 2,048 seeded icospheres (subdivision 2-5) + 500,000
soup triangles (4.24M triangles, fp32), 63-bit Morton LBVH build + refit;
then 1920x1080 primary rays + one cosine-weighted diffuse bounce; then a
short plain Bittner13 optimisation (1% of the nodes per step).

A "step" is one LBVH build of the whole scene, timed like LBVH::runImpl
(lib/builder/LBVH.cpp:41-50: encode -> sort -> leaves/topology -> refit ->
sync) with the mesh already resident in HBM.  `value` is Mtri/s of the
build.  Per-kernel durations come from HIP events on the library's stream;
the dominant kernel's algorithmic bytes per launch over its mean duration
give the roofline figure.  The CPU baseline is the oracle's OpenMP LBVH
(oracle/oracle_baseline.cpp) on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LBVH Mtri/s + Mrays/s; time to reach plain-Bittner SAH (learned vs plain)"
HBM_PEAK_GBS = 8000.0
CAM = dict(pos=(0.5, 0.5, 2.0), look_at=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov=45.0)


SORT_STAGES = ("k_msdScan", "k_msdScatter", "k_msdWindows", "k_msdSplit", "k_msdJobs", "k_onesweep", "k_splitRank",
               "k_splitLong")


def kernel_bytes(name, n, nv, morton_bits, tile_log2, pass_idx=0, passes=0):
    """Algorithmic HBM bytes of one launch, per the byte model of the build
    (n triangles, nv vertices, each vertex counted once):
    scene box: faces 12 n + vertices 12 nv (from the triangles) or vertices
    12 nv (from every vertex, the full build's pass); Morton: faces 12 n +
    vertices 12 nv plus the key write; MSD scatter: key read, key + index written; onesweep pass: key +
    value read and written (first pass reads no value, last pass writes no
    key; 63-bit keys: the 5 passes over bits 24-62, all writing keys); group
    rank: key 8 + index 4 read, index 4 written (each group member read once
    per group, in cache; the index-based topology built here needs no sorted
    keys, so none are written).  63-bit keys of a full build below 2^24
    triangles that fit one round of wide sort tiles are packed (the key's
    low 24 bits replaced by the index): Morton writes the packed key 8 n and
    the low bits 4 n, a onesweep pass reads and writes the 8-byte words only,
    and the group rank reads the words 8 n and the low bits 4 n and writes
    the index 4 n.  Level-0 build: sorted index 4 n + faces 12 n + vertices 12 nv read,
    48-byte leaves (n) and the inner nodes below the top tiles
    (n - ceil(n / 2^tile_log2)) written.  The MSD window and big-bucket sorts
    move a data-dependent share of the pairs, so they have no byte model."""
    kb = 8 if morton_bits == 63 else 4
    packed = morton_bits == 63 and n < (1 << 24) and n <= 256 * 17408
    if name == "k_sceneBox":
        return 12 * n + 12 * nv
    if name == "k_vertBox":
        return 12 * nv
    if name in ("k_morton", "k_mortonMsd"):
        return 12 * n + 12 * nv + kb * n + (4 * n if packed and name == "k_morton" else 0)
    if name == "k_msdScatter":
        return kb * n + (kb + 4) * n
    if name.startswith("k_onesweep"):
        if packed:
            return 16 * n
        keep_last = morton_bits == 63  # 63-bit keys: the group step needs the sorted keys
        rd = kb + (0 if pass_idx == 0 else 4)
        wr = (0 if pass_idx == passes - 1 and not keep_last else kb) + 4
        return n * (rd + wr)
    if name == "k_splitRank":
        return 8 * n + 4 * n + 4 * n
    if name == "k_buildLevel0":
        inner = n - -(-n // (1 << tile_log2))
        return 16 * n + 12 * nv + 48 * (n + inner)
    return None


def profile_build(args):
    """Counter runs (rocprofv3 --pmc): the timed LBVH builds only, then three
    calibration kernels of known bytes whose counters per byte scale the
    build's FETCH_SIZE / WRITE_SIZE to bytes (libbvht_calibrate.so): a copy
    of 256 MiB (reads and writes 256 MiB), a streaming read (a sum over
    256 MiB) and a gather (every 128-byte row of 256 MiB read once, in a
    random order given by 4-byte row indices, and written contiguously)."""
    import ctypes

    import numpy as np

    import bvhtest_amd as bt

    L = bt.lib()
    v, f = bt.scene_spheres_soup(2048, 500000, 2)
    mesh = bt.DeviceMesh(v, f)
    builder = bt.LBVHBuilder(mesh, args.morton_bits, digit_bits=args.digit_bits)
    tree = builder.allocate_bvh()
    for _ in range(args.steps):
        builder.build(tree)
    L.LBVH_doCUDASyc()
    cal = ctypes.CDLL(os.path.join(ROOT, "bvhtest_amd", "lib", "libbvht_calibrate.so"))
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    cal.calib_copy.argtypes, cal.calib_copy.restype = [vp, vp, sz], ctypes.c_bool
    cal.calib_read.argtypes, cal.calib_read.restype = [vp, sz, vp], ctypes.c_bool
    cal.calib_gather.argtypes, cal.calib_gather.restype = [vp, vp, sz, vp], ctypes.c_bool
    nbytes = 256 << 20
    num_rows = nbytes // 128
    a, b = bt.DeviceBuffer(nbytes), bt.DeviceBuffer(nbytes)
    partial = bt.DeviceBuffer(2048 * 4)
    rows = np.random.default_rng(7).permutation(num_rows).astype(np.uint32)
    rows_dev = bt.DeviceBuffer(rows.nbytes).upload(rows)  # every index < num_rows: reads stay in `a`
    for _ in range(3):
        if not cal.calib_copy(b.ptr, a.ptr, nbytes):
            raise SystemExit("calibration copy failed")
    for _ in range(3):
        if not cal.calib_read(a.ptr, nbytes, partial.ptr):
            raise SystemExit("calibration read failed")
    for _ in range(3):
        if not cal.calib_gather(a.ptr, rows_dev.ptr, num_rows, b.ptr):
            raise SystemExit("calibration gather failed")
    print(json.dumps({"profile_build": True, "steps": args.steps, "copy_bytes": nbytes, "read_bytes": nbytes,
                      "gather_rows": num_rows, "gather_row_bytes": 128,
                      "gather_index_bytes": int(rows.nbytes)}), flush=True)


def profile_batch(args):
    """Kernel-trace runs of the batched build: the 512-mesh batch built
    `steps` times without graph capture (tile 2^10 and 8-bit digits, the
    autotuner's choice on MI355X), so every kernel shows up under its own
    name."""
    import bvhtest_amd as bt

    L = bt.lib()
    v, f, fc, vc = bt.scene_batch(512, 1000, 1000, 500000)
    b = bt.BatchBuilder(v, f, fc, vc, tile_log2=10, radix_bits=8)
    for _ in range(args.steps):
        b.build(use_graph=False)
    L.LBVH_doCUDASyc()
    print(json.dumps({"profile_batch": True, "steps": args.steps, "triangles": int(fc.sum())}), flush=True)


def measured_traffic(kernel, bits):
    """HBM bytes per launch of `kernel` from the committed counter runs
    (profiles/lbvh63_4M_traffic.json: FETCH_SIZE and WRITE_SIZE, calibrated
    with a 256 MiB device copy); they cover the 63-bit workload only."""
    path = os.path.join(ROOT, "profiles", "lbvh63_4M_traffic.json")
    if bits != 63 or not os.path.exists(path):
        return None
    with open(path) as fh:
        entry = json.load(fh).get("kernels", {}).get(kernel)
    return entry["traffic_bytes"] if entry else None


def run_bench(args):
    import bvhtest_amd as bt

    L = bt.lib()
    if L.bvht_deviceCount() < 1:
        raise SystemExit("bench.py needs a HIP device")
    if args.profile_build:
        return profile_build(args)
    if args.profile_batch:
        return profile_batch(args)
    t0 = time.time()
    v, f = bt.scene_spheres_soup(2048, 500000, 2)
    n = len(f)
    print(f"# scene: {n} triangles, {len(v)} vertices ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    mesh = bt.DeviceMesh(v, f)
    bits = args.morton_bits
    builder = bt.LBVHBuilder(mesh, bits, digit_bits=args.digit_bits)
    tree = builder.allocate_bvh()

    # warmup
    for _ in range(args.warmup):
        builder.build(tree)
    L.LBVH_doCUDASyc()
    # timed region: K builds, stream-synchronised on both sides
    t0 = time.perf_counter()
    for _ in range(args.steps):
        builder.build(tree)
    L.LBVH_doCUDASyc()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    mtri = n * args.steps / wall / 1e6
    print(f"# build: {ms_per_step:.3f} ms/step, {mtri:.1f} Mtri/s", file=sys.stderr, flush=True)

    # per-kernel timing (separate, instrumented builds)
    L.bvht_LBVH_setTiming(True)
    per = {}
    reps = max(3, min(args.steps, 10))
    for _ in range(reps):
        builder.build(tree)
        buf = (ctypes.c_float * 32)()
        k = L.bvht_LBVH_stageTimes(buf, 32)
        pass_idx = 0
        for i in range(k):
            name = L.bvht_LBVH_stageName(i).decode()
            key = name
            if name == "k_onesweep":
                key = f"k_onesweep[{pass_idx}]"
                pass_idx += 1
            per.setdefault(key, []).append(buf[i])
    passes = 5 if bits == 63 else 4  # onesweep digit passes (63-bit keys: bits 24-62)
    # the digit passes again as one timed entry: events bracket every kernel
    # above, which adds their overhead to each short pass; here there is no
    # event between the passes, and each pass gets the mean of the group
    L.bvht_LBVH_setTimingPassGroups(True)
    grouped = []
    for _ in range(reps):
        builder.build(tree)
        buf = (ctypes.c_float * 32)()
        k = L.bvht_LBVH_stageTimes(buf, 32)
        grouped += [buf[i] for i in range(k) if L.bvht_LBVH_stageName(i).decode() == "k_onesweep"]
    L.bvht_LBVH_setTimingPassGroups(False)
    L.bvht_LBVH_setTiming(False)
    stages = {k: float(np.mean(x)) for k, x in per.items()}
    pass_events_ms = {k: round(v, 4) for k, v in stages.items() if k.startswith("k_onesweep[")}
    if len(grouped) == reps and len(pass_events_ms) == passes:
        for k in pass_events_ms:
            stages[k] = float(np.mean(grouped)) / passes
    stages["sort_total"] = sum(x for k, x in stages.items() if k.startswith(SORT_STAGES))
    nv = len(v)
    tile_log2 = builder.mem.opts.tileLog2

    def nbytes(name):
        if name.startswith("k_onesweep["):
            p = int(name[len("k_onesweep["):-1])
            return kernel_bytes("k_onesweep", n, nv, bits, tile_log2, p, passes)
        return kernel_bytes(name, n, nv, bits, tile_log2)

    # dominant kernel: the modelled kernel with the most time per build
    # (launches of one kernel summed, e.g. the onesweep passes); its figure
    # is the mean launch: mean algorithmic bytes over mean duration
    family = {}
    for name, ms in stages.items():
        if name == "sort_total" or not nbytes(name):
            continue
        fam = name.split("[")[0]
        family.setdefault(fam, []).append((ms, nbytes(name)))
    dom = max(family, key=lambda k: sum(ms for ms, _ in family[k]))
    launches = family[dom]
    dom_ms = sum(ms for ms, _ in launches) / len(launches)
    dom_bytes = sum(b for _, b in launches) / len(launches)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    label = dom if len(launches) == 1 else f"{dom} (mean of {len(launches)} launches)"
    roofline = {"bound": "hbm", "kernel": label, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dom, bits), "bytes_per_launch": int(dom_bytes),
                "ms_per_launch": round(dom_ms, 4)}
    # every modelled kernel the same way, most time per build first
    roofline_kernels = {}
    for fam in sorted(family, key=lambda k: -sum(ms for ms, _ in family[k])):
        ls = family[fam]
        f_ms = sum(ms for ms, _ in ls) / len(ls)
        f_bytes = sum(b for _, b in ls) / len(ls)
        f_gbs = f_bytes / (f_ms * 1e-3) / 1e9
        roofline_kernels[fam] = {"launches": len(ls), "ms_per_build": round(sum(ms for ms, _ in ls), 4),
                                 "ms_per_launch": round(f_ms, 4), "bytes_per_launch": int(f_bytes),
                                 "gbs": round(f_gbs, 1), "frac": round(f_gbs / HBM_PEAK_GBS, 4),
                                 "traffic": measured_traffic(fam, bits)}
    print(f"# stages (ms): {json.dumps({k: round(x, 4) for k, x in stages.items()})}", file=sys.stderr, flush=True)

    # per-stage and end-to-end fractions of the HBM peak: modelled bytes over
    # the stage's measured time (kernels without a byte model add time only)
    stage_groups = {"encode": ("k_sceneBox", "k_vertBox", "k_morton"), "sort": SORT_STAGES, "build": ("k_buildLevel",)}
    stage_roofline = {}
    total_bytes = 0
    for g, prefixes in stage_groups.items():
        names = [k for k in stages if k != "sort_total" and k.startswith(prefixes)]
        g_ms = sum(stages[k] for k in names)
        g_bytes = sum(nbytes(k) or 0 for k in names)
        total_bytes += g_bytes
        if g_ms > 0:
            stage_roofline[g] = {"ms": round(g_ms, 4), "bytes": int(g_bytes),
                                 "gbs": round(g_bytes / (g_ms * 1e-3) / 1e9, 1),
                                 "frac": round(g_bytes / (g_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    e2e = total_bytes / (ms_per_step * 1e-3) / 1e9
    stage_roofline["end_to_end"] = {"ms": round(ms_per_step, 4), "bytes": int(total_bytes), "bytes_per_tri": round(total_bytes / n, 1),
                                    "gbs": round(e2e, 1), "frac": round(e2e / HBM_PEAK_GBS, 4)}
    print(f"# stage roofline: {json.dumps(stage_roofline)}", file=sys.stderr, flush=True)

    # primary + bounce rays
    w, h = 1920, 1080
    nrays = w * h
    rays_dev = bt.generate_rays_device(w, h, **CAM)
    hits_dev = bt.DeviceBuffer(nrays * bt.HIT_DTYPE.itemsize)
    img_dev = bt.DeviceBuffer(nrays * bt.PIXEL_DTYPE.itemsize)
    brays_dev = bt.DeviceBuffer(nrays * bt.RAY_DTYPE.itemsize)
    skip_dev = bt.DeviceBuffer(nrays * 4)
    bhits_dev = bt.DeviceBuffer(nrays * bt.HIT_DTYPE.itemsize)
    light = bt.vec3((2.0, 2.0, 2.0))

    def trace_once():
        L.generateRays(rays_dev.ptr, w, h, bt.vec3(CAM["pos"]), bt.vec3(CAM["look_at"]), bt.vec3(CAM["up"]), CAM["fov"])
        L.bvht_traceRays(rays_dev.ptr, nrays, tree.ptr.nodes, 0, mesh.raw, light, 0, 0.0, None, hits_dev.ptr, img_dev.ptr)
        L.bvht_makeBounceRays(rays_dev.ptr, hits_dev.ptr, nrays, mesh.raw, 1, brays_dev.ptr, skip_dev.ptr)
        L.bvht_traceRays(brays_dev.ptr, nrays, tree.ptr.nodes, 0, mesh.raw, light, 1, 0.0, skip_dev.ptr, bhits_dev.ptr,
                         None)

    trace_once()
    L.tracerDoCudaSync()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        trace_once()
    L.tracerDoCudaSync()
    trace_s = (time.perf_counter() - t0) / reps
    hits = hits_dev.download(bt.HIT_DTYPE, nrays)
    bhits = bhits_dev.download(bt.HIT_DTYPE, nrays)
    mrays = 2 * nrays / trace_s / 1e6
    trace = {"mrays_per_s": round(mrays, 1), "primary_rays": nrays, "bounce_rays": nrays,
             "ms_per_frame": round(trace_s * 1e3, 3),
             "primary_hit_frac": round(float((hits["triId"] != 0xFFFFFFFF).mean()), 4),
             "bounce_hit_frac": round(float((bhits["triId"] != 0xFFFFFFFF).mean()), 4)}
    print(f"# trace: {trace}", file=sys.stderr, flush=True)

    # plain Bittner13 (1% of the nodes per step), bounded number of steps
    bittner = None
    if args.bittner_steps > 0:
        btree = builder.allocate_bvh()
        builder.build(btree)
        L.LBVH_doCUDASyc()
        opt = bt.Bittner13GPU(maxNumStepps=args.bittner_steps, numChunks=32, batchPercent=1.0)
        sah0 = bt.calc_sah(btree, 1.2, 1.0)
        opt.setup(btree, n)
        t0 = time.perf_counter()
        for _ in range(args.bittner_steps):
            opt.step(btree)
        L.doCudaDevSync()
        bt_s = time.perf_counter() - t0
        sah1 = bt.calc_sah(btree, 1.2, 1.0)
        bittner = {"steps": args.bittner_steps, "ms_per_step": round(bt_s * 1e3 / args.bittner_steps, 3),
                   "sah_1.2_1.0_initial": round(sah0, 4), "sah_1.2_1.0_final": round(sah1, 4),
                   "stats": opt.stats()}
        opt.free()
        btree.free()
        print(f"# bittner: {bittner}", file=sys.stderr, flush=True)

    batch = None if args.no_batch else batch_bench(bt, L)
    learned = None if args.no_learned else learned_bench(bt)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(v, f, bits, tree.download())

    out = {"metric": METRIC, "value": round(mtri, 2), "unit": "Mtri/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "icosphere clusters + triangle soup, LBVH build + refit, 1920x1080 primary + "
                                  "1 diffuse bounce", "triangles": n, "spheres": 2048, "soup": 500000,
                      "morton_bits": bits, "topology": "reference (index Karras)", "seed": 2,
                      "sort": ("MSD: 16-bit buckets, in-LDS sorts by (key, index)" if builder.mem.opts.digitBits == 16
                               else "onesweep over bits 24-62 (8-bit digits), then (key, index) rank in groups of "
                                    "equal top bits" if bits == 63 else "onesweep, 8-bit digits")},
           "roofline": roofline, "roofline_kernels": roofline_kernels, "roofline_stages": stage_roofline, "stages_ms": {k: round(x, 4) for k, x in stages.items()},
           "onesweep_event_ms": pass_events_ms,
           "trace": trace, "bittner": bittner, "batched": batch, "learned_priority": learned, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def batch_bench(bt, L, reps=5):
    """512 meshes of log-uniform size in [1K, 500K] triangles: one segmented
    build replayed from a hipGraph (tile size and sort digit width
    autotuned) vs. 512 sequential single-mesh builds of the same meshes
    (30-bit keys both)."""
    t0 = time.time()
    v, f, fc, vc = bt.scene_batch(512, 1000, 1000, 500000)
    total = int(fc.sum())
    print(f"# batch scene: {len(fc)} meshes, {total} triangles ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    b = bt.BatchBuilder(v, f, fc, vc)
    best, ms = b.autotune()
    b.build(use_graph=True)
    L.LBVH_doCUDASyc()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.build(use_graph=True)
    L.LBVH_doCUDASyc()
    graph_ms = (time.perf_counter() - t0) * 1e3 / reps
    # the same meshes one by one, mesh data and trees already on the device
    meshes, builders, trees = [], [], []
    for i in range(len(fc)):
        fo, vo, _ = b.offsets(i)
        mv = v[vo:vo + int(vc[i])]
        mf = f[fo:fo + int(fc[i])]
        m = bt.DeviceMesh(mv, mf)
        lb = bt.LBVHBuilder(m, 30)
        meshes.append(m)
        builders.append(lb)
        trees.append(lb.allocate_bvh())
    for lb, tr in zip(builders, trees):
        lb.build(tr)
    L.LBVH_doCUDASyc()
    t0 = time.perf_counter()
    for _ in range(reps):
        for lb, tr in zip(builders, trees):
            lb.build(tr)
    L.LBVH_doCUDASyc()
    seq_ms = (time.perf_counter() - t0) * 1e3 / reps
    for lb in builders:
        lb.free()
    # modelled bytes of the whole batched build (u32 keys, one pass per
    # digit): mesh boxes 12 V (every vertex once); Morton 12 n + 12 V + 4 n key + 24 n
    # leaf-box record; each sort pass key + value read and written (the
    # first reads no value, the last writes no key); level-0 build sorted
    # index 4 n + record 24 n + 48-byte leaves and inner nodes below the top
    # tiles
    nv_all = int(vc.sum())
    passes = 4 if int(b.mem.radixBits) == 8 else 3
    sort_b = sum(total * ((4 + (0 if p == 0 else 4)) + ((0 if p == passes - 1 else 4) + 4)) for p in range(passes))
    inner = sum(int(x) - -(-int(x) // (1 << int(best))) for x in fc)
    batch_bytes = 12 * nv_all + (40 * total + 12 * nv_all) + sort_b + (28 * total
                                                                                      + 48 * (total + inner))
    batch_gbs = batch_bytes / (graph_ms * 1e-3) / 1e9
    out = {"meshes": int(len(fc)), "triangles": total, "tile_log2": int(best), "radix_bits": int(b.mem.radixBits),
           "roofline_end_to_end": {"bytes_per_tri": round(batch_bytes / total, 1), "gbs": round(batch_gbs, 1),
                                   "frac": round(batch_gbs / HBM_PEAK_GBS, 4)},
           "autotune_ms": {"radix8_tiles_8_to_11": [round(x, 3) for x in ms[:4]],
                           "radix10_tiles_8_to_11": [round(x, 3) for x in ms[4:]]},
           "graph_ms": round(graph_ms, 3),
           "graph_mtri_per_s": round(total / graph_ms / 1e3, 1), "sequential_ms": round(seq_ms, 3),
           "sequential_mtri_per_s": round(total / seq_ms / 1e3, 1), "speedup": round(seq_ms / graph_ms, 2)}
    print(f"# batched: {out}", file=sys.stderr, flush=True)
    return out


def learned_bench(bt, seed=200, plain_steps=100):
    """Time to reach the plain optimizer's final SAH(1.2, 1) after 100 steps
    within 0.5%: learned priority vs. plain, on one held-out ~16.9M-triangle
    scene (8,192 icospheres + 2M soup triangles)."""
    path = os.path.join(ROOT, "bvhtest_amd", "data", "priority_mlp.npy")
    if not os.path.exists(path):
        return {"status": "no trained weights at bvhtest_amd/data/priority_mlp.npy"}
    from bvhtest_amd import priority_training as pt

    model = bt.PriorityModel.load(path)
    r = pt.evaluate(model, seed, 8192, 2000000, plain_steps)
    r["weights"] = "bvhtest_amd/data/priority_mlp.npy"
    print(f"# learned priority: {r}", file=sys.stderr, flush=True)
    return r


def host_cpu():
    """Model name (as lscpu prints it, from /proc/cpuinfo) and the number of
    logical CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 0
    return model, avail


def cpu_baseline(v, f, bits, nodes, seconds=12.0):
    """Oracle OpenMP LBVH timed on this host: repeated full builds of the same
    scene for about `seconds` with every host thread, then about a third of
    that with one thread.  `value` is the all-thread rate.  Beside it, on the
    same tree: the oracle tracer (the reference CPUTracer's traversal, OpenMP
    dynamic 512) on every 8th primary ray of the 1920x1080 frame, and one step
    of the sequential Bittner13 (1% of the nodes, single-threaded search)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc

    def sample(threads, budget):
        orc.lbvh_omp(v[:3000], f[:1000] % 3000, bits, threads)  # warm the thread pool
        reps, secs = 0, 0.0
        while secs < budget and reps < 1000:
            s, _, _ = orc.lbvh_omp(v, f, bits, threads)
            secs += s
            reps += 1
        return reps, secs

    threads = orc.lib().orc_omp_max_threads()
    reps1, secs1 = sample(1, seconds / 3)
    reps, secs = sample(threads, seconds)  # last: leaves every thread enabled for the tracer below
    mtri = len(f) * reps / secs / 1e6
    print(f"# cpu baseline: {reps} builds in {secs:.2f}s on {threads} threads; {reps1} in {secs1:.2f}s on 1",
          file=sys.stderr, flush=True)
    rays = orc.gen_rays(1920, 1080, CAM["pos"], CAM["look_at"], CAM["up"], CAM["fov"])
    sample_rays = np.ascontiguousarray(rays[::8])
    t0 = time.perf_counter()
    orc.trace(sample_rays, nodes, 0, v, f)
    trace_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.bittner_seq(nodes, 0, 1.0, 1)
    bn_s = time.perf_counter() - t0
    print(f"# cpu baseline: trace {len(sample_rays) / trace_s / 1e6:.2f} Mrays/s, sequential Bittner13 {bn_s:.2f} s/step",
          file=sys.stderr, flush=True)
    model, avail = host_cpu()
    return {"value": round(mtri, 3), "unit": "Mtri/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_logical_cpus": avail,
            "sample": f"{reps} full builds of the same {len(f)}-triangle scene ({bits}-bit keys, encode to refit), "
                      f"oracle OpenMP LBVH, {secs:.2f}s of build time",
            "single_thread": {"value": round(len(f) * reps1 / secs1 / 1e6, 3), "unit": "Mtri/s",
                              "sample": f"{reps1} builds, {secs1:.2f}s"},
            "trace": {"value": round(len(sample_rays) / trace_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                      "sample": f"every 8th primary ray of the 1920x1080 frame ({len(sample_rays)} rays), "
                                f"oracle tracer on the same tree, {trace_s:.2f}s"},
            "bittner_seq": {"value": round(bn_s * 1e3, 1), "unit": "ms/step", "cores": 1,
                            "sample": "one step of the sequential Bittner13 (1% of the nodes, selection by the "
                                      "reference priority) on the same tree"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--morton-bits", type=int, default=63, choices=[30, 63])
    ap.add_argument("--digit-bits", type=int, default=0, choices=[0, 8, 16],
                    help="LBVH sort: 8 (default, 0) onesweep, 16 MSD")
    ap.add_argument("--bittner-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--profile-build", action="store_true", help="builds + a 256 MiB copy only (counter runs)")
    ap.add_argument("--profile-batch", action="store_true", help="batched builds only (kernel-trace runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); runs on more GPUs are not "
              f"implemented", file=sys.stderr)
        raise SystemExit(2)
    run_bench(args)


if __name__ == "__main__":
    main()
