"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the checker, never the product path.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

NODE_DTYPE = np.dtype([("bbox", "<f4", (6,)), ("parent", "<u4"), ("left", "<u4"), ("right", "<u4"),
                       ("isLeafFlag", "<u2"), ("isLeft", "<u2"), ("level", "<u2"), ("pad", "<u2"),
                       ("surfaceArea", "<f4")])
RAY_DTYPE = np.dtype([("pos", "<f4", (3,)), ("dir", "<f4", (3,)), ("invDir", "<f4", (3,)),
                      ("sign", "u1", (3,)), ("pad", "u1"), ("px", "<u4"), ("py", "<u4")])
PIXEL_DTYPE = np.dtype([("hit", "u1"), ("diffuse", "u1"), ("intCount", "<u2")])
HIT_DTYPE = np.dtype([("triId", "<u4"), ("t", "<f4")])

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_f = ctypes.c_float

_SIG = {
    "orc_lbvh": (ctypes.c_int, [_vp, _u32, _vp, _u32, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "orc_spread3": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_int]),
    "orc_init_tri": (None, [_vp, _vp, _u32, _vp, _vp]),
    "orc_box_sa": (_f, [_vp]),
    "orc_calc_sah": (ctypes.c_double, [_vp, _u32, _u32, _f, _f]),
    "orc_fix_levels": (_u32, [_vp, _u32, _u32]),
    "orc_validate": (_u32, [_vp, _u32, _u32, _u32, _vp, _vp, _u32, _vp]),
    "orc_gen_rays": (None, [_u32, _u32, _vp, _vp, _vp, _f, _vp]),
    "orc_set_ray": (None, [_vp, _vp, _vp, _u32, _u32]),
    "orc_trace": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _f, _vp, _vp, _vp]),
    "orc_trace_bruteforce": (None, [_vp, _u32, _vp, _vp, _u32, ctypes.c_int, _f, _vp]),
    "orc_trace_bundle": (None, [_vp, _u32, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "orc_refit_leaf_tris": (None, [_vp, _vp, _vp, _vp, _u32]),
    "orc_fix_tree1": (None, [_vp, _u32]),
    "orc_set_stale_refit": (None, [ctypes.c_double, ctypes.c_uint64]),
    "orc_last_run_info": (None, [_vp, _vp]),
    "orc_bittner13_seq": (ctypes.c_int, [_vp, _u32, ctypes.POINTER(_u32), _f, _u32, ctypes.c_int, _vp, _vp]),
    "orc_bittner13_gpuref": (ctypes.c_int, [_vp, _u32, _f, _u32, _u32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                            _vp, _vp]),
    "orc_bn13_priorities": (None, [_vp, _u32, ctypes.c_int, _vp]),
    "orc_bn13_select": (None, [_vp, _u32, _f, _u32, ctypes.c_int, _vp, _vp, _vp]),
    "orc_gpuref_patch": (None, [_vp, _u32, _u32, ctypes.c_int, _u32, _vp]),
    "orc_gpuref_search": (_u32, [_vp, _u32, _u32, ctypes.c_int, ctypes.c_int, _u32, _vp]),
    "orc_seq_patch": (None, [_vp, _u32, ctypes.POINTER(_u32), _u32, _vp]),
    "orc_tree_is_valid": (ctypes.c_int, [_vp, _u32, _u32]),
    "orc_seq_find_node_box": (_u32, [_vp, _u32, _u32, _vp, _f]),
    "orc_lbvh_omp": (ctypes.c_double, [_vp, _vp, _u32, ctypes.c_int, _vp, _vp, ctypes.c_int]),
    "orc_omp_max_threads": (ctypes.c_int, []),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"oracle not built: {ORACLE_PATH} (make -C oracle)")
        L = ctypes.CDLL(ORACLE_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def lbvh(vert, faces, morton_bits=30, topology=0):
    """Returns (sorted_keys u64, sorted_faces u32, nodes, scene[12])."""
    vert = np.ascontiguousarray(vert, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    n = len(faces)
    keys = np.empty(n, np.uint64)
    sf = np.empty(n, np.uint32)
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    scene = np.empty(12, np.float32)
    rc = lib().orc_lbvh(p(vert), len(vert), p(faces), n, morton_bits, topology, p(keys), p(sf), p(nodes), p(scene))
    assert rc == 0
    return keys, sf, nodes, scene


def fix_levels(nodes, root=0):
    return lib().orc_fix_levels(p(nodes), len(nodes), root)


def validate(nodes, root, max_level, vert, faces):
    err = np.zeros(6, np.uint32)
    vert = np.ascontiguousarray(vert, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    total = lib().orc_validate(p(nodes), len(nodes), root, max_level, p(vert), p(faces), len(faces), p(err))
    return total, err


def calc_sah(nodes, root=0, c_inner=1.0, c_leaf=1.0):
    return lib().orc_calc_sah(p(nodes), len(nodes), root, c_inner, c_leaf)


def gen_rays(w, h, pos, look_at, up, fov):
    rays = np.empty(w * h, RAY_DTYPE)
    f3 = lambda x: np.ascontiguousarray(x, np.float32)
    pos, look_at, up = f3(pos), f3(look_at), f3(up)
    lib().orc_gen_rays(w, h, p(pos), p(look_at), p(up), fov, p(rays))
    return rays


def trace(rays, nodes, root, vert, faces, light=(2, 2, 2), culling=True, mode=0, tmin=0.0, skip=None):
    n = len(rays)
    pix = np.empty(n, PIXEL_DTYPE)
    hits = np.empty(n, HIT_DTYPE)
    light = np.ascontiguousarray(light, np.float32)
    vert = np.ascontiguousarray(vert, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    skip = None if skip is None else np.ascontiguousarray(skip, np.uint32)
    lib().orc_trace(p(rays), n, p(nodes), root, p(vert), p(faces), p(light), int(culling), mode, tmin, p(skip),
                    p(pix), p(hits))
    return pix, hits


def trace_bruteforce(rays, vert, faces, culling=True, tmin=0.0):
    hits = np.empty(len(rays), HIT_DTYPE)
    vert = np.ascontiguousarray(vert, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    lib().orc_trace_bruteforce(p(rays), len(rays), p(vert), p(faces), len(faces), int(culling), tmin, p(hits))
    return hits


def trace_bundle(rays, w, h, nodes, root, vert, faces, light=(2, 2, 2)):
    """kTraceRayBundle: 8x8 tiles of rays walking one node sequence."""
    pix = np.empty(w * h, PIXEL_DTYPE)
    hits = np.empty(w * h, HIT_DTYPE)
    light = np.ascontiguousarray(light, np.float32)
    vert = np.ascontiguousarray(vert, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    lib().orc_trace_bundle(p(rays), w, h, p(nodes), root, p(vert), p(faces), p(light), p(pix), p(hits))
    return pix, hits


def refit_leaf_tris(nodes, vert, faces, ids):
    """kRefitLeafBVHTris on a copy of the nodes (surfaceArea kept)."""
    nodes = nodes.copy()
    ids = np.ascontiguousarray(ids, np.uint32)
    lib().orc_refit_leaf_tris(p(nodes), p(np.ascontiguousarray(vert, np.float32)),
                              p(np.ascontiguousarray(faces, np.uint32)), p(ids), len(ids))
    return nodes


def fix_tree1(nodes):
    """kFixTree1 on a copy: every inner node from its children."""
    nodes = nodes.copy()
    lib().orc_fix_tree1(p(nodes), len(nodes))
    return nodes


def bittner_seq(nodes, root, batch_percent, steps, sort=True):
    nodes = nodes.copy()
    r = _u32(root)
    sah = np.empty(steps + 1, np.float64)
    ov = np.zeros(1, np.uint64)
    rc = lib().orc_bittner13_seq(p(nodes), len(nodes), ctypes.byref(r), batch_percent, steps, int(sort), p(sah), p(ov))
    assert rc == 0
    return nodes, r.value, sah


def bittner_gpuref(nodes, batch_percent, num_chunks, steps, offset_access=False, search=0, chunk_trees=False,
                   alt_sort=False):
    """search: 0 = the reference's findNode1 (heap), 1 = replay of the device
    wave search (frontier 1024), 2 = the reference's findNode2 (altFindNode,
    16 slots).  alt_sort: the reference's threshold selection.
    stats: skipped (lock conflicts), dropped, contained, overflows, runaway
    searches.  With chunk_trees the tree after every chunk is returned too
    (array of shape (steps * num_chunks, N))."""
    nodes = nodes.copy()
    sah = np.empty(steps + 1, np.float64)
    stats = np.zeros(5, np.uint64)
    trees = np.zeros((steps * num_chunks, len(nodes)), NODE_DTYPE) if chunk_trees else None
    rc = lib().orc_bittner13_gpuref(p(nodes), len(nodes), batch_percent, num_chunks, steps, int(offset_access),
                                    int(search), int(alt_sort), p(trees) if chunk_trees else None, p(sah), p(stats))
    if chunk_trees:
        return rc, nodes, sah, stats, trees
    return rc, nodes, sah, stats


def bn13_priorities(nodes, seq=False):
    out = np.empty(len(nodes), np.float32)
    lib().orc_bn13_priorities(p(nodes), len(nodes), int(seq), p(out))
    return out


def bn13_select(nodes, batch_percent, num_chunks=1, mode=0):
    """mode 0: GPU sort, 1: GPU altSort, 2: sequential sorted, 3: sequential
    unsorted.  Returns (order, used)."""
    out = np.empty(len(nodes), np.uint32)
    cnt, used = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
    lib().orc_bn13_select(p(nodes), len(nodes), batch_percent, num_chunks, mode, p(out), p(cnt), p(used))
    return out[: int(cnt[0])], int(used[0])


def gpuref_patch(nodes, node, search=0, alt_slots=16):
    """One GPU-optimizer patch applied and refit.  Returns (nodes, info):
    info = [status, best1, best2, ins1, ins2, unused1, unused2, grandparent]."""
    nodes = nodes.copy()
    info = np.zeros(8, np.uint32)
    lib().orc_gpuref_patch(p(nodes), len(nodes), node, search, alt_slots, p(info))
    return nodes, info


def gpuref_search(nodes, node, stage=0, search=0, alt_slots=16):
    """(insertion point, its induced cost) of the first (stage 0) or second
    (stage 1) reinsertion after removing `node`."""
    cost = np.zeros(1, np.float32)
    best = lib().orc_gpuref_search(p(nodes), len(nodes), node, stage, search, alt_slots, p(cost))
    return best, float(cost[0])


def seq_patch(nodes, root, node):
    """One removal + both reinsertions of the sequential optimizer.
    Returns (nodes, root, info = [status, ins1, ins2, unused1, unused2])."""
    nodes = nodes.copy()
    r = _u32(root)
    info = np.zeros(5, np.uint32)
    lib().orc_seq_patch(p(nodes), len(nodes), ctypes.byref(r), node, p(info))
    return nodes, r.value, info


def set_stale_refit(prob, seed=1):
    """Failure model of bittner_gpuref's per-chunk refit (prob 0: off)."""
    lib().orc_set_stale_refit(prob, seed)


def last_run_info():
    """(stale child reads, longest search in steps) of the last bittner_gpuref run."""
    a, b = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    lib().orc_last_run_info(p(a), p(b))
    return int(a[0]), int(b[0])


def tree_is_valid(nodes, root=0):
    return bool(lib().orc_tree_is_valid(p(nodes), len(nodes), root))


def lbvh_omp(vert, faces, morton_bits=30, threads=0):
    vert = np.ascontiguousarray(vert, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    n = len(faces)
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    sf = np.empty(n, np.uint32)
    secs = lib().orc_lbvh_omp(p(vert), p(faces), n, morton_bits, p(nodes), p(sf), threads)
    return secs, nodes, sf


def seq_find_node_box(nodes, root, box6, sa):
    box6 = np.ascontiguousarray(box6, np.float32)
    return lib().orc_seq_find_node_box(p(nodes), len(nodes), root, p(box6), sa)
