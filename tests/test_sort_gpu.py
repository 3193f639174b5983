"""The LBVH build's sort alone (bvht_sortKeyIndexPairs), both algorithms,
against numpy's stable order by (key, index).

The reference sorts Morton keys with index values by CUB's stable radix sort
(lib/builder/LBVH_CUDA.cu:185-193); the result is the permutation
np.lexsort((index, key)).  Bar: bit-exact indices and keys.  The key sets
exercise the MSD sort's structure: buckets of the top 16 bits that fill a
window exactly, sit one past the LDS limit, need several split levels, or
hold only equal keys (split down to the index bits).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, LDS = 2048, 4096  # window and in-LDS limit of the MSD sort


def top(bits, b):
    return np.uint64(b) << np.uint64(bits - 16)


def cases(bits):
    rng = np.random.default_rng(bits)
    hi = 1 << bits
    lo_bits = bits - 16
    c = {
        "one": np.array([5]),
        "two_equal": np.array([7, 7]),
        "random_3000": rng.integers(0, hi, 3000, dtype=np.uint64),
        "random_1M": rng.integers(0, hi, 1 << 20, dtype=np.uint64),
        # 256 tiles of 17,408 keys (the most the one-round tiles take on a
        # 256-CU device), then one key more (tiles of 5,632 / 5,120 keys)
        "random_256_wide_tiles": rng.integers(0, hi, 256 * 17408, dtype=np.uint64),
        "random_past_wide_tiles": rng.integers(0, hi, 256 * 17408 + 1, dtype=np.uint64),
        "all_equal_20k": np.full(20000, hi // 3, np.uint64),
        "few_distinct": (rng.integers(0, 3, 50000).astype(np.uint64) << np.uint64(lo_bits + 2)),
        "one_bucket_spread": top(bits, 77) | rng.integers(0, 1 << lo_bits, 60000, dtype=np.uint64),
        "one_bucket_narrow": top(bits, 77) | rng.integers(0, 1 << min(lo_bits, 10), 60000, dtype=np.uint64),
        "window_sized_buckets": np.concatenate([np.full(W, top(bits, 3)), np.full(W, top(bits, 9)),
                                                np.full(W + 1, top(bits, 11)), rng.integers(0, hi, 3000, dtype=np.uint64)]),
        "lds_limit_buckets": np.concatenate([np.full(LDS, top(bits, 3)), np.full(LDS + 1, top(bits, 5)),
                                             rng.integers(0, hi, 5000, dtype=np.uint64)]),
        "sorted": np.sort(rng.integers(0, hi, 200000, dtype=np.uint64)),
        "reversed": np.sort(rng.integers(0, hi, 200000, dtype=np.uint64))[::-1].copy(),
    }
    # clustered like icosphere clusters: dense blobs plus uniform soup
    blobs = [top(bits, int(b)) | rng.integers(0, 1 << max(lo_bits - 6, 1), 20000, dtype=np.uint64)
             for b in rng.integers(0, 1 << 16, 12)]
    c["clusters"] = np.concatenate(blobs + [rng.integers(0, hi, 100000, dtype=np.uint64)])
    rng.shuffle(c["clusters"])
    # runs of equal high bits (24 and up for 63-bit keys) of many lengths,
    # shuffled: the 63-bit group step's groups inside one wave's positions,
    # across its neighbours', past them (galloping search) and longer than
    # 4096 (the long-group sort)
    lens = np.concatenate([rng.integers(1, 64, 300), rng.integers(64, 193, 60), rng.integers(193, 4097, 12),
                           [4096, 4097, 5000]])
    shift = 24 if bits == 63 else 14
    grp = np.repeat(np.arange(len(lens), dtype=np.uint64), lens) << np.uint64(shift)
    mg = grp | rng.integers(0, 1 << shift, len(grp), dtype=np.uint64)
    rng.shuffle(mg)
    c["mixed_groups"] = mg
    return c


CASES = ["one", "two_equal", "random_3000", "random_1M", "random_256_wide_tiles", "random_past_wide_tiles",
         "all_equal_20k", "few_distinct", "one_bucket_spread",
         "one_bucket_narrow", "window_sized_buckets", "lds_limit_buckets", "sorted", "reversed", "clusters",
         "mixed_groups"]


@pytest.mark.parametrize("digits", [16, 8], ids=["msd", "onesweep"])
@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("name", CASES)
def test_sort_matches_stable_order(bvh, name, bits, digits):
    keys = cases(bits)[name].astype(np.uint64 if bits == 63 else np.uint32)
    assert int(keys.max()) < (1 << bits)
    print(f"case {name} n={len(keys)} bits={bits} digits={digits}", flush=True)
    idx, sk = bvh.sort_key_index_pairs(keys, bits, digits)
    ref = np.lexsort((np.arange(len(keys)), keys)).astype(np.uint32)
    assert np.array_equal(idx, ref)
    assert np.array_equal(sk, keys[ref])


def test_sort_rejects_bad_arguments(bvh):
    with pytest.raises(bvh.BVHTestError, match="keyBits"):
        bvh.sort_key_index_pairs(np.zeros(4, np.uint32), 31)
    with pytest.raises(bvh.BVHTestError, match="digitBits"):
        bvh.sort_key_index_pairs(np.zeros(4, np.uint32), 30, 11)
