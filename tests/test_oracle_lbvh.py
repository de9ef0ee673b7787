"""Pins the oracle's Karras LBVH build and refit against a brute-force top-down radix tree.

The brute force splits each range at the highest bit where the first and last (key, index)
differ, which is the radix tree Karras's bottom-up construction must reproduce node for node
(left children are numbered by their last leaf, right children by their first, the root is 0).
"""
import numpy as np
import pytest

import oracle


def brute_tree(keys, key_bits, base=None):
    n = len(keys)
    keys = [int(k) for k in keys]
    left = {}
    right = {}
    rng = {}

    def aug(i):
        li = i - (base[i] if base is not None else 0)
        return (keys[i] << 32) | li

    def build(lo, hi, idx):
        a, b = aug(lo), aug(hi)
        top = (a ^ b).bit_length() - 1
        gamma = lo
        while gamma + 1 <= hi and ((aug(gamma + 1) >> top) & 1) == ((a >> top) & 1):
            gamma += 1
        rng[idx] = (lo, hi)
        if lo == gamma:
            left[idx] = gamma
        else:
            left[idx] = gamma + n
            build(lo, gamma, gamma)
        if gamma + 1 == hi:
            right[idx] = gamma + 1
        else:
            right[idx] = gamma + 1 + n
            build(gamma + 1, hi, gamma + 1)

    if n >= 2:
        build(0, n - 1, 0)
    return left, right, rng


def check_tree(keys, dtype):
    keys = np.sort(np.asarray(keys, dtype))
    n = keys.size
    t = oracle.karras(keys)
    bl, br, brng = brute_tree(keys, 32 if dtype == np.uint32 else 64)
    assert sorted(brng) == list(range(n - 1))
    for i in range(n - 1):
        node = i + n
        assert t["left"][node] == bl[i], (i, keys)
        assert t["right"][node] == br[i]
        assert (t["min"][node], t["max"][node]) == brng[i]
        assert t["parent"][t["left"][node]] == node and t["parent"][t["right"][node]] == node
    assert t["parent"][n] == -1
    return t


@pytest.mark.parametrize("seed", range(12))
def test_karras_random_with_duplicates(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 65))
    hi = int(rng.choice([4, 64, 1 << 30]))
    check_tree(rng.integers(0, hi, n), np.uint32)
    check_tree(rng.integers(0, min(hi, 1 << 62), n, dtype=np.int64).astype(np.uint64), np.uint64)


def test_karras_edge_cases():
    check_tree([5, 5], np.uint32)
    check_tree([1, 2], np.uint32)
    check_tree([0, 0, 0, 0, 0, 0, 0], np.uint32)
    check_tree(np.arange(33), np.uint32)
    check_tree([0, 0xFFFFFFFF, 0xFFFFFFFF], np.uint32)
    check_tree([0, (1 << 63) - 1], np.uint64)


def test_refit_is_union_of_leaf_boxes():
    rng = np.random.default_rng(5)
    n = 300
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, 50, (n, 3))
    codes = oracle.morton30(pos, [0, 0, 0], [50, 50, 50])
    sk, si = oracle.sort_pairs(codes, np.arange(n, dtype=np.uint32))
    t = oracle.karras(sk)
    cms, bounds = oracle.refit(si.astype(np.int32), pos, 0.5, t["left"], t["right"])
    leaf_lo = pos[si, :3] - np.float32(0.5)
    leaf_hi = pos[si, :3] + np.float32(0.5)
    assert np.array_equal(cms, pos[si])
    for node in range(n, 2 * n - 1):
        lo, hi = t["min"][node], t["max"][node]
        assert np.array_equal(bounds[node, :3], leaf_lo[lo:hi + 1].min(0))
        assert np.array_equal(bounds[node, 4:7], leaf_hi[lo:hi + 1].max(0))


def test_forest_subtrees_match_standalone_trees():
    """Scene id in the high key bits + scene-local tie-break: every scene's subtree is the
    standalone tree of that scene with indices shifted by the scene start."""
    rng = np.random.default_rng(6)
    sizes = [17, 1, 40, 2, 23]
    keys, base, local = [], [], []
    start = 0
    for s, m in enumerate(sizes):
        k = np.sort(rng.integers(0, 8, m).astype(np.uint64))   # many duplicates
        local.append(k)
        keys.append((np.uint64(s) << np.uint64(30)) | k)
        base += [start] * m
        start += m
    keys = np.concatenate(keys)
    n = keys.size
    t = oracle.karras(keys, base=np.array(base, np.int32))
    start = 0
    for s, m in enumerate(sizes):
        if m >= 2:
            ts = oracle.karras(local[s])
            # find the node covering exactly this scene
            nodes = [v for v in range(n, 2 * n - 1) if t["min"][v] == start and t["max"][v] == start + m - 1]
            assert len(nodes) == 1
            # map standalone -> forest by ranges
            fmap = {(t["min"][v] - start, t["max"][v] - start): v for v in range(n, 2 * n - 1)
                    if start <= t["min"][v] and t["max"][v] <= start + m - 1}
            for v in range(m, 2 * m - 1):
                fv = fmap[(ts["min"][v], ts["max"][v])]
                for side in ("left", "right"):
                    c = ts[side][v]
                    fc = t[side][fv]
                    if c < m:
                        assert fc == c + start
                    else:
                        assert (t["min"][fc] - start, t["max"][fc] - start) == (ts["min"][c], ts["max"][c])
        start += m
