"""Pins the oracle's integer stages: grid hash, Morton codes, stable sort, cell bounds.

Each check is independent of the oracle's own formula: naive bit interleaving, numpy's stable
argsort, a closed-form hash written with numpy, and set invariants.
"""
import numpy as np
import pytest

import oracle
from diploma_thesis_amd import capi


def naive_interleave(x, y, z, bits):
    """code bit 3k+2 = x bit k, 3k+1 = y bit k, 3k = z bit k (morton3D's xx*4 + yy*2 + zz)."""
    code = np.zeros(np.shape(x), np.uint64)
    for k in range(bits):
        code |= ((np.asarray(x, np.uint64) >> k) & 1) << np.uint64(3 * k + 2)
        code |= ((np.asarray(y, np.uint64) >> k) & 1) << np.uint64(3 * k + 1)
        code |= ((np.asarray(z, np.uint64) >> k) & 1) << np.uint64(3 * k)
    return code


def naive_spread(v, bits):
    out = np.zeros(np.shape(v), np.uint64)
    for k in range(bits):
        out |= ((np.asarray(v, np.uint64) >> k) & 1) << np.uint64(3 * k)
    return out


def test_expand10_exhaustive():
    v = np.arange(1024, dtype=np.uint32)
    assert np.array_equal(oracle.expand10(v).astype(np.uint64), naive_spread(v, 10))
    assert int(oracle.expand10(np.array([1023], np.uint32))[0]) == 0x09249249


def test_expand21_exhaustive():
    v = np.arange(1 << 21, dtype=np.uint64)
    assert np.array_equal(oracle.expand21(v), naive_spread(v, 21))


def test_morton30_known_answers_and_clamps():
    pts = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5], [-3, 2, np.nan],
                    [0.999999, 0.999999, 0.999999]], np.float32)
    c = oracle.morton30_unit(pts)
    assert c[0] == 0
    assert c[1] == 0x3FFFFFFF                       # u >= 1 clamps to 1023 on every axis
    assert c[2] == 1 << 29 and c[3] == 1 << 28 and c[4] == 1 << 27
    assert c[5] == naive_interleave(0, 1023, 0, 10)  # negative -> 0, >1 -> 1023, NaN -> 0
    assert c[6] == 0x3FFFFFFF


def test_morton30_random_vs_naive():
    rng = np.random.default_rng(1)
    u = rng.uniform(-0.1, 1.1, size=(20000, 3)).astype(np.float32)
    q = np.clip(np.minimum(np.maximum(u * np.float32(1024.0), 0), 1023), 0, 1023).astype(np.uint32)
    assert np.array_equal(oracle.morton30_unit(u).astype(np.uint64), naive_interleave(q[:, 0], q[:, 1], q[:, 2], 10))


def test_morton63_random_vs_naive():
    rng = np.random.default_rng(2)
    u = rng.uniform(-0.1, 1.1, size=(20000, 3)).astype(np.float32)
    q = np.minimum(np.maximum(u * np.float32(2097152.0), 0), 2097151).astype(np.uint64)
    assert np.array_equal(oracle.morton63_unit(u), naive_interleave(q[:, 0], q[:, 1], q[:, 2], 21))
    one = oracle.morton63_unit(np.ones((1, 3), np.float32))
    assert int(one[0]) == (1 << 63) - 1


def test_morton_box_normalisation_and_zero_extent():
    pos = np.array([[32, 16, 8, 0], [64, 64, 64, 0], [0, 5, 5, 0]], np.float32)
    c = oracle.morton30(pos, [0, 0, 0], [64, 64, 64])
    ref = oracle.morton30_unit(pos[:, :3] / np.float32(64.0))
    assert np.array_equal(c, ref)
    flat = oracle.morton30(pos, [0, 5, 0], [64, 5, 64])   # zero extent on y: treated as 0
    assert np.array_equal(flat, oracle.morton30_unit(np.stack([pos[:, 0] / 64, 0 * pos[:, 1], pos[:, 2] / 64], 1)))


def _params(grid=(64, 64, 64), origin=(0, 0, 0), cell=1.0):
    p = capi.ArSimParams()
    p.gridSize.x, p.gridSize.y, p.gridSize.z = grid
    p.numCells = grid[0] * grid[1] * grid[2]
    p.worldOrigin.x, p.worldOrigin.y, p.worldOrigin.z = origin
    p.cellSize.x = p.cellSize.y = p.cellSize.z = cell
    p.particleRadius = cell / 2
    return p


def test_grid_hash_closed_form_with_wrap():
    rng = np.random.default_rng(3)
    pos = np.zeros((50000, 4), np.float32)
    pos[:, :3] = rng.uniform(-300, 300, size=(50000, 3)).astype(np.float32)
    for grid in [(64, 64, 64), (256, 128, 32), (4, 4, 4)]:
        p = _params(grid=grid, origin=(-1.0, 2.0, 0.5), cell=1.0)
        h, idx = oracle.calc_hash(pos, p)
        g = np.floor((pos[:, :3] - np.array([-1.0, 2.0, 0.5], np.float32)) / np.float32(1.0)).astype(np.int64)
        g &= np.array(grid, np.int64) - 1                   # two's-complement wrap of negative cells
        ref = (g[:, 2] * grid[1] * grid[0] + g[:, 1] * grid[0] + g[:, 0]).astype(np.uint32)
        assert np.array_equal(h, ref)
        assert np.array_equal(idx, np.arange(50000, dtype=np.uint32))


@pytest.mark.parametrize("n", [0, 1, 2, 17, 4096, 100003])
def test_sort_matches_numpy_stable(n):
    rng = np.random.default_rng(n)
    for keys in [rng.integers(0, 1 << 18, n, dtype=np.uint32), np.full(n, 7, np.uint32),
                 np.arange(n, dtype=np.uint32)[::-1].copy(), rng.integers(0, 4, n, dtype=np.uint32),
                 rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)]:
        vals = np.arange(n, dtype=np.uint32)
        ok, ov = oracle.sort_pairs(keys, vals)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(ok, keys[order]) and np.array_equal(ov, vals[order])
    k64 = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    ok, ov = oracle.sort_pairs(k64, np.arange(n, dtype=np.uint32))
    order = np.argsort(k64, kind="stable")
    assert np.array_equal(ok, k64[order]) and np.array_equal(ov, order.astype(np.uint32))


def test_sort_bit_range():
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 32, 5000, dtype=np.uint64).astype(np.uint32)
    ok, ov = oracle.sort_pairs(keys, np.arange(5000, dtype=np.uint32), 4, 12)
    order = np.argsort((keys >> 4) & 0xFF, kind="stable")
    assert np.array_equal(ov, order.astype(np.uint32))


def test_cell_bounds_invariants():
    rng = np.random.default_rng(4)
    h = np.sort(rng.integers(0, 512, 3000).astype(np.uint32))
    cs, ce = oracle.cell_bounds(h, 512)
    for c in range(512):
        idx = np.nonzero(h == c)[0]
        if idx.size == 0:
            assert cs[c] == 0xFFFFFFFF
        else:
            assert cs[c] == idx[0] and ce[c] == idx[-1] + 1
