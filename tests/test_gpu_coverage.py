"""GPU parity for the remaining rows: AABB, the dynamic Morton box, the reference-compatibility
flags, and voxelised ellipsoid bodies (anisotropic inertia, uneven particle counts per body)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import Engine, capi, scenes
from gpu_util import Dev, sync, zeros
from test_gpu_engine import assert_state_equal

pytestmark = pytest.mark.gpu
L = capi.lib


def test_aabb_exact():
    print("stage: aabb", flush=True)
    rng = np.random.default_rng(3)
    for n in (1, 1000, 777777):
        pos = rng.normal(scale=50.0, size=(n, 4)).astype(np.float32)
        d = Dev(pos)
        mn, mx = zeros(4, np.float32), zeros(4, np.float32)
        wsb = L.arsim_aabb_workspace_bytes(n)
        ws = Dev(shape=(wsb,), dtype=np.uint8)
        capi.check(L.arsim_aabb(d.p, n, mn.p, mx.p, ws.p, wsb, None))
        sync()
        assert np.array_equal(mn.get(), pos.min(0)) and np.array_equal(mx.get(), pos.max(0))


def test_dynamic_morton_box_step():
    print("engine: dynamic AABB", flush=True)
    sc = scenes.mix_scene(6000, 2, capi.METHOD_LBVH30)
    sc.cfg.flags = capi.DYNAMIC_AABB
    with Engine(sc, use_graph=False) as e:
        e.step(3)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(3)
    assert_state_equal(g, st, "dynamic aabb")


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
def test_reference_compat_flags_step(method):
    print(f"engine: compat flags method={method}", flush=True)
    sc = scenes.mix_scene(8000, 3, method)
    sc.cfg.flags = capi.COMPAT_REF_UNITS | capi.COMPAT_DROP_FIRST | capi.COMPAT_GRAIN_CLAMP
    # forces act as F / m per step in the reference units: a soft spring keeps 4 steps finite
    sc.params.spring, sc.params.damping, sc.params.shear = 20.0, 0.2, 0.05
    # grains outside the walls, so that the clamp acts on the first step
    g_idx = np.where(sc.rb < 0)[0]
    sc.pos[g_idx[:50], 0] = sc.cfg.boxMax[0] + 0.3
    sc.pos[g_idx[50:100], 2] = sc.cfg.boxMin[2] - 0.3
    sc.rb_ang[:, :3] = np.float32(2.0)
    with Engine(sc, use_graph=False) as e:
        e.step(4)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(4)
    assert np.isfinite(st.pos).all() and np.isfinite(st.vel).all()
    assert st.pos[g_idx[:50], 0].max() < sc.cfg.boxMax[0]
    assert_state_equal(g, st, "compat")


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH63])
def test_ellipsoid_pile_steps(method):
    print(f"engine: ellipsoid pile method={method}", flush=True)
    sc = scenes.config2(method, seed=1, scale=0.01)
    assert sc.b > 100 and sc.ppo.min() >= 30 and sc.ppo.max() > 8 * sc.ppo.min()
    I = sc.rb_inertia.reshape(-1, 3, 3)
    assert np.any(np.abs(I[:, 0, 0] - I[:, 1, 1]) > 1e-6)       # anisotropic inverse inertia
    sc.rb_ang[:, :3] = np.random.default_rng(0).normal(scale=2.0, size=(sc.b, 3)).astype(np.float32)
    with Engine(sc, use_graph=True) as e:
        e.step(10)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(10)
    assert_state_equal(g, st, "ellipsoids")
    s64 = oracle.State(sc, np.float64)
    s64.step(10)
    assert np.max(np.abs(g["rb_ang"][:, :3] - s64.rb_ang[:, :3])) < 1e-4 * max(1.0, np.abs(s64.rb_ang).max())


def test_engine_cell_index_is_cell_start():
    """The engine's cell table after a step: scenes' cell ranges packed back to back, and
    table[c] = first sorted index with key >= c for every cell c and c = numCells (exact CSR)."""
    print("engine: cell index", flush=True)
    mix = [scenes.mix_scene(n, 70 + n, capi.METHOD_GRID) for n in (1500, 9000, 3000, 40000)]
    cells = [s.params.gridSize.x * s.params.gridSize.y * s.params.gridSize.z for s in mix]
    assert len(set(cells)) > 1
    total = int(sum(cells))
    n = sum(s.n for s in mix)
    with Engine(mix, use_graph=False) as e:
        e.step(1)
        e.sync()
        kp, tp = C.c_void_p(), C.c_void_p()
        nul = None
        capi.check(L.arsim_engine_debug_buffers(e.h, C.byref(kp), nul, C.byref(tp), nul, nul, nul, nul, nul, nul, nul))
        keys = np.zeros(n, np.uint32)
        table = np.zeros(total + 1, np.uint32)
        capi.check(L.arsim_memcpy_dtoh(capi.ptr(keys), kp, keys.nbytes, None))
        capi.check(L.arsim_memcpy_dtoh(capi.ptr(table), tp, table.nbytes, None))
    assert np.all(keys[1:] >= keys[:-1]) and int(keys[-1]) < total
    base = np.concatenate([[0], np.cumsum(cells)])
    counts = np.histogram(keys, bins=base)[0]
    assert counts.tolist() == [s.n for s in mix]
    expect = np.searchsorted(keys, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32)
    assert np.array_equal(table, expect)


def test_engine_cell_table_with_long_empty_runs():
    """A scene in one corner of a 256^3 grid: the empty runs between its z slabs (~55,000 cells) and the
    table's tail (~16 M cells) are filled by the long-run pass, over graph replays too."""
    print("engine: cell table, long empty runs", flush=True)
    sc = scenes.mix_scene(6000, 91, capi.METHOD_GRID)
    p = sc.params
    p.gridSize.x = p.gridSize.y = p.gridSize.z = 256
    p.numCells = 256 ** 3
    total = 256 ** 3
    with Engine(sc, use_graph=True) as e:
        for steps in (1, 3):
            e.step(steps)
            e.sync()
            kp, tp = C.c_void_p(), C.c_void_p()
            nul = None
            capi.check(L.arsim_engine_debug_buffers(e.h, C.byref(kp), nul, C.byref(tp), nul, nul, nul, nul, nul, nul,
                                                    nul))
            keys = np.zeros(sc.n, np.uint32)
            table = np.zeros(total + 1, np.uint32)
            capi.check(L.arsim_memcpy_dtoh(capi.ptr(keys), kp, keys.nbytes, None))
            capi.check(L.arsim_memcpy_dtoh(capi.ptr(table), tp, table.nbytes, None))
            gaps = np.diff(np.unique(keys).astype(np.int64))
            assert gaps.max() > 4096 and total - int(keys[-1]) > 4096
            expect = np.searchsorted(keys, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32)
            assert np.array_equal(table, expect)
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(4)
    assert_state_equal(g, st, "sparse grid")


def overflowing_scene():
    """A scene of 64^3 cells with 9,000 grains in one slab (z in [10, 11), y in [1, 15)): even the
    narrowest buckets (256 of 1,024 cells) leave one bucket with more keys than LDS holds (7,776)."""
    sc = scenes.mix_scene(8000, 95, capi.METHOD_GRID)
    rng = np.random.default_rng(95)
    g = np.where(sc.rb < 0)[0]
    assert len(g) >= 1900
    sc2 = sc.copy()
    # put every grain, and more grains up to 9,000 in all, into z in [10, 11), y in [1, 15)
    k = len(g)
    sc2.pos[g, 0] = rng.uniform(1.0, 39.0, k).astype(np.float32)
    sc2.pos[g, 1] = rng.uniform(1.0, 15.0, k).astype(np.float32)
    sc2.pos[g, 2] = rng.uniform(10.05, 10.95, k).astype(np.float32)
    extra = 9000 - k
    add_pos = np.zeros((extra, 4), np.float32)
    add_pos[:, 0] = rng.uniform(1.0, 39.0, extra)
    add_pos[:, 1] = rng.uniform(1.0, 15.0, extra)
    add_pos[:, 2] = rng.uniform(10.05, 10.95, extra)
    sc2.pos = np.concatenate([sc2.pos, add_pos])
    sc2.vel = np.concatenate([sc2.vel, np.zeros_like(add_pos)])
    sc2.rb = np.concatenate([sc2.rb, np.full(extra, -1, np.int32)])
    sc2.body_frame = np.concatenate([sc2.body_frame, np.zeros_like(add_pos)])
    keys, _ = oracle.calc_hash(sc2.pos, sc2.params)
    assert np.bincount(keys >> 10).max() > 7776 and sc2.n / 256 < 0.75 * 7776
    return sc2


def read_keys_and_table(e, n, total):
    kp, tp = C.c_void_p(), C.c_void_p()
    nul = None
    capi.check(L.arsim_engine_debug_buffers(e.h, C.byref(kp), nul, C.byref(tp), nul, nul, nul, nul, nul, nul, nul))
    keys = np.zeros(n, np.uint32)
    table = np.zeros(total + 1, np.uint32)
    capi.check(L.arsim_memcpy_dtoh(capi.ptr(keys), kp, keys.nbytes, None))
    capi.check(L.arsim_memcpy_dtoh(capi.ptr(table), tp, table.nbytes, None))
    return keys, table


def test_bucket_sort_overflowing_bucket():
    """The overflowing bucket is sorted in global memory; keys, cell table and state stay exact."""
    print("engine: bucket sort, one bucket past the LDS capacity", flush=True)
    sc2 = overflowing_scene()
    total = int(sc2.params.numCells)
    with Engine(sc2, use_graph=False) as e:
        e.step(2)
        e.sync()
        kk, table = read_keys_and_table(e, sc2.n, total)
        g2 = e.download()
    assert np.all(kk[1:] >= kk[:-1])
    assert np.array_equal(table, np.searchsorted(kk, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32))
    st = oracle.State(sc2, np.float32)
    st.step(2)
    assert_state_equal(g2, st, "overflowing bucket")


def test_bucket_sort_ten_bit_tile_rank():
    """A scene of 200,000 particles on a 128^3 grid with its 50,000 grains packed into one z plane: at
    512 buckets (2^12 cells each) one bucket would hold 12,500 keys, so the plan takes 1,024 buckets of
    2^11 cells and the pose kernel ranks its tiles by a 10-bit bucket index.  Keys, cell table and state
    stay exact after graph-replayed steps."""
    print("engine: bucket sort, 1,024 buckets in one scene", flush=True)
    sc = scenes.mix_scene(200000, 64, capi.METHOD_GRID)
    p = sc.params
    p.gridSize.x = p.gridSize.y = p.gridSize.z = 128
    p.numCells = 128 ** 3
    rng = np.random.default_rng(64)
    g = np.where(sc.rb < 0)[0]
    assert len(g) >= 45000
    sc.pos[g, 0] = rng.uniform(1.0, 127.0, len(g)).astype(np.float32)
    sc.pos[g, 1] = rng.uniform(1.0, 127.0, len(g)).astype(np.float32)
    sc.pos[g, 2] = rng.uniform(10.05, 10.95, len(g)).astype(np.float32)
    keys, _ = oracle.calc_hash(sc.pos, p)
    assert np.bincount(keys >> 12).max() > 7776 and np.bincount(keys >> 11).max() < 7776
    total = 128 ** 3
    with Engine(sc, use_graph=True) as e:
        e.step(2)
        e.sync()
        kk, table = read_keys_and_table(e, sc.n, total)
        got = e.download()
        stages = [name for name, _ in e.profile()]
    assert "bucket_sort_reorder" in stages, stages
    assert np.all(kk[1:] >= kk[:-1])
    assert np.array_equal(table, np.searchsorted(kk, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32))
    st = oracle.State(sc, np.float32)
    st.step(2)
    assert_state_equal(got, st, "1,024 buckets")


def test_bucket_sort_scene_buckets_in_a_batch():
    """Buckets per scene in a batch: a scene of 8^3 cells first, so the later scenes' cell ranges (and
    their buckets) start at offsets that are not multiples of the bucket width; then the overflowing
    scene (its full bucket sorted in global memory) and a scene of 4,000 particles.  The batch's keys are
    sorted, its cell table is exact CSR over all scenes' cells, and every scene equals the oracle alone
    after graph-replayed steps, for one and two workgroups per bucket."""
    print("engine: bucket sort, per-scene buckets in a batch", flush=True)
    small = scenes.mix_scene(1100, 61, capi.METHOD_GRID)
    small.params.gridSize.x = small.params.gridSize.y = small.params.gridSize.z = 8
    small.params.numCells = 512
    mix = [small, overflowing_scene(), scenes.mix_scene(4000, 62, capi.METHOD_GRID)]
    total = sum(int(s.params.numCells) for s in mix)
    n = sum(s.n for s in mix)
    for split in (1, 2):
        with Engine(mix, use_graph=True, bucket_split=split) as e:
            e.step(3)
            e.sync()
            keys, table = read_keys_and_table(e, n, total)
            got = [e.download(k) for k in range(len(mix))]
            stages = [name for name, _ in e.profile()]
        assert "bucket_sort_reorder" in stages, stages
        assert np.all(keys[1:] >= keys[:-1]) and int(keys[-1]) < total
        assert np.array_equal(table, np.searchsorted(keys, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32))
        for k, sc in enumerate(mix):
            st = oracle.State(sc, np.float32)
            st.step(3)
            assert_state_equal(got[k], st, f"split {split}, scene {k}")


@pytest.mark.parametrize("g", [8, 16, 32, 128])
def test_bucket_sort_key_widths(g):
    """The bucket path at key widths 9, 12, 15 and 21 bits (grids 8^3 ... 128^3): low-key sorts of one
    pass (1 and 4 bits), of a full and a partial pass (7 bits) and of two passes (13 bits).  The cell
    table is exact CSR and the state matches the oracle after graph-replayed steps."""
    print(f"engine: bucket sort, grid {g}^3", flush=True)
    sc = scenes.mix_scene(6000, 100 + g, capi.METHOD_GRID)
    p = sc.params
    p.gridSize.x = p.gridSize.y = p.gridSize.z = g
    p.numCells = g ** 3
    total = g ** 3
    with Engine(sc, use_graph=True) as e:
        e.step(3)
        e.sync()
        kp, tp = C.c_void_p(), C.c_void_p()
        nul = None
        capi.check(L.arsim_engine_debug_buffers(e.h, C.byref(kp), nul, C.byref(tp), nul, nul, nul, nul, nul, nul, nul))
        keys = np.zeros(sc.n, np.uint32)
        table = np.zeros(total + 1, np.uint32)
        capi.check(L.arsim_memcpy_dtoh(capi.ptr(keys), kp, keys.nbytes, None))
        capi.check(L.arsim_memcpy_dtoh(capi.ptr(table), tp, table.nbytes, None))
        got = e.download()
        stages = [name for name, _ in e.profile()]   # one more step, after the state was read
    assert "bucket_sort_reorder" in stages, stages
    assert np.all(keys[1:] >= keys[:-1]) and int(keys[-1]) < total
    assert np.array_equal(table, np.searchsorted(keys, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32))
    st = oracle.State(sc, np.float32)
    st.step(3)
    assert_state_equal(got, st, f"grid {g}")


def test_benchmark_scene_two_steps_bit_exact():
    """The benchmark's own workload: 2^20 particles (8,192 cubes + 524,288 grains) on the 256^3 hashed
    grid, two graph-replayed steps, every state array bit-identical to the oracle, and the cell table
    exact CSR over all 2^24 cells (buckets of about 4096 keys, two rounds per workgroup)."""
    print("engine: 2^20-particle benchmark scene, 2 steps", flush=True)
    sc = scenes.config1(capi.METHOD_GRID)
    assert sc.n == 1 << 20
    total = int(sc.params.numCells)
    with Engine(sc, use_graph=True) as e:
        e.step(2)
        e.sync()
        kp, tp = C.c_void_p(), C.c_void_p()
        nul = None
        capi.check(L.arsim_engine_debug_buffers(e.h, C.byref(kp), nul, C.byref(tp), nul, nul, nul, nul, nul, nul, nul))
        keys = np.zeros(sc.n, np.uint32)
        table = np.zeros(total + 1, np.uint32)
        capi.check(L.arsim_memcpy_dtoh(capi.ptr(keys), kp, keys.nbytes, None))
        capi.check(L.arsim_memcpy_dtoh(capi.ptr(table), tp, table.nbytes, None))
        got = e.download()
        stages = [name for name, _ in e.profile()]
    assert "bucket_sort_reorder" in stages, stages
    assert np.all(keys[1:] >= keys[:-1])
    assert np.array_equal(table, np.searchsorted(keys, np.arange(total + 1, dtype=np.int64), "left").astype(np.uint32))
    st = oracle.State(sc, np.float32)
    st.step(2)
    assert_state_equal(got, st, "benchmark scene")


def test_benchmark_batch_one_step_bit_exact():
    """The benchmark's batch: 512 scenes log-uniform in [1024, 65536] particles (8.85 M) stepped as one
    graph-replayed batch.  Every scene's state after a step equals the oracle's step of that scene
    alone (27-bit keys: the bucket sort with buckets per scene, and long empty runs of the cell table)."""
    print("engine: 512-scene benchmark batch, 1 step", flush=True)
    mix = scenes.batch_mix(512, seed=0, method=capi.METHOD_GRID)
    assert sum(s.n for s in mix) > 8_000_000
    with Engine(mix, use_graph=True) as e:
        e.step(1)
        e.sync()
        got = [e.download(k) for k in range(len(mix))]
    for k, sc in enumerate(mix):
        st = oracle.State(sc, np.float32)
        st.step(1)
        assert_state_equal(got[k], st, f"batch scene {k}")


def test_segmented_cell_table_past_the_exact_limit():
    """Two scenes on 1024^3 grids (2^31 cells in all, past the 2^30 cells of an exact table): the table
    holds segments of two cells, the DEM scans the sorted keys of each row, and every scene's state
    after two graph-replayed steps equals the oracle's step of that scene alone."""
    print("engine: segmented cell table, 2 x 1024^3 cells", flush=True)
    mix = [scenes.mix_scene(n, 300 + n, capi.METHOD_GRID) for n in (5000, 7000)]
    for sc in mix:
        p = sc.params
        p.gridSize.x = p.gridSize.y = p.gridSize.z = 1024
        p.numCells = 1024 ** 3
    with Engine(mix, use_graph=True) as e:
        e.step(2)
        e.sync()
        got = [e.download(k) for k in range(len(mix))]
        stages = [name for name, _ in e.profile()]
    assert "dem_grid" in stages and "reduce_apply" in stages, stages   # the unfused kernels
    for k, sc in enumerate(mix):
        st = oracle.State(sc, np.float32)
        st.step(2)
        assert_state_equal(got[k], st, f"segmented table, scene {k}")


@pytest.mark.parametrize("scene_counts", [(30000,), (9000, 20000, 4000)])
def test_engine_lbvh_node_boxes(scene_counts):
    """The engine's step keeps node boxes only inside the traversal records; the debug view restates
    them.  Every leaf's box must be its centre -/+ r and every internal node's box the min / max of its
    children's, over a tree of many refit workgroups (one scene, and a forest of three)."""
    print(f"engine: LBVH node boxes, scenes {scene_counts}", flush=True)
    mix = [scenes.mix_scene(n, 500 + n, capi.METHOD_LBVH30) for n in scene_counts]
    n = sum(s.n for s in mix)
    r = np.float32(mix[0].params.particleRadius)
    with Engine(mix if len(mix) > 1 else mix[0], use_graph=True) as e:
        e.step(2)
        e.sync()
        lp, rp, bp, cp = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        nul = None
        capi.check(L.arsim_engine_debug_buffers(e.h, nul, nul, nul, nul, nul, nul, C.byref(lp), C.byref(rp),
                                                C.byref(bp), C.byref(cp)))
        left = np.zeros(2 * n, np.int32)
        right = np.zeros(2 * n, np.int32)
        bounds = np.zeros((2 * n, 8), np.float32)
        cms = np.zeros((n, 4), np.float32)
        for host, dev in ((left, lp), (right, rp), (bounds, bp), (cms, cp)):
            capi.check(L.arsim_memcpy_dtoh(capi.ptr(host), dev, host.nbytes, None))
    m = 2 * n - 1
    lo = np.full((m, 3), np.inf, np.float32)
    hi = np.full((m, 3), -np.inf, np.float32)
    lo[:n] = cms[:, :3] - r
    hi[:n] = cms[:, :3] + r
    L_, R_ = left[n:m], right[n:m]
    assert L_.min() >= 0 and L_.max() < m and R_.min() >= 0 and R_.max() < m
    for _ in range(200):   # bottom-up to a fixed point (tree depth passes)
        nlo = np.minimum(lo[L_], lo[R_])
        nhi = np.maximum(hi[L_], hi[R_])
        if np.array_equal(nlo, lo[n:]) and np.array_equal(nhi, hi[n:]):
            break
        lo[n:], hi[n:] = nlo, nhi
    else:
        raise AssertionError("node boxes did not settle")
    assert np.isfinite(lo).all() and np.isfinite(hi).all()
    assert np.array_equal(bounds[:m, :3], lo) and np.array_equal(bounds[:m, 4:7], hi)
