"""The benchmark configurations themselves, stepped on the GPU and compared bit for bit with the
oracle: the 2^20-particle scene with the 30- and 63-bit LBVH, the 512-scene batch as one LBVH forest,
and the full 4,194,304-particle dense pile.  The engine's synchronize reports an LBVH traversal-stack
overflow as an error, so a passing LBVH test also shows that no traversal overflowed."""
import ctypes as C

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import Engine, capi, scenes
from test_gpu_engine import assert_state_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [capi.METHOD_LBVH30, capi.METHOD_LBVH63])
def test_benchmark_scene_lbvh_two_steps_bit_exact(method):
    """8,192 cubes + 524,288 grains in [0,512]^3: 2^20 leaves whose 30-bit codes hold many duplicates
    (the index tie-break decides those subtrees), two graph-replayed steps."""
    print(f"engine: 2^20-particle benchmark scene, LBVH method={method}, 2 steps", flush=True)
    sc = scenes.config1(method)
    assert sc.n == 1 << 20
    with Engine(sc, use_graph=True) as e:
        e.step(2)
        e.sync()                      # raises on a traversal-stack overflow
        got = e.download()
    st = oracle.State(sc, np.float32)
    st.step(2)
    assert_state_equal(got, st, f"benchmark scene, method {method}")


@pytest.mark.parametrize("method", [capi.METHOD_LBVH30, capi.METHOD_LBVH63])
def test_lbvh_bucket_past_the_lds_capacity(method):
    """100,000 particles with 25,000 grains packed into one corner: their codes share the top 10 bits,
    so one bucket of the per-bucket Morton sort holds far more keys than LDS does (4,096) and is sorted in
    global memory; with 63-bit codes, grains sharing the top 30 bits are then ranked by the full code.
    Two graph-replayed steps equal the oracle's."""
    print(f"engine: Morton bucket past the LDS capacity, method {method}, 2 steps", flush=True)
    sc = scenes.mix_scene(100000, 33, method)
    g = np.where(sc.rb < 0)[0]
    assert len(g) >= 20000
    rng = np.random.default_rng(33)
    lo = sc.pos[:, :3].min(0)
    sc.pos[g, :3] = (lo + 1.0 + rng.uniform(0.0, 6.0, (len(g), 3))).astype(np.float32)
    with Engine(sc, use_graph=True) as e:
        e.step(2)
        e.sync()
        got = e.download()
        stages = [name for name, _ in e.profile()]
    assert "bucket_sort" in stages, stages
    st = oracle.State(sc, np.float32)
    st.step(2)
    assert_state_equal(got, st, f"Morton bucket past the LDS capacity, method {method}")


@pytest.mark.parametrize("method", [capi.METHOD_LBVH30, capi.METHOD_LBVH63])
def test_lbvh_bucket_long_runs_of_equal_top_bits(method):
    """3,000 grains inside a cube a third of a particle radius wide: they share a few 30-bit codes, so the
    63-bit path ranks runs of hundreds of equal top bits by their full codes.  A step equals the oracle's."""
    print(f"engine: Morton bucket, long runs of equal 30-bit codes, method {method}", flush=True)
    sc = scenes.mix_scene(100000, 34, method)
    g = np.where(sc.rb < 0)[0][:3000]
    rng = np.random.default_rng(34)
    lo = sc.pos[:, :3].min(0)
    sc.pos[g, :3] = (lo + 1.0 + rng.uniform(0.0, 0.15, (len(g), 3))).astype(np.float32)
    with Engine(sc, use_graph=True) as e:
        e.step(1)
        e.sync()
        got = e.download()
        stages = [name for name, _ in e.profile()]
    assert "bucket_sort" in stages, stages
    st = oracle.State(sc, np.float32)
    st.step(1)
    assert_state_equal(got, st, f"long runs of equal 30-bit codes, method {method}")


def test_benchmark_batch_lbvh_forest_one_step():
    """The 512-scene batch (8.85 M particles) as one LBVH forest: 64-bit keys with the scene id above
    bit 30, one tree per scene.  Every scene's state after a step equals the oracle's step of that
    scene alone."""
    print("engine: 512-scene benchmark batch, LBVH30 forest, 1 step", flush=True)
    mix = scenes.batch_mix(512, seed=0, method=capi.METHOD_LBVH30)
    assert sum(s.n for s in mix) > 8_000_000
    with Engine(mix, use_graph=True) as e:
        e.step(1)
        e.sync()
        got = [e.download(k) for k in range(len(mix))]
    for k, sc in enumerate(mix):
        st = oracle.State(sc, np.float32)
        st.step(1)
        assert_state_equal(got[k], st, f"forest scene {k}")


def dense_pile_stats(sc):
    """Body sizes and contacts per particle of a pile (contact = the DEM predicate: closer than 2r and
    not two particles of one body)."""
    from scipy.spatial import cKDTree
    pos = sc.pos[:, :3].astype(np.float64)
    pairs = cKDTree(pos).query_pairs(2.0 * sc.params.particleRadius, output_type="ndarray")
    a, b = sc.rb[pairs[:, 0]], sc.rb[pairs[:, 1]]
    dem = pairs[(a != b) | (a < 0)]
    per = np.bincount(dem.ravel(), minlength=sc.n)
    sizes = np.bincount(sc.rb[sc.rb >= 0])
    return {"sizes": sizes, "contacts": per.mean(), "grain_contacts": per[sc.rb < 0].mean(),
            "pairs_within_2r": 2.0 * len(pairs) / sc.n}


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
def test_dense_pile_full_size_one_step(method):
    """16,384 ellipsoids voxelised at spacing r (33 to ~2,000 particles each, heavy-tailed) in a
    packed bed of grains, 4,194,304 particles: one step equals the oracle's."""
    print(f"engine: dense pile, 4,194,304 particles, method={method}, 1 step", flush=True)
    sc = scenes.config2(method)
    assert sc.n == 4194304 and sc.b == 16384
    st = dense_pile_stats(sc)
    print(f"  body particles {st['sizes'].min()}..{st['sizes'].max()} (median {np.median(st['sizes']):.0f}), "
          f"contacts/particle {st['contacts']:.2f} (grains {st['grain_contacts']:.2f}), "
          f"pairs within 2r/particle {st['pairs_within_2r']:.1f}", flush=True)
    assert st["sizes"].min() >= 30 and st["sizes"].max() > 1000 and st["grain_contacts"] > 6.0
    with Engine(sc, use_graph=True) as e:
        e.step(1)
        e.sync()
        got = e.download()
    ref = oracle.State(sc, np.float32)
    ref.step(1)
    assert_state_equal(got, ref, f"dense pile, method {method}")


@pytest.mark.parametrize("method", [capi.METHOD_LBVH30, capi.METHOD_LBVH63])
def test_lbvh_bucket_sort_stable_on_equal_codes(method):
    """Grains stacked five deep on identical positions give equal codes, 30- and 63-bit alike: after a
    step the engine's sorted codes are non-decreasing, every particle appears once, and equal codes
    keep particle order (the stable sort).  The state itself is not compared: coincident grains have
    no contact normal."""
    print(f"engine: Morton bucket sort stability on equal codes, method {method}", flush=True)
    sc = scenes.mix_scene(100000, 35, method)
    g = np.where(sc.rb < 0)[0]
    groups = g[: (len(g) // 5) * 5].reshape(-1, 5)
    sc.pos[groups[:, 1:], :3] = sc.pos[groups[:, :1], :3]
    sc.vel[groups[:, 1:], :3] = sc.vel[groups[:, :1], :3]   # still together after the step's integration
    n = sc.n
    with Engine(sc, use_graph=False) as e:
        e.step(1)
        e.sync()
        kp, ip = C.c_void_p(), C.c_void_p()
        nul = None
        capi.check(capi.lib.arsim_engine_debug_buffers(e.h, C.byref(kp), C.byref(ip), nul, nul, nul, nul, nul, nul,
                                                       nul, nul))
        keys = np.zeros(n, np.uint64 if method == capi.METHOD_LBVH63 else np.uint32)
        idx = np.zeros(n, np.uint32)
        capi.check(capi.lib.arsim_memcpy_dtoh(capi.ptr(keys), kp, keys.nbytes, None))
        capi.check(capi.lib.arsim_memcpy_dtoh(capi.ptr(idx), ip, idx.nbytes, None))
        stages = [name for name, _ in e.profile()]
    assert "bucket_sort" in stages, stages
    k = keys.astype(np.uint64)
    assert np.all(k[1:] >= k[:-1])
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))
    eq = k[1:] == k[:-1]
    assert eq.sum() >= len(groups) * 4   # the stacked grains share their codes
    assert np.all(idx[1:][eq] > idx[:-1][eq])
