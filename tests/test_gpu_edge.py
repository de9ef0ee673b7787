"""Edge cases of the engine step against the oracle: one particle, two touching grains, scenes with
only bodies or only grains, and every sort / DEM tuning the autotuner may pick."""
import numpy as np
import pytest

import oracle
from diploma_thesis_amd import Engine, capi, scenes
from test_gpu_engine import assert_state_equal

pytestmark = pytest.mark.gpu
METHODS = [capi.METHOD_GRID, capi.METHOD_LBVH30, capi.METHOD_LBVH63]


def only_grains(sc, count):
    """The scene's first `count` grains alone (no bodies), same parameters and box."""
    g = np.where(sc.rb < 0)[0][:count]
    out = sc.copy()
    out.pos, out.vel = sc.pos[g].copy(), sc.vel[g].copy()
    out.rb = np.full(len(g), -1, np.int32)
    out.body_frame = np.zeros((len(g), 4), np.float32)
    for k in ("rb_pos", "rb_vel", "rb_ang", "rb_quat"):
        setattr(out, k, np.zeros((0, 4), np.float32))
    out.rb_inertia = np.zeros((0, 9), np.float32)
    out.rb_mass = np.zeros(0, np.float32)
    out.ppo = np.zeros(0, np.int32)
    out.params.numBodies = 0
    return out


def only_bodies(sc):
    nb = int(sc.ppo.sum())
    out = sc.copy()
    for k in ("pos", "vel", "rb", "body_frame"):
        setattr(out, k, getattr(sc, k)[:nb].copy())
    return out


def run_both(sc, steps, **kw):
    with Engine(sc, use_graph=False, **kw) as e:
        e.step(steps)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(steps)
    return g, st


@pytest.mark.parametrize("method", METHODS)
def test_single_particle(method):
    print(f"edge: single particle method={method}", flush=True)
    sc = only_grains(scenes.mix_scene(1024, 1, method), 1)
    g, st = run_both(sc, 5)
    assert_state_equal(g, st, "one grain")


@pytest.mark.parametrize("method", METHODS)
def test_two_touching_grains(method):
    print(f"edge: two grains method={method}", flush=True)
    sc = only_grains(scenes.mix_scene(1024, 2, method), 2)
    r = float(sc.params.particleRadius)
    sc.pos[1, :3] = sc.pos[0, :3] + np.float32([1.5 * r, 0.2 * r, 0.0])
    sc.vel[0, :3], sc.vel[1, :3] = (1.0, 0.0, 0.0), (-1.0, 0.0, 0.0)
    g, st = run_both(sc, 3)
    assert_state_equal(g, st, "two grains")
    assert not np.array_equal(st.vel[:, :3], sc.vel[:, :3])   # they did push each other


@pytest.mark.parametrize("method", METHODS)
def test_grains_only_and_bodies_only(method):
    print(f"edge: grains only / bodies only method={method}", flush=True)
    base = scenes.mix_scene(6000, 3, method)
    for sc, what in ((only_grains(base, 1500), "grains only"), (only_bodies(base), "bodies only")):
        g, st = run_both(sc, 4)
        assert_state_equal(g, st, what)


@pytest.mark.parametrize("block,ipt,rbits", [(64, 4, 8), (256, 8, 7), (128, 12, 8), (192, 16, 6), (128, 8, 6)])
@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
def test_every_tuning_gives_the_same_bits(method, block, ipt, rbits):
    print(f"edge: tuning {block}/{ipt}/{rbits} method={method}", flush=True)
    sc = scenes.mix_scene(9000, 4, method)
    g, st = run_both(sc, 2, dem_block=block, sort_ipt=ipt, sort_radix_bits=rbits)
    assert_state_equal(g, st, f"tuning {block}/{ipt}/{rbits}")
