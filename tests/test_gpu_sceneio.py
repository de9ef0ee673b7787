"""Engine parity on scenes built from the reference's particle templates, and snapshot / restore:
a run interrupted by a snapshot and continued in a new engine gives the same bits as the run that
was never interrupted (and as the oracle)."""
import os

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import Engine, capi, scenes
from test_gpu_engine import assert_state_equal

pytestmark = pytest.mark.gpu
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "OBJparticles")


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
def test_template_pile_vs_oracle(method):
    print(f"templates: pile method={method}", flush=True)
    sc = scenes.template_pile(DATA, grains=300, method=method)
    with Engine(sc) as e:
        e.step(40)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(40)
    assert_state_equal(g, st, "template pile")


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
def test_snapshot_restore_continues_bit_exact(method):
    print(f"snapshot: method={method}", flush=True)
    sc = scenes.mix_scene(6000, 5, method)
    c = sc.rb_pos[:, :3].mean(0)
    sc.rb_pos[:, :3] = c + (sc.rb_pos[:, :3] - c) * np.float32(0.65)
    gi = sc.rb < 0
    sc.pos[gi, :3] = c + (sc.pos[gi, :3] - c) * np.float32(0.65)
    with Engine(sc) as a:
        a.step(15)
        snap = a.snapshot()
        a.step(15)
        a.sync()
        full = a.download()
    with Engine(snap) as b:
        b.step(15)
        b.sync()
        resumed = b.download()
    for k in ("pos", "vel", "rel", "rb_pos", "rb_vel", "rb_ang", "rb_quat", "rb_iw"):
        assert np.array_equal(full[k], resumed[k]), k


def test_snapshot_of_batched_scene_continues_alone():
    print("snapshot: batched scene continued alone", flush=True)
    mix = [scenes.mix_scene(n, 40 + n, capi.METHOD_GRID) for n in (2000, 5000, 3000)]
    with Engine(mix) as a:
        a.step(10)
        snap = a.snapshot(1)
        a.step(10)
        a.sync()
        full = a.download(1)
    with Engine(snap) as b:
        b.step(10)
        b.sync()
        resumed = b.download()
    for k in ("pos", "vel", "rb_pos", "rb_vel", "rb_ang", "rb_quat"):
        assert np.array_equal(full[k], resumed[k]), k


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
def test_render_state_matches_oracle(method):
    """Model matrices of every body after a run equal the oracle's bit for bit (rotation of the
    cumulative quaternion, position in the last column), and the rotation is orthonormal."""
    print(f"render state: method={method}", flush=True)
    sc = scenes.template_pile(DATA, grains=100, method=method)
    with Engine(sc) as e:
        e.step(25)
        m, q = e.render_state()
    st = oracle.State(sc, np.float32)
    st.step(25)
    assert np.array_equal(m, st.model) and np.array_equal(q, st.qcum)
    M = m.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)   # row-major view of the mat4
    R = M[:, :3, :3]
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
    assert np.array_equal(m.reshape(-1, 4, 4)[:, 3, :3], st.rb_pos[:, :3])
