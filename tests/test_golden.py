"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): seeded squeezed
scenes with the oracle's FP32 and FP64 state after a few steps.  The CPU test pins the oracle
against its own stored outputs (any change to its arithmetic shows up here); the GPU test steps the
engine on the same inputs and requires the FP32 state bit for bit, and the FP64 state to 1e-4."""
import glob
import os

import numpy as np
import pytest

import oracle
from diploma_thesis_amd.scenes import HostScene

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))
KEYS = [("pos", "pos"), ("vel", "vel"), ("rel", "rel"), ("rb_pos", "rb_pos"), ("rb_vel", "rb_vel"),
        ("rb_ang", "rb_ang"), ("rb_quat", "qcum"), ("rb_iw", "Iw")]


def load(path):
    d = np.load(path, allow_pickle=False)
    sc = HostScene.from_arrays({k[3:]: d[k] for k in d.files if k.startswith("in_")})
    return sc, d


def test_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_golden(path):
    sc, d = load(path)
    steps = int(d["steps"])
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        st = oracle.State(sc, dt)
        st.step(steps)
        for _, ok in KEYS:
            assert np.array_equal(getattr(st, ok), d[f"{tag}_{ok}"]), f"{tag} {ok}"
    # the fixture is a contact case: bodies moved off their free-flight path
    assert np.max(np.abs(d["f32_rb_vel"] - d["in_rb_vel"])) > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_engine_matches_golden(path):
    from diploma_thesis_amd import Engine
    print(f"golden: {os.path.basename(path)}", flush=True)
    sc, d = load(path)
    with Engine(sc, use_graph=True) as e:
        e.step(int(d["steps"]))
        e.sync()
        g = e.download()
    for gk, ok in KEYS:
        assert np.array_equal(g[gk], d[f"f32_{ok}"]), gk
        ref = d[f"f64_{ok}"]
        scale = max(1.0, float(np.max(np.abs(ref))))
        assert float(np.max(np.abs(g[gk] - ref))) <= 1e-4 * scale, gk
