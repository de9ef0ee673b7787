"""Template loader (the reference's OBJparticles format, initObject semantics) and scene builder.

The loader is pinned by an independent numpy restatement run on the same files: the FP32 sequential
centre of mass, the per-model offsets and the FP32 inertia sum must agree bit for bit, the inverse to
2 ulp (the library inverts in double, numpy by LU), and A A^-1 = I; the hard-coded tensors of the
reference are checked as closed forms.  The file suffix follows the reference's comparison chain."""
import os

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import capi, scenes

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "OBJparticles")
R = 1.0 / 64.0
f32 = np.float32


def ulp_dist(a, b):
    """Distance in units in the last place between float32 arrays (+0 and -0 are 0 apart)."""
    def ordered(x):
        i = np.asarray(x, np.float32).view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)
    return np.abs(ordered(a) - ordered(b))


def test_suffix_follows_reference_comparison_chain():
    # abs(scale - s < 0.01): each test is scale < s + 0.01 and the first that holds wins
    cases = {0.3: "1_0", 1.0: "1_0", 1.005: "1_0", 1.2: "1_5", 1.5: "1_5", 1.8: "2_0", 2.0: "2_0",
             2.3: "2_5", 2.5: "2_5", 2.6: "1_5", 4.0: "1_5"}
    for s, want in cases.items():
        assert scenes.template_suffix(s) == want, s


def numpy_template(model, scale):
    path = os.path.join(DATA, model, f"{model}_{scenes.template_suffix(scale)}.txt")
    lines = open(path).read().split("\n")
    n = int(lines[0])
    pts = np.array([[f32(t) for t in lines[1 + i].split()[:3]] for i in range(n)], np.float32)
    c = [f32(0), f32(0), f32(0)]
    for p in pts:
        for k in range(3):
            c[k] = f32(c[k] + p[k])
    c = [f32(v / f32(n)) for v in c]
    off = {"bunny": (-R, R, -R), "banana": (0, 3 * R, 0), "teapot": (0, 2 * R, 0)}.get(model, (0, 0, 0))
    s = [f32(c[k] + f32(off[k])) for k in range(3)]
    rel = (pts - np.array(s, np.float32)).astype(np.float32)
    I = np.eye(3, dtype=np.float32)
    for x, y, z in rel:
        I[0, 0] = f32(I[0, 0] + f32(f32(y * y) + f32(z * z)))
        I[1, 1] = f32(I[1, 1] + f32(f32(x * x) + f32(z * z)))
        I[2, 2] = f32(I[2, 2] + f32(f32(x * x) + f32(y * y)))
        I[0, 1] = I[1, 0] = f32(I[0, 1] - f32(x * y))
        I[0, 2] = I[2, 0] = f32(I[0, 2] - f32(x * z))
        I[1, 2] = I[2, 1] = f32(I[1, 2] - f32(y * z))
    return n, rel, I


@pytest.mark.parametrize("model,scale", [("bunny", 1.0), ("bunny", 2.5), ("teapot", 1.0), ("teapot", 2.5),
                                         ("banana", 2.5), ("cube", 2.5)])
def test_loader_matches_numpy_restatement(model, scale):
    t = scenes.Template(DATA, model, scale, R)
    n, rel, I = numpy_template(model, scale)
    assert t.n == n
    assert np.array_equal(t.rel[:, :3], rel) and not t.rel[:, 3].any()
    inv = t.inv_inertia.reshape(3, 3).T          # column-major -> row-major (symmetric anyway)
    ref = np.linalg.inv(I.astype(np.float64)).astype(np.float32)
    assert ulp_dist(inv, ref).max() <= 2
    assert np.allclose(I.astype(np.float64) @ inv.astype(np.float64), np.eye(3), atol=1e-6)
    assert t.radius == float(rel.max())           # the largest coordinate, as the reference keeps it
    assert t.mass == 1.0
    # centring: the mean of the moved points is minus the per-model offset
    off = {"bunny": (-R, R, -R), "teapot": (0, 2 * R, 0), "banana": (0, 3 * R, 0)}.get(model, (0, 0, 0))
    assert np.allclose(rel.astype(np.float64).mean(0), -np.array(off), atol=1e-6)


@pytest.mark.parametrize("model,scale,diag", [
    ("banana", 2.0, (0.0140314, 0.0218576, 0.0309458)), ("banana", 1.5, (0.00596544, 0.00895867, 0.0125508)),
    ("banana", 1.0, (0.00321249, 0.00464048, 0.00641495)), ("bunny", 1.5, (0.00989096, 0.00979847, 0.0119784)),
    ("teapot", 2.0, (0.0122465, 0.0232885, 0.0208828)), ("teapot", 1.5, (0.0075797, 0.0140341, 0.012614))])
def test_hard_coded_tensors(model, scale, diag):
    t = scenes.Template(DATA, model, scale, R)
    inv = t.inv_inertia.reshape(3, 3)
    want = np.diag([np.float32(1.0 / np.float64(np.float32(d))) for d in diag]).astype(np.float32)
    assert np.array_equal(inv, want)


@pytest.mark.parametrize("scale,edge", [(1.0, 0.1), (1.5, 0.15), (2.0, 0.2)])
def test_cube_inverse_inertia_is_fixed(scale, edge):
    t = scenes.Template(DATA, "cube", scale, R)
    v = np.float32(6.0 / (edge * edge))
    assert np.array_equal(t.inv_inertia.reshape(3, 3), np.diag([v, v, v]).astype(np.float32))


def test_loader_errors():
    with pytest.raises(RuntimeError):
        scenes.Template(DATA, "no_such_model", 1.0, R)


def test_builder_layout_and_oracle_step():
    sc = scenes.template_pile(DATA, grains=100)
    assert sc.b == 16 and sc.n == int(sc.ppo.sum()) + 100
    nb = int(sc.ppo.sum())
    assert np.array_equal(sc.rb[:nb], np.repeat(np.arange(sc.b), sc.ppo)) and np.all(sc.rb[nb:] == -1)
    assert np.array_equal(sc.rb_quat, np.tile(np.float32([0, 0, 0, 1]), (sc.b, 1)))
    # particles sit at body position + template offset
    st = 0
    for b in range(sc.b):
        k = int(sc.ppo[b])
        assert np.array_equal(sc.pos[st:st + k, :3], (sc.rb_pos[b, :3] + sc.body_frame[st:st + k, :3]).astype(np.float32))
        st += k
    s = oracle.State(sc, np.float32)
    s.step(20)
    assert np.isfinite(s.pos).all() and np.isfinite(s.rb_ang).all()
    assert len(oracle.contacts_brute(s.pos, s.rb, sc.params.particleRadius)) > 50
