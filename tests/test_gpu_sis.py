"""Sequential-impulse contact model on the GPU against the oracle, bit for bit: the deepest-contact
finders and the gather / level-scheduled sweeps through the C ABI, and the engine's SIS step (one
scene with a static floor, and a batch of scenes against each scene alone)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import Engine, capi, scenes
from gpu_util import Dev, sync, zeros
from test_gpu_engine import assert_state_equal
from test_gpu_static import floor_points

pytestmark = pytest.mark.gpu
L = capi.lib


def cube_offsets(k=4):
    g = np.arange(k, dtype=np.float32) - (k - 1) / 2
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


def quat_rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def dense_bodies(count, lo, span, seed, k=4, vel_sigma=1.0):
    """`count` cubes of k^3 particles, centres uniform in [lo, lo + span]^3 (they overlap), random
    orientations and velocities; bodies only, sequential-impulse flags set."""
    rng = np.random.default_rng(seed)
    base = scenes.mix_scene(4096, seed)
    offs = cube_offsets(k)
    m = offs.shape[0]
    inertia = (offs ** 2).sum() * np.eye(3) - offs.T.astype(np.float64) @ offs
    inv = np.linalg.inv(inertia).astype(np.float32)
    sc = base.copy()
    cms = rng.uniform(lo, lo + span, size=(count, 3)).astype(np.float32)
    q = random_quats(rng, count)
    sc.rb = np.repeat(np.arange(count, dtype=np.int32), m)
    sc.body_frame = np.zeros((count * m, 4), np.float32)
    sc.body_frame[:, :3] = np.tile(offs, (count, 1))
    sc.pos = np.zeros((count * m, 4), np.float32)
    for b in range(count):
        sc.pos[b * m:(b + 1) * m, :3] = cms[b] + offs @ quat_rot(q[b]).T
    sc.vel = np.zeros_like(sc.pos)
    sc.rb_pos = np.zeros((count, 4), np.float32)
    sc.rb_pos[:, :3] = cms
    sc.rb_vel = np.zeros((count, 4), np.float32)
    sc.rb_vel[:, :3] = rng.normal(scale=vel_sigma, size=(count, 3))
    sc.rb_ang = np.zeros((count, 4), np.float32)
    sc.rb_ang[:, :3] = rng.normal(scale=0.5, size=(count, 3))
    sc.rb_quat = q
    sc.rb_inertia = np.tile(inv.T.reshape(1, 9), (count, 1))    # column-major
    sc.rb_mass = np.full(count, float(m), np.float32)
    sc.ppo = np.full(count, m, np.int32)
    sc.params.numBodies = count
    sc.params.ARrestitution, sc.params.RBrestitution = 0.5, 0.8
    sc.params.ARfriction, sc.params.RBfriction = 0.3, 0.1
    sc.cfg.flags = capi.CONTACT_SIS
    sc.cfg.dt = 2e-3
    return sc


def grid_arrays(pos, p):
    h, i = oracle.calc_hash(pos, p)
    sh, si = oracle.sort_pairs(h, i)
    cs, ce = oracle.cell_bounds(sh, p.numCells)
    return si, cs, ce, np.ascontiguousarray(pos[si])


def test_find_deepest_reference_layout_vs_oracle():
    print("sis: deepest contacts, reference layout", flush=True)
    sc = dense_bodies(60, 8.0, 10.0, seed=1)
    grains = np.random.default_rng(2).uniform(8, 18, size=(300, 4)).astype(np.float32)
    grains[:, 3] = 0
    pos = np.concatenate([sc.pos, grains])
    rb = np.concatenate([sc.rb, np.full(len(grains), -1, np.int32)])
    n = len(rb)
    si, cs, ce, spos = grid_arrays(pos, sc.params)
    want = oracle.sis_find_deepest(si, cs, ce, spos, rb, sc.params)
    dRb, dSi, dCs, dCe, dSp = Dev(rb), Dev(si), Dev(cs), Dev(ce), Dev(spos)
    dDist, dBody, dPart = zeros(n, np.float32), Dev(np.full(n, -7, np.int32)), Dev(np.full(n, -7, np.int32))
    dCol = zeros((n, 4), np.float32)
    capi.check(L.arsim_find_deepest_contacts_grid(dRb.p, dBody.p, dPart.p, dDist.p, dCol.p, dSp.p, dSi.p, dCs.p,
                                                  dCe.p, n, C.byref(sc.params), None), "find_deepest")
    sync()
    dist, body, part, col = dDist.get(), dBody.get(), dPart.get(), dCol.get()
    assert np.array_equal(dist, want[0])
    hit = want[0] > 0
    assert hit.sum() > 500
    assert np.array_equal(body[hit], want[1][hit]) and np.array_equal(part[hit], want[2][hit])
    assert (body[~hit] == -7).all() and (part[~hit] == -7).all()          # written only on a contact
    assert np.array_equal(col[:, 0] == 1.0, hit)


def test_find_static_reference_layout_vs_oracle():
    print("sis: static contacts, reference layout", flush=True)
    rng = np.random.default_rng(3)
    sc = dense_bodies(20, 6.0, 12.0, seed=3)
    P, N = floor_points(4.0, 20.0, 4.0, 20.0, 9.0, 0.45, rng)
    n = sc.n
    si, cs, ce, spos = grid_arrays(sc.pos, sc.params)
    ssi, scs, sce, sspos = grid_arrays(P, sc.params)
    want = oracle.sis_find_static(si, spos, ssi, scs, sce, sspos, sc.params)
    dDist, dPart = zeros(n, np.float32), Dev(np.full(n, -7, np.int32))
    bufs = [Dev(spos), Dev(sspos), Dev(si), Dev(ssi), Dev(scs), Dev(sce)]
    capi.check(L.arsim_find_static_contacts_grid(dPart.p, dDist.p, None, *[b.p for b in bufs], n, C.byref(sc.params),
                                                 None), "find_static")
    sync()
    hit = want[0] > 0
    assert hit.sum() > 100
    assert np.array_equal(dDist.get(), want[0]) and np.array_equal(dPart.get()[hit], want[1][hit])


def test_gather_and_sweeps_vs_oracle():
    print("sis: gather + level-scheduled sweeps", flush=True)
    rng = np.random.default_rng(5)
    sc = dense_bodies(150, 8.0, 9.0, seed=5)
    P, N = floor_points(6.0, 20.0, 6.0, 20.0, 10.0, 0.45, rng)
    st = oracle.State(sc, np.float32)
    st.integrate_bodies()
    st.pose_grains()
    p = sc.params
    si, cs, ce, spos = grid_arrays(st.pos, p)
    ssi, scs, sce, sspos = grid_arrays(P, p)
    dA, pA = oracle.sis_find_static(si, spos, ssi, scs, sce, sspos, p)
    dB, bB, pB = oracle.sis_find_deepest(si, cs, ce, spos, st.rb, p)
    want = st.sis_gather(dA, pA, N, dB, bB, pB)
    cnt = len(want["bias"])
    assert cnt > 3000 and (want["bodies"][:, 1] < 0).sum() > 100
    n, nb = st.N, st.B
    cap = 2 * n
    ws = int(L.arsim_sis_workspace_bytes(cap, n))
    dWs = zeros(ws, np.uint8)
    body = [Dev(st.rb_pos), Dev(st.rb_vel), Dev(st.rb_ang), Dev(st.rel), Dev(st.rb)]
    fin = [Dev(dA), Dev(pA), Dev(N), Dev(dB), Dev(bB), Dev(pB)]
    out = [zeros((cap, 4), np.float32), zeros((cap, 4), np.float32), zeros((cap, 2), np.int32), zeros(cap, np.float32),
           Dev(np.full((cap, 4), 9.0, np.float32))]
    dCnt = zeros(1, np.int32)
    capi.check(L.arsim_sis_gather(body[0].p, body[1].p, body[2].p, body[3].p, body[4].p, n, *[f.p for f in fin],
                                  *[o.p for o in out], dCnt.p, cap, C.byref(p), dWs.p, ws, None), "sis_gather")
    sync()
    assert int(dCnt.get()[0]) == cnt
    got = [o.get()[:cnt] for o in out]
    assert np.array_equal(got[0], want["normal"]) and np.array_equal(got[1], want["point"])
    assert np.array_equal(got[2], want["bodies"]) and np.array_equal(got[3], want["bias"])
    assert (got[4] == 0).all()
    dIw, dM = Dev(st.Iw), Dev(st.mass)
    capi.check(L.arsim_sis_solve(*[o.p for o in out], dCnt.p, cap, body[0].p, body[1].p, body[2].p, dIw.p, dM.p, nb,
                                 18, C.byref(p), dWs.p, ws, None), "sis_solve")
    sync()
    st.sis_solve(want, 18)
    assert np.array_equal(body[1].get(), st.rb_vel) and np.array_equal(body[2].get(), st.rb_ang)
    assert np.array_equal(out[4].get()[:cnt], want["impulse"])


def test_engine_sis_with_static_floor_vs_oracle():
    print("sis: engine step with a static floor", flush=True)
    rng = np.random.default_rng(8)
    sc = dense_bodies(120, 8.0, 12.0, seed=8)
    P, N = floor_points(4.0, 24.0, 4.0, 24.0, 8.5, 0.45, rng)
    with Engine(sc) as e:
        e.set_static(P, N)
        e.step(12)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.set_static(P, N)
    st.step(12)
    assert_state_equal(g, st, "sis + static floor")
    free = oracle.State(sc, np.float32)
    free.cfg.flags = 0
    free.step(12)
    assert not np.array_equal(free.rb_vel, st.rb_vel)


def test_engine_sis_batched_matches_each_scene_alone():
    print("sis: batched scenes vs alone", flush=True)
    mix = [dense_bodies(c, 6.0, s, seed=20 + k) for k, (c, s) in enumerate([(40, 6.0), (5, 9.0), (90, 10.0), (1, 4.0)])]
    with Engine(mix) as e:
        e.step(6)
        e.sync()
        many = [e.download(s) for s in range(len(mix))]
    for s, sc in enumerate(mix):
        st = oracle.State(sc, np.float32)
        st.step(6)
        assert_state_equal(many[s], st, f"scene {s} of the batch")


def test_engine_sis_refuses_grains():
    sc = scenes.mix_scene(2048, 1)
    sc.cfg.flags = capi.CONTACT_SIS
    with pytest.raises(capi.ArSimError):
        Engine(sc)


def test_engine_sis_without_contacts_and_single_body():
    print("sis: no contacts, one body", flush=True)
    far = dense_bodies(8, 4.0, 30.0, seed=31, vel_sigma=0.0)
    far.rb_pos[:, :3] = np.float32([[4 + 8 * (b % 4), 4 + 8 * (b // 4), 20.0] for b in range(8)])
    one = dense_bodies(1, 10.0, 1.0, seed=32)
    for sc, what in ((far, "no contacts"), (one, "one body")):
        with Engine(sc) as e:
            e.step(5)
            e.sync()
            g = e.download()
        st = oracle.State(sc, np.float32)
        st.step(5)
        assert_state_equal(g, st, what)


def test_engine_sis_more_bodies_than_lds_holds():
    # 2100 bodies of 8 particles: the sweeps keep velocities in global memory (more than 2048 bodies)
    print("sis: 2100 bodies, velocities in global memory", flush=True)
    sc = dense_bodies(2100, 6.0, 26.0, seed=33, k=2)
    with Engine(sc) as e:
        e.step(3)
        e.sync()
        g = e.download()
    st = oracle.State(sc, np.float32)
    st.step(3)
    assert_state_equal(g, st, "2100 bodies")
