"""The oracle's sequential-impulse model (deepest contact per particle, contact gather, Gauss-Seidel
impulses) pinned on the CPU by brute force, closed forms and conservation laws."""
import types

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import capi, scenes


def params(r=0.5, cells=16, cell=1.0, ar_rest=0.5, rb_rest=1.0, ar_fric=0.3, g=-9.81):
    p = capi.ArSimParams()
    p.particleRadius = r
    p.gridSize.x = p.gridSize.y = p.gridSize.z = cells
    p.numCells = cells ** 3
    p.cellSize.x = p.cellSize.y = p.cellSize.z = cell
    p.gravity.y = g
    p.globalDamping = p.Wdamping = 1.0
    p.ARrestitution, p.RBrestitution, p.ARfriction, p.RBfriction = ar_rest, rb_rest, ar_fric, 0.1
    return p


def cfg(dt=1e-3, flags=capi.CONTACT_SIS):
    c = capi.ArSimStepConfig()
    c.dt = dt
    c.boxMax[:] = [16.0, 16.0, 16.0]
    c.method = capi.METHOD_GRID
    c.flags = flags
    return c


def body_scene(offsets, cms, vels=None, angs=None, inv_inertia=None, masses=None, p=None, c=None):
    """Bodies made of the given body-frame particle offsets (one array per body), identity orientation."""
    b = len(offsets)
    ppo = np.array([len(o) for o in offsets], np.int32)
    frame = np.zeros((int(ppo.sum()), 4), np.float32)
    frame[:, :3] = np.concatenate(offsets)
    rb = np.repeat(np.arange(b, dtype=np.int32), ppo)
    pos = frame.copy()
    pos[:, :3] += np.asarray(cms, np.float32)[rb]
    s = types.SimpleNamespace()
    s.n, s.b = len(rb), b
    s.pos, s.vel, s.body_frame, s.rb = pos, np.zeros_like(pos), frame, rb
    s.rb_pos = np.zeros((b, 4), np.float32)
    s.rb_pos[:, :3] = cms
    s.rb_vel = np.zeros((b, 4), np.float32)
    s.rb_ang = np.zeros((b, 4), np.float32)
    if vels is not None:
        s.rb_vel[:, :3] = vels
    if angs is not None:
        s.rb_ang[:, :3] = angs
    s.rb_quat = np.tile(np.float32([0, 0, 0, 1]), (b, 1))
    inv = np.eye(3, dtype=np.float32).reshape(9) if inv_inertia is None else inv_inertia
    s.rb_inertia = np.tile(np.asarray(inv, np.float32).reshape(1, 9), (b, 1)) if np.ndim(inv) == 1 else \
        np.asarray(inv, np.float32)
    s.rb_mass = np.ones(b, np.float32) if masses is None else np.asarray(masses, np.float32)
    s.ppo = ppo
    s.pstart = (np.cumsum(ppo) - ppo).astype(np.int32)
    s.params = p if p is not None else params()
    s.cfg = c if c is not None else cfg()
    return s


def cube(k=2, spacing=1.0):
    g = (np.arange(k) - (k - 1) / 2) * spacing
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def grid_arrays(pos, p):
    h, i = oracle.calc_hash(pos, p)
    sh, si = oracle.sort_pairs(h, i)
    cs, ce = oracle.cell_bounds(sh, p.numCells)
    return si, cs, ce, np.ascontiguousarray(pos[si])


# ------------------------------------------------------------------ deepest contact: brute force
def test_find_deepest_matches_brute_force():
    rng = np.random.default_rng(3)
    p = params(cells=16)
    offs = [cube(2) for _ in range(12)]
    cms = rng.uniform(3.0, 6.0, size=(12, 3)).astype(np.float32)    # crowded: bodies interpenetrate
    sc = body_scene(offs, cms, p=p)
    # a few grains too: they are partners of bodies but never of each other
    grains = rng.uniform(3.0, 6.0, size=(20, 4)).astype(np.float32)
    grains[:, 3] = 0
    pos = np.concatenate([sc.pos, grains])
    rb = np.concatenate([sc.rb, np.full(20, -1, np.int32)])
    si, cs, ce, spos = grid_arrays(pos, p)
    dist, body, part = oracle.sis_find_deepest(si, cs, ce, spos, rb, p)
    x = pos[:, :3].astype(np.float64)
    d = 2 * 0.5 - np.linalg.norm(x[:, None] - x[None], axis=2)
    d[rb[:, None] == rb[None]] = -np.inf                 # same body, and grain-grain
    best = d.max(1)
    want = np.where(best > 0, best, 0.0)
    assert np.allclose(dist, want, atol=2e-6)
    hit = best > 0
    assert (dist[~hit] == 0).all() and (dist[hit] > 0).all()
    srt = np.sort(d, 1)
    unique = hit & (srt[:, -1] - srt[:, -2] > 1e-5)
    assert unique.sum() > 40
    assert np.array_equal(part[unique], d.argmax(1)[unique])
    assert np.array_equal(body[unique], rb[d.argmax(1)][unique])


def test_find_static_matches_brute_force():
    rng = np.random.default_rng(4)
    p = params(cells=16)
    sc = body_scene([cube(3) for _ in range(4)], rng.uniform(4.0, 9.0, size=(4, 3)).astype(np.float32), p=p)
    stat = rng.uniform(3.0, 10.0, size=(3000, 4)).astype(np.float32)
    stat[:, 3] = 0
    si, cs, ce, spos = grid_arrays(sc.pos, p)
    ssi, scs, sce, sspos = grid_arrays(stat, p)
    dist, part = oracle.sis_find_static(si, spos, ssi, scs, sce, sspos, p)
    x, y = sc.pos[:, :3].astype(np.float64), stat[:, :3].astype(np.float64)
    d = 1.0 - np.linalg.norm(x[:, None] - y[None], axis=2)
    best = d.max(1)
    assert np.allclose(dist, np.where(best > 0, best, 0), atol=2e-6)
    srt = np.sort(d, 1)
    unique = (best > 0) & (srt[:, -1] - srt[:, -2] > 1e-5)
    assert unique.sum() > 20
    assert np.array_equal(part[unique], d.argmax(1)[unique])


# ------------------------------------------------------------------ gather
def test_gather_points_normals_bias_and_order():
    p = params(rb_rest=0.7, ar_rest=0.4)
    offs = [np.float32([[0, 0, 0], [1, 0, 0]]), np.float32([[0, 0, 0]]), np.float32([[0, 0, 0]])]
    cms = np.float32([[4, 4, 4], [5.8, 4.3, 4], [9, 9, 9]])
    sc = body_scene(offs, cms, vels=[[1, 0, 0], [-1, 0, 0], [0, 0, 0]], angs=[[0, 0, 2], [0, 0, 0], [0, 0, 0]], p=p)
    st = oracle.State(sc)
    st.pose_grains()
    # particle 1 (body 0) touches particle 2 (body 1); particle 3 (body 2) has a static contact
    dist = np.float32([0, 0.2, 0.2, 0])
    body = np.int32([-1, 1, 0, -1])
    part = np.int32([-1, 2, 1, -1])
    sdist = np.float32([0, 0, 0, 0.1])
    spart = np.int32([-1, -1, -1, 0])
    snorm = np.float32([[0, 1, 0, 0]])
    c = st.sis_gather(sdist, spart, snorm, dist, body, part)
    assert c["bodies"].tolist() == [[2, -1], [0, 1], [1, 0]]       # static contacts first, then body order
    p1, p2 = st.pos[1, :3].astype(np.float64), st.pos[2, :3].astype(np.float64)
    n = (p1 - p2) / np.linalg.norm(p1 - p2)                          # from the partner towards the particle
    assert np.allclose(c["point"][1, :3], (p1 + p2) / 2, atol=1e-6)
    assert np.allclose(c["normal"][1, :3], n, atol=1e-6)
    assert np.allclose(c["normal"][2, :3], -n, atol=1e-6)
    v1 = np.float64([1, 0, 0]) + np.cross([0, 0, 2], st.rel[1, :3])
    v2 = np.float64([-1, 0, 0])
    assert np.isclose(c["bias"][1], 0.7 * (v1 @ n - v2 @ n), rtol=1e-5)
    assert np.allclose(c["point"][0, :3], st.pos[3, :3]) and c["bias"][0] == 0.0
    assert np.allclose(c["normal"][0], snorm[0])


def test_gather_skips_grain_partners():
    p = params()
    sc = body_scene([np.float32([[0, 0, 0]])], np.float32([[4, 4, 4]]), p=p)
    st = oracle.State(sc)
    st.pose_grains()
    c = st.sis_gather(None, None, None, np.float32([0.3]), np.int32([-1]), np.int32([5]))
    assert len(c["bias"]) == 0


# ------------------------------------------------------------------ solver: closed forms
def one_static_contact(v, w, point, normal, mass=2.0, inv_i=0.25, ar_rest=0.5, ar_fric=0.0, iterations=1):
    p = params(ar_rest=ar_rest, ar_fric=ar_fric)
    sc = body_scene([np.float32([[0, 0, 0]])], np.float32([[0, 0, 0]]), vels=[v], angs=[w],
                    inv_inertia=(np.eye(3) * inv_i).reshape(9), masses=[mass], p=p)
    st = oracle.State(sc)
    r = np.float64(point)
    n = np.float64(normal)
    vrel0 = (np.float64(v) + np.cross(w, r)) @ n
    c = {"normal": np.float32([[*n, 0]]), "point": np.float32([[*r, 0]]), "bodies": np.int32([[0, -1]]),
         "bias": np.float32([ar_rest * vrel0]), "impulse": np.zeros((1, 4), np.float32)}
    st.sis_solve(c, iterations)
    vel = st.rb_vel[0, :3].astype(np.float64) + np.cross(st.rb_ang[0, :3].astype(np.float64), r)
    return st, c, vrel0, vel


def test_head_on_static_contact_restitution():
    st, c, vrel0, vel = one_static_contact([0, -2, 0], [0, 0, 0], [0, -1, 0], [0, 1, 0])
    assert vrel0 == -2.0
    assert np.isclose(vel[1], 1.0, atol=1e-6)                         # v_n' = -e v_n
    assert np.isclose(c["impulse"][0, 0], 3.0 * 2.0, rtol=1e-6)       # j = (1 + e) |v_n| m
    assert np.allclose(st.rb_ang[0, :3], 0)


def test_offset_static_contact_hits_target_in_one_sweep():
    # the effective mass 1/m + ((I^-1 (p x n)) x p) . n must make one sweep land exactly on -e v_rel
    st, c, vrel0, vel = one_static_contact([0.3, -2, 0.1], [0.5, 0, -0.2], [0.6, -1, 0.2], [0, 1, 0])
    assert vrel0 < 0
    assert np.isclose(vel[1], -0.5 * vrel0, rtol=1e-5)
    assert np.linalg.norm(st.rb_ang[0, :3]) > 0.05                    # off-centre impulse spins the body


def test_separating_contact_gets_no_impulse():
    st, c, vrel0, vel = one_static_contact([0, 2, 0], [0, 0, 0], [0, -1, 0], [0, 1, 0], iterations=18)
    assert c["impulse"][0, 0] == 0 and np.allclose(st.rb_vel[0, :3], [0, 2, 0])


@pytest.mark.parametrize("normal,tang", [([0, 1, 0], [1, 0, 0.5]), ([1, 0, 0], [0, 1, 0.3]), ([-1, 0, 0], [0, 0.7, 1])])
def test_static_friction_stops_tangential_slip(normal, tang):
    # contact point on the line through the centre along -n, isotropic inertia: the two friction
    # directions decouple and a wide friction bound removes the tangential velocity in one sweep
    n = np.float64(normal)
    v = -1.0 * n + np.float64(tang)
    st, c, vrel0, vel = one_static_contact(v, [0, 0, 0], -n, n, ar_fric=50.0)
    assert np.isclose(vel @ n, -0.5 * vrel0, atol=1e-5)
    assert np.linalg.norm(vel - (vel @ n) * n) < 1e-5


def test_static_friction_is_bounded():
    st, c, vrel0, vel = one_static_contact([3, -1, 0], [0, 0, 0], [0, -1, 0], [0, 1, 0], ar_fric=0.01)
    mc = 1 / 2.0
    bound = 0.01 * mc * 9.81
    assert np.isclose(abs(c["impulse"][0, 1]) + abs(c["impulse"][0, 2]), bound, rtol=1e-5)
    assert np.isclose(st.rb_vel[0, 0], 3 - bound / 2.0, rtol=1e-5)


def test_body_contact_conserves_momentum():
    rng = np.random.default_rng(7)
    inv = rng.uniform(0.2, 0.6, size=(2, 3))
    invI = np.stack([np.diag(a) for a in inv])                      # diagonal (symmetric) inverse inertia
    p = params(rb_rest=1.0)
    sc = body_scene([np.float32([[0, 0, 0]]), np.float32([[0, 0, 0]])], np.float32([[1, 2, 3], [2.1, 2.3, 2.8]]),
                    vels=[[1, 0.2, 0], [-1, 0, 0.3]], angs=[[0.1, 0.2, 0.3], [0, -0.4, 0.2]],
                    inv_inertia=invI.transpose(0, 2, 1).reshape(2, 9), masses=[3.0, 5.0], p=p)
    st = oracle.State(sc)
    x = st.rb_pos[:, :3].astype(np.float64)
    m = np.float64([3.0, 5.0])
    I = np.stack([np.linalg.inv(a) for a in invI])

    def momenta():
        v, w = st.rb_vel[:, :3].astype(np.float64), st.rb_ang[:, :3].astype(np.float64)
        P = (m[:, None] * v).sum(0)
        L = sum(I[k] @ w[k] + m[k] * np.cross(x[k], v[k]) for k in range(2))
        return P, L

    P0, L0 = momenta()
    pt = (x[0] + x[1]) / 2
    n = (x[0] - x[1]) / np.linalg.norm(x[0] - x[1])
    c = {"normal": np.float32([[*n, 0]]), "point": np.float32([[*pt, 0]]), "bodies": np.int32([[0, 1]]),
         "bias": np.float32([0.0]), "impulse": np.zeros((1, 4), np.float32)}
    st.sis_solve(c, 18)
    P1, L1 = momenta()
    assert c["impulse"][0, 0] > 0.1
    assert np.allclose(P1, P0, atol=1e-5) and np.allclose(L1, L0, atol=1e-5)
    # with bias 0 the normal relative velocity at the contact ends at 0
    v, w = st.rb_vel[:, :3].astype(np.float64), st.rb_ang[:, :3].astype(np.float64)
    vrel = (v[0] + np.cross(w[0], pt - x[0])) @ n - (v[1] + np.cross(w[1], pt - x[1])) @ n
    assert abs(vrel) < 1e-5


def test_sweeps_are_sequential_gauss_seidel():
    # two static contacts on one body: the second sees the first one's update within the same sweep,
    # so one sweep differs from applying both impulses computed from the initial state
    p = params(ar_fric=0.0)
    sc = body_scene([np.float32([[0, 0, 0]])], np.float32([[0, 0, 0]]), vels=[[0, -2, 0]], masses=[1.0], p=p)
    st = oracle.State(sc)
    c = {"normal": np.float32([[0, 1, 0, 0], [0, 1, 0, 0]]), "point": np.float32([[0.5, -1, 0, 0], [-0.5, -1, 0, 0]]),
         "bodies": np.int32([[0, -1], [0, -1]]), "bias": np.float32([-1.0, -1.0]), "impulse": np.zeros((2, 4), np.float32)}
    st.sis_solve(c, 1)
    j1, j2 = c["impulse"][:, 0]
    assert j1 > 0 and j2 >= 0 and not np.isclose(j1, j2)


def test_f32_mirror_tracks_f64_solve():
    rng = np.random.default_rng(11)
    nb = 6
    cms = rng.uniform(2, 8, size=(nb, 3)).astype(np.float32)
    sc = body_scene([np.float32([[0, 0, 0]])] * nb, cms, vels=rng.normal(size=(nb, 3)), angs=rng.normal(size=(nb, 3)),
                    masses=rng.uniform(1, 4, nb), p=params(ar_fric=0.3))
    pairs = [(0, 1), (1, 2), (2, 0), (3, 4), (4, -1), (5, -1), (0, -1), (3, 5)]
    k = len(pairs)
    n = rng.normal(size=(k, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pts = cms[[a for a, _ in pairs]] + rng.normal(scale=0.3, size=(k, 3))
    base = {"normal": np.c_[n, np.zeros(k)].astype(np.float32), "point": np.c_[pts, np.zeros(k)].astype(np.float32),
            "bodies": np.int32(pairs), "bias": rng.normal(scale=0.2, size=k).astype(np.float32)}
    out = []
    for dt in (np.float32, np.float64):
        st = oracle.State(sc, dt)
        c = dict(base, impulse=np.zeros((k, 4), np.float32))
        st.sis_solve(c, 18)
        out.append((st.rb_vel[:, :3].astype(np.float64), st.rb_ang[:, :3].astype(np.float64)))
    scale = max(np.abs(out[1][0]).max(), np.abs(out[1][1]).max())
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e-4 * scale
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-4 * scale


# ------------------------------------------------------------------ full step
def floor_points(lo=1.0, hi=15.0, y=1.0, pitch=0.5):
    g = np.arange(lo, hi, pitch, dtype=np.float32)
    xx, zz = np.meshgrid(g, g, indexing="ij")
    pts = np.zeros((xx.size, 4), np.float32)
    pts[:, 0], pts[:, 1], pts[:, 2] = xx.ravel(), y, zz.ravel()
    nrm = np.zeros_like(pts)
    nrm[:, 1] = 1
    return pts, nrm


def test_sis_step_rests_bodies_on_static_floor():
    p = params(ar_fric=0.3, rb_rest=0.3)
    cms = np.float32([[5, 2.2, 5], [9, 2.6, 9], [5.5, 4.5, 5.2]])
    sc = body_scene([cube(2)] * 3, cms, vels=[[0, -1, 0], [0.5, -2, 0], [0, -1, 0]], masses=[8.0] * 3,
                    inv_inertia=(np.eye(3) / 4.0).reshape(9), p=p, c=cfg(dt=2e-3))
    st = oracle.State(sc)
    st.set_static(*floor_points())
    ymin, gap = [], []
    for _ in range(400):
        st.step(1)
        ymin.append(st.pos[:, 1].min())
        x = st.pos[:, :3].astype(np.float64)
        d = np.linalg.norm(x[:, None] - x[None], axis=2)
        d[st.rb[:, None] == st.rb[None]] = np.inf
        gap.append(d.min())
    assert np.isfinite(st.pos).all()
    # the floor's points sit at y = 1; particles of radius 0.5 rest about 1 above them
    assert min(ymin) > 1.5 and st.pos[:, 1].min() < 2.3
    assert np.abs(st.rb_vel[:2, 1]).max() < 0.2
    # the dropped body lands on the first one without sinking into it
    assert min(gap) > 0.7 and st.rb_pos[2, 1] > st.rb_pos[0, 1] + 1.5


def test_sis_step_refuses_grains_and_lbvh():
    sc = scenes.mix_scene(1024, 1)
    sc.cfg.flags = capi.CONTACT_SIS
    with pytest.raises(AssertionError):
        oracle.State(sc).step(1)
