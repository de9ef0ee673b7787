"""Pins the oracle's floating-point stages with closed forms, known answers and invariants.

- pair force: head-on overlap (spring + dashpot), pure tangential slide (shear only), no contact
  at d >= 2r, exact antisymmetry F(a,b) = -F(b,a);
- walls: overlap with the floor gives k (r - d) upward plus damping of the normal velocity;
- integration: free flight under gravity, rotation of a body at constant ω by θ = |ω| dt,
  unit-norm quaternions, inverse inertia of a 4x4x4 cube = I / 160;
- orientation (Kernels/rigidBodyKernels.cu:319-322): from a non-identity orientation the
  increment composes on the world side (R_z(θ) R_x(90°), checked against Rodrigues matrices),
  and the world inverse inertia is R I_b^-1 R^T for anisotropic I_b^-1 (a quarter turn about z
  swaps two moments; 120° about (1,1,1) permutes them exactly; a generic rotation against
  float64), and a torque on such a body gives the hand-computed Δω.  Reversing the quaternion
  product or transposing the inertia product fails these tests;
- segmented reduction: a single particle force at r gives Δv = F dt / m, Δω = I^-1 (r x F) dt;
- FP32 mirror step vs FP64 step on the small scene: agreement to FP32 rounding.
"""
import math

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import capi, scenes


def pc(k=5e4, c=50.0, s=10.0, r=0.5, dt=1e-4):
    p = capi.ArSimParams()
    p.spring, p.damping, p.shear, p.attraction, p.particleRadius = k, c, s, 0.0, r
    p.globalDamping = p.Wdamping = 1.0
    p.gravity.y = -9.81
    for a in ("x", "y", "z"):
        setattr(p.gridSize, a, 64)
        setattr(p.cellSize, a, 2 * r)
    p.numCells = 64 ** 3
    cfg = capi.ArSimStepConfig()
    cfg.dt = dt
    for a in range(3):
        cfg.boxMin[a], cfg.boxMax[a] = 0.0, 64.0
    cfg.wallMask = capi.WALLS_OPEN_TOP
    return p, cfg


def test_pair_force_head_on():
    p, cfg = pc()
    d = 0.75
    hit, F = oracle.pair_force_f32([0, 0, 0], [d, 0, 0], [1.0, 0, 0], [-1.0, 0, 0], p, cfg)
    assert hit
    # F = -k (2r - d) n + c v_rel, n = +x, v_rel = -2 x
    expect = -5e4 * (1.0 - d) + 50.0 * (-2.0)
    assert F[0] == pytest.approx(expect, rel=1e-6)
    assert F[1] == 0 and F[2] == 0


def test_pair_force_tangential_slide():
    p, cfg = pc(k=0.0, c=0.0)
    hit, F = oracle.pair_force_f32([0, 0, 0], [0, 0.9, 0], [0, 0, 0], [2.0, 0, 0], p, cfg)
    assert hit
    assert F[0] == pytest.approx(10.0 * 2.0, rel=1e-6) and F[1] == 0 and F[2] == 0


def test_pair_force_no_contact_at_contact_distance():
    p, cfg = pc()
    hit, F = oracle.pair_force_f32([0, 0, 0], [1.0, 0, 0], [0, 0, 0], [0, 0, 0], p, cfg)
    assert not hit and not F.any()
    hit, _ = oracle.pair_force_f32([0, 0, 0], [0.99999, 0, 0], [0, 0, 0], [0, 0, 0], p, cfg)
    assert hit


def test_pair_force_antisymmetric_bitwise():
    p, cfg = pc()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        pa = rng.uniform(0, 10, 3).astype(np.float32)
        pb = (pa + rng.uniform(-0.6, 0.6, 3)).astype(np.float32)
        va, vb = rng.normal(size=3).astype(np.float32), rng.normal(size=3).astype(np.float32)
        h1, f1 = oracle.pair_force_f32(pa, pb, va, vb, p, cfg)
        h2, f2 = oracle.pair_force_f32(pb, pa, vb, va, p, cfg)
        assert h1 == h2
        assert np.array_equal(f1, -f2)


def test_wall_floor_known_answer():
    p, cfg = pc()
    F = oracle.wall_force_f64([10.0, 0.3, 10.0], [0.5, -2.0, 0.0], p, cfg)
    # floor: overlap r - d = 0.2, normal +y: k*0.2 up; dashpot -c v; shear -s v_t
    assert F[1] == pytest.approx(5e4 * 0.2 + 50 * 2.0)
    assert F[0] == pytest.approx(-50 * 0.5 - 10 * 0.5)
    F = oracle.wall_force_f64([10.0, 0.7, 10.0], [0.5, -2.0, 0.0], p, cfg)
    assert not F.any()
    cfg.wallMask = 0
    assert not oracle.wall_force_f64([10.0, 0.1, 10.0], [0, 0, 0], p, cfg).any()


def test_sincos_accuracy():
    x = np.concatenate([np.linspace(-20, 20, 20001), np.geomspace(1e-8, 1e-2, 500)]).astype(np.float32)
    s, c = oracle.sincos(x)
    xd = x.astype(np.float64)
    assert np.max(np.abs(s - np.sin(xd))) < 2e-7
    assert np.max(np.abs(c - np.cos(xd))) < 2e-7


def _single_body_scene():
    sc = scenes.config0()
    return sc


def test_cube_inertia_and_mass():
    sc = scenes.config0()
    assert sc.n == 16384 and sc.b == 256
    assert np.all(sc.ppo == 64) and np.all(sc.rb_mass == 64)
    # Σ over 4x4x4 unit-spaced points of (y² + z²) = 160, isotropic, off-diagonals zero
    I = sc.rb_inertia[0].reshape(3, 3)
    assert np.allclose(I, np.eye(3) / 160.0, atol=1e-9)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_free_flight_grain(dtype):
    sc = scenes.mix_scene(2048, 3)
    st = oracle.State(sc, dtype)
    g = np.where(st.rb < 0)[0][0]
    st.pos[g, :3] = [10, 20, 30]
    st.vel[g, :3] = [1, 2, 3]
    st.pose_grains()
    dt = dtype(1e-4)
    v = np.array([1, 2 + dtype(-9.81) * dt, 3], dtype)
    assert np.allclose(st.vel[g, :3], v, rtol=1e-6)
    assert np.allclose(st.pos[g, :3], np.array([10, 20, 30], dtype) + v * dt, rtol=1e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_body_rotation_closed_form(dtype):
    sc = scenes.config0()
    st = oracle.State(sc, dtype)
    theta = 0.01
    dt = float(sc.cfg.dt)
    st.rb_ang[:] = 0
    st.rb_ang[:, 2] = theta / dt
    st.qcum[:] = [0, 0, 0, 1]
    st.rb_vel[:] = 0
    st.integrate_bodies()
    # the incremental quaternion rotates about z by theta
    assert np.allclose(st.qinc[0], [0, 0, math.sin(theta / 2), math.cos(theta / 2)], atol=2e-7)
    st.body_frame[:64] = 0
    st.body_frame[:64, 0] = 1.0
    st.pose_grains()
    r = st.rel[0, :3]
    assert np.allclose(r, [math.cos(theta), math.sin(theta), 0], atol=2e-7)
    assert np.allclose(np.linalg.norm(st.qcum, axis=1), 1, atol=2e-7)


def _rot(axis, angle):
    """Rotation matrix about a unit axis (Rodrigues), float64: independent of quaternions."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)


def _quat(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2)
    return np.array([a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2)])


def _body_state(dtype, q0, inv_inertia=None, omega=(0.0, 0.0, 0.0)):
    """Oracle state of the small scene with body 0 at orientation q0 (x, y, z, w), inverse body
    inertia inv_inertia (3x3) and angular velocity omega; no linear velocity."""
    sc = scenes.config0()
    st = oracle.State(sc, dtype)
    st.qcum[0] = np.asarray(q0, dtype)
    if inv_inertia is not None:
        st.Ib[0] = np.asarray(inv_inertia, np.float64).T.reshape(9).astype(dtype)   # column-major
    st.rb_ang[0, :3] = omega
    st.rb_vel[0, :3] = 0
    return sc, st


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_body_rotation_composes_in_world_frame(dtype):
    # a body turned 90 degrees about x, then spun about the world z axis by theta in one step:
    # the new orientation is R_z(theta) R_x(90) (the increment acts after the orientation it
    # updates), so the body-frame x axis ends at (cos theta, sin theta, 0); the reversed
    # product R_x(90) R_z(theta) would put it at (cos theta, 0, sin theta)
    theta = 0.3
    sc, st = _body_state(dtype, _quat((1, 0, 0), math.pi / 2), omega=(0.0, 0.0, theta / float(sc_dt())))
    st.integrate_bodies()
    R = _rot((0, 0, 1), theta) @ _rot((1, 0, 0), math.pi / 2)
    Rm = st.model[0].reshape(4, 4).T[:3, :3]   # column-major mat4 of the cumulative rotation
    tol = 2e-6 if dtype == np.float32 else 1e-12
    assert np.allclose(Rm, R, atol=tol)
    st.body_frame[0, :3] = [1.0, 0.0, 0.0]
    st.body_frame[1, :3] = [0.0, 1.0, 0.0]
    st.pose_grains()
    assert np.allclose(st.rel[0, :3], [math.cos(theta), math.sin(theta), 0.0], atol=tol)
    assert np.allclose(st.rel[1, :3], R @ [0.0, 1.0, 0.0], atol=tol)
    assert np.allclose(st.rel[1, :3], [0.0, 0.0, 1.0], atol=tol)
    # a second step about z again: R_z(2 theta) R_x(90)
    st.integrate_bodies()
    R2 = _rot((0, 0, 1), 2 * theta) @ _rot((1, 0, 0), math.pi / 2)
    assert np.allclose(st.model[0].reshape(4, 4).T[:3, :3], R2, atol=2 * tol)


def sc_dt():
    return scenes.config0().cfg.dt


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_world_inverse_inertia_quarter_turn_about_z(dtype):
    # I_b^-1 = diag(a, b, c) turned 90 degrees about z exchanges the x and y moments: diag(b, a, c)
    a, b, c = 0.5, 1.0 / 3.0, 0.2
    _, st = _body_state(dtype, _quat((0, 0, 1), math.pi / 2), np.diag([a, b, c]))
    st.integrate_bodies()
    Iw = st.Iw[0].reshape(3, 3).T
    assert np.allclose(Iw, np.diag([b, a, c]), atol=2e-7 if dtype == np.float32 else 1e-14)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_world_inverse_inertia_is_r_i_rt(dtype):
    # 120 degrees about (1,1,1) sends x -> y -> z -> x, and its quaternion (1/2, 1/2, 1/2, 1/2) is
    # exact: R diag(a, b, c) R^T = diag(c, a, b) exactly, while R^T diag(a, b, c) R = diag(b, c, a)
    a, b, c = 0.5, 1.0 / 3.0, 0.2
    _, st = _body_state(dtype, [0.5, 0.5, 0.5, 0.5], np.diag([a, b, c]))
    st.integrate_bodies()
    Iw = st.Iw[0].reshape(3, 3).T
    assert np.array_equal(Iw, np.diag([c, a, b]).astype(dtype))
    # a generic orientation and a full symmetric inverse inertia against R I R^T in float64
    rng = np.random.default_rng(5)
    M = rng.normal(size=(3, 3))
    Ib = M @ M.T + 3 * np.eye(3)
    axis, ang = rng.normal(size=3), 1.1
    _, st = _body_state(dtype, _quat(axis, ang), Ib)
    st.integrate_bodies()
    R = _rot(axis, ang)
    Ib_used = st.Ib[0].reshape(3, 3).T.astype(np.float64)
    assert np.allclose(st.Iw[0].reshape(3, 3).T, R @ Ib_used @ R.T, rtol=1e-5 if dtype == np.float32 else 1e-12,
                       atol=1e-6 if dtype == np.float32 else 1e-13)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_torque_on_anisotropic_rotated_body(dtype):
    # the body of test_world_inverse_inertia_is_r_i_rt (I_w^-1 = diag(c, a, b)); a particle at body
    # offset (1, 2, 1/2) sits at r = (1/2, 1, 2) in the world; a force F there turns the body by
    # dω = diag(c, a, b) (r x F) dt, computed here by hand
    a, b, c = 0.5, 1.0 / 3.0, 0.2
    sc, st = _body_state(dtype, [0.5, 0.5, 0.5, 0.5], np.diag([a, b, c]))
    st.body_frame[3, :3] = [1.0, 2.0, 0.5]
    st.integrate_bodies()
    st.pose_grains()
    assert np.array_equal(st.rel[3, :3], np.array([0.5, 1.0, 2.0], dtype))
    st.force[:] = 0
    st.torque[:] = 0
    F = np.array([3.0, -1.0, 2.0])
    r = np.array([0.5, 1.0, 2.0])
    st.force[3, :3] = F
    st.torque[3, :3] = np.cross(r, F)
    st.reduce_apply()
    dt = float(sc.cfg.dt)
    tau = np.array([1.0 * 2.0 - 2.0 * -1.0, 2.0 * 3.0 - 0.5 * 2.0, 0.5 * -1.0 - 1.0 * 3.0])   # r x F = (4, 5, -3.5)
    dw = np.array([c * tau[0], a * tau[1], b * tau[2]]) * dt
    assert np.allclose(st.rb_ang[0, :3], dw, rtol=1e-6 if dtype == np.float32 else 1e-13, atol=0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reduce_single_particle_force(dtype):
    sc = scenes.config0()
    st = oracle.State(sc, dtype)
    st.integrate_bodies()
    st.pose_grains()
    st.force[:] = 0
    st.torque[:] = 0
    F = np.array([3.0, -1.0, 2.0], dtype)
    r = st.rel[5, :3].copy()
    st.force[5, :3] = F
    st.torque[5, :3] = np.cross(r, F)
    v0, w0 = st.rb_vel[0, :3].copy(), st.rb_ang[0, :3].copy()
    st.reduce_apply()
    dt = float(sc.cfg.dt)
    # the difference v - v0 is exact to one rounding of v: allow 2 ulp of |v| in FP32
    atol = 2 * np.spacing(np.abs(st.rb_vel[0, :3]).astype(dtype)).max() if dtype == np.float32 else 1e-15
    assert np.allclose(st.rb_vel[0, :3] - v0, F * dt / 64.0, rtol=1e-5, atol=atol)
    Iw = st.Iw[0].reshape(3, 3).T     # column-major storage
    assert np.allclose(st.rb_ang[0, :3] - w0, Iw @ np.cross(r, F) * dt, rtol=1e-4, atol=atol)
    # bodies without a force only feel gravity
    st2 = oracle.State(sc, dtype)
    st2.integrate_bodies()
    assert np.array_equal(st.rb_vel[1], st2.rb_vel[1])


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30, capi.METHOD_LBVH63])
def test_fp32_mirror_tracks_fp64_over_steps(method):
    sc = scenes.mix_scene(4096, 11, method)
    a, b = oracle.State(sc, np.float32), oracle.State(sc, np.float64)
    a.step(20)
    b.step(20)
    assert np.max(np.abs(a.pos - b.pos)) < 1e-3
    assert np.max(np.abs(a.rb_vel - b.rb_vel)) < 1e-3


def test_grid_and_lbvh_paths_give_same_contacts_brute():
    sc = scenes.mix_scene(3000, 5)
    st = oracle.State(sc, np.float32)
    st.pos[:, :3] += np.random.default_rng(0).normal(scale=0.15, size=(sc.n, 3)).astype(np.float32)
    pairs = oracle.contacts_brute(st.pos, st.rb, 0.5)
    # numpy brute force
    P = st.pos[:, :3].astype(np.float32)
    ref = set()
    for i in range(sc.n):
        d = P[i + 1:] - P[i]
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        js = np.nonzero(dist < np.float32(1.0))[0] + i + 1
        for j in js:
            if st.rb[i] != st.rb[j] or st.rb[i] == -1:
                ref.add((i, int(j)))
    assert pairs == ref and len(ref) > 0
