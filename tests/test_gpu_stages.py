"""Stage-level parity of the gfx950 kernels against the oracle, through the C ABI.

Every stage gets identical inputs on both sides.  Integer / index stages (hash, Morton codes,
sort, cell bounds, LBVH topology and bounds, contact sets) must be bit-exact.  FP32 stages must
match the oracle's FP32 mirror bit for bit and the FP64 model within 1e-5 relative to
max(|F|, k r) for forces (k r is the spring force at half a radius of overlap) and 1e-5 relative
for integrated states.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import capi, scenes
from gpu_util import Dev, rel_err_vs_scale, sync, zeros

pytestmark = pytest.mark.gpu
L = capi.lib


def jittered_scene(n=None, seed=0, sigma=0.25, squeeze=0.55):
    """Small-scene generator output squeezed towards its centre and jittered, so that many pairs
    (grain-grain, grain-body, body-body) overlap; positions are used as they are."""
    sc = scenes.config0() if n is None else scenes.mix_scene(n, seed)
    rng = np.random.default_rng(seed)
    c = sc.pos[:, :3].mean(0)
    old = sc.rb_pos[:, :3].copy()
    sc.rb_pos[:, :3] = c + (old - c) * np.float32(squeeze)       # bodies move closer, keep their shape
    body = sc.rb >= 0
    sc.pos[body, :3] += sc.rb_pos[sc.rb[body], :3] - old[sc.rb[body]]
    sc.pos[~body, :3] = c + (sc.pos[~body, :3] - c) * np.float32(squeeze)
    sc.pos[:, :3] += rng.normal(scale=sigma, size=(sc.n, 3)).astype(np.float32)
    sc.vel[:, :3] = rng.normal(scale=0.5, size=(sc.n, 3)).astype(np.float32)
    return sc


def test_calc_hash_bit_exact():
    print("stage: calc_hash", flush=True)
    rng = np.random.default_rng(0)
    for sc_pos, grid in [(scenes.config0().pos, (64, 64, 64)),
                         (np.concatenate([rng.uniform(-200, 600, (200000, 3)), np.zeros((200000, 1))], 1), (256, 128, 64))]:
        pos = np.ascontiguousarray(sc_pos, np.float32)
        p = scenes.config0().params
        p.gridSize.x, p.gridSize.y, p.gridSize.z = grid
        p.numCells = grid[0] * grid[1] * grid[2]
        n = pos.shape[0]
        dp, dh, di = Dev(pos), zeros(n, np.uint32), zeros(n, np.uint32)
        capi.check(L.arsim_calc_hash(dh.p, di.p, dp.p, n, C.byref(p), None))
        sync()
        h, i = oracle.calc_hash(pos, p)
        assert np.array_equal(dh.get(), h)
        assert np.array_equal(di.get(), i)


def test_calc_hash_rejects_non_pow2_grid():
    p = scenes.config0().params
    p.gridSize.x = 48
    d = zeros(4, np.float32)
    assert L.arsim_calc_hash(d.p, d.p, d.p, 1, C.byref(p), None) == capi.ARSIM_ERR_NOT_POW2


def test_morton_bit_exact():
    print("stage: morton", flush=True)
    rng = np.random.default_rng(1)
    n = 300000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-10, 530, (n, 3))
    bmin, bmax = np.zeros(3, np.float32), np.full(3, 512, np.float32)
    dp, dc, di = Dev(pos), zeros(n, np.uint32), zeros(n, np.int32)
    capi.check(L.arsim_morton30(dp.p, dc.p, di.p, n, capi.ptr(bmin), capi.ptr(bmax), None, None))
    sync()
    assert np.array_equal(dc.get(), oracle.morton30(pos, bmin, bmax))
    assert np.array_equal(di.get(), np.arange(n, dtype=np.int32))
    dc64 = zeros(n, np.uint64)
    capi.check(L.arsim_morton63(dp.p, dc64.p, di.p, n, capi.ptr(bmin), capi.ptr(bmax), None, None))
    sync()
    assert np.array_equal(dc64.get(), oracle.morton63(pos, bmin, bmax))


def _gpu_sort(keys, vals, begin, end):
    n = keys.size
    kb = 4 if keys.dtype == np.uint32 else 8
    dk, dv = Dev(keys), Dev(vals)
    ak, av = zeros(n, keys.dtype), zeros(n, np.uint32)
    ws_bytes = L.arsim_sort_workspace_bytes(n, kb)
    ws = Dev(shape=(max(ws_bytes, 4),), dtype=np.uint8)
    pk, pv, pak, pav = (C.c_void_p(x.p.value) for x in (dk, dv, ak, av))
    fn = L.arsim_sort_pairs_u32 if kb == 4 else L.arsim_sort_pairs_u64
    capi.check(fn(C.byref(pk), C.byref(pv), C.byref(pak), C.byref(pav), n, begin, end, ws.p, ws_bytes, None))
    sync()
    # the result is where *keys points after the call (pointer exchange like cub::DoubleBuffer)
    src_k = dk if pk.value == dk.p.value else ak
    src_v = dv if pv.value == dv.p.value else av
    assert {pk.value, pak.value} == {dk.p.value, ak.p.value}
    return src_k.get(), src_v.get()


@pytest.mark.parametrize("n", [1, 2, 1000, 4096, 4097, 65537, 1 << 20])
def test_sort_u32_bit_exact(n):
    print(f"stage: sort u32 n={n}", flush=True)
    rng = np.random.default_rng(n)
    cases = [(rng.integers(0, 1 << 18, n, dtype=np.uint32), 0, 18),
             (rng.integers(0, 1 << 24, n, dtype=np.uint32), 0, 24),
             (rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32), 0, 32),
             (rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32), 8, 20),
             (np.full(n, 12345, np.uint32), 0, 32),
             (np.arange(n, dtype=np.uint32)[::-1].copy(), 0, 32),
             (rng.integers(0, 3, n, dtype=np.uint32), 0, 2)]
    for keys, b, e in cases:
        vals = rng.integers(0, 1 << 30, n, dtype=np.uint32)
        gk, gv = _gpu_sort(keys, vals, b, e)
        digits = (keys.astype(np.uint64) >> np.uint64(b)) & np.uint64((1 << (e - b)) - 1)
        order = np.argsort(digits, kind="stable")
        assert np.array_equal(gk, keys[order]), (n, b, e)
        assert np.array_equal(gv, vals[order]), (n, b, e)


@pytest.mark.parametrize("n", [3, 5000, 300001])
def test_sort_u64_bit_exact(n):
    print(f"stage: sort u64 n={n}", flush=True)
    rng = np.random.default_rng(7 + n)
    for keys, b, e in [(rng.integers(0, 1 << 63, n, dtype=np.uint64), 0, 63),
                       (rng.integers(0, 1 << 39, n, dtype=np.uint64), 0, 39),
                       (np.full(n, 1 << 40, np.uint64), 0, 64)]:
        vals = np.arange(n, dtype=np.uint32)
        gk, gv = _gpu_sort(keys, vals, b, e)
        ok, ov = oracle.sort_pairs(keys, vals, b, e)
        assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def test_sort_workspace_check():
    d = zeros(16, np.uint32)
    p = [C.c_void_p(d.p.value) for _ in range(4)]
    rc = L.arsim_sort_pairs_u32(C.byref(p[0]), C.byref(p[1]), C.byref(p[2]), C.byref(p[3]), 16, 0, 32, d.p, 8, None)
    assert rc == capi.ARSIM_ERR_WORKSPACE


def grid_pipeline(sc):
    """hash -> sort -> reorder on the GPU; returns device buffers and host copies."""
    n = sc.n
    p = sc.params
    dpos, dvel, drb = Dev(sc.pos), Dev(sc.vel), Dev(sc.rb)
    dh, di, ah, ai = zeros(n, np.uint32), zeros(n, np.uint32), zeros(n, np.uint32), zeros(n, np.uint32)
    capi.check(L.arsim_calc_hash(dh.p, di.p, dpos.p, n, C.byref(p), None))
    ws_bytes = L.arsim_sort_workspace_bytes(n, 4)
    ws = Dev(shape=(ws_bytes,), dtype=np.uint8)
    pk, pv, pak, pav = (C.c_void_p(x.p.value) for x in (dh, di, ah, ai))
    capi.check(L.arsim_sort_pairs_u32(C.byref(pk), C.byref(pv), C.byref(pak), C.byref(pav), n, 0, 18, ws.p,
                                      ws_bytes, None))
    sk = dh if pk.value == dh.p.value else ah
    si = di if pv.value == di.p.value else ai
    ncell = p.numCells
    cs, ce = zeros(ncell, np.uint32), zeros(ncell, np.uint32)
    spos, svel = zeros((n, 4), np.float32), zeros((n, 4), np.float32)
    capi.check(L.arsim_reorder_find_cell_start(drb.p, cs.p, ce.p, spos.p, svel.p, sk.p, si.p, dpos.p, dvel.p, n,
                                               ncell, None))
    sync()
    return dict(pos=dpos, vel=dvel, rb=drb, sk=sk, si=si, cs=cs, ce=ce, spos=spos, svel=svel, keep=(ws, dh, di, ah, ai))


def test_reorder_cell_start_bit_exact():
    print("stage: reorder", flush=True)
    sc = jittered_scene()
    g = grid_pipeline(sc)
    h, i = oracle.calc_hash(sc.pos, sc.params)
    sh, si = oracle.sort_pairs(h, i, 0, 18)
    assert np.array_equal(g["sk"].get(), sh) and np.array_equal(g["si"].get(), si)
    cs, ce = oracle.cell_bounds(sh, sc.params.numCells)
    gcs, gce = g["cs"].get(), g["ce"].get()
    assert np.array_equal(gcs, cs)
    occupied = cs != 0xFFFFFFFF
    assert np.array_equal(gce[occupied], ce[occupied])
    spos, svel = g["spos"].get(), g["svel"].get()
    assert np.array_equal(spos[:, :3], sc.pos[si, :3])
    assert np.array_equal(spos[:, 3].view(np.int32), sc.rb[si])
    assert np.array_equal(svel, sc.vel[si])


def _forces_case(sc, flags=0):
    sc.cfg.flags = flags
    n = sc.n
    g = grid_pipeline(sc)
    # relative positions: give every body particle its offset from the body centre
    rel = np.zeros((n, 4), np.float32)
    body = sc.rb >= 0
    rel[body, :3] = sc.pos[body, :3] - sc.rb_pos[sc.rb[body], :3]
    drel = Dev(rel)
    dF, dT = zeros((n, 4), np.float32), zeros((n, 4), np.float32)
    capi.check(L.arsim_dem_grid(g["rb"].p, dF.p, dT.p, g["vel"].p, g["spos"].p, g["svel"].p, drel.p, g["si"].p,
                                g["cs"].p, g["ce"].p, n, C.byref(sc.params), C.byref(sc.cfg), None))
    sync()
    h, i = oracle.calc_hash(sc.pos, sc.params)
    sh, si = oracle.sort_pairs(h, i, 0, 18)
    cs, ce = oracle.cell_bounds(sh, sc.params.numCells)
    out = {}
    for dt in (np.float32, np.float64):
        st = oracle.State(sc, dt)
        st.rel[:] = rel
        st.dem_grid(si, cs, ce, sc.pos[si].astype(dt), sc.vel[si].astype(dt), sc.rb[si])
        out[dt] = st
    return dF.get(), dT.get(), g["vel"].get(), out


@pytest.mark.parametrize("flags", [0, capi.COMPAT_REF_UNITS | capi.COMPAT_GRAIN_CLAMP])
def test_dem_grid_forces(flags):
    print(f"stage: dem_grid flags={flags}", flush=True)
    sc = jittered_scene(20000, seed=3)
    F, T, V, ref = _forces_case(sc, flags)
    body = sc.rb >= 0
    m32, m64 = ref[np.float32], ref[np.float64]
    assert np.array_equal(F[body], m32.force[body])
    assert np.array_equal(T[body], m32.torque[body])
    assert np.array_equal(V[~body], m32.vel[~body])
    kr = sc.params.spring * sc.params.particleRadius
    assert rel_err_vs_scale(F[body, :3], m64.force[body, :3], kr) < 1e-5
    assert rel_err_vs_scale(T[body, :3], m64.torque[body, :3], kr) < 1e-5
    assert np.count_nonzero(np.linalg.norm(F[body, :3], axis=1)) > 1000   # contacts really happened


def lbvh_pipeline(sc, bits=30):
    n = sc.n
    bmin = np.array(sc.cfg.boxMin[:], np.float32)
    bmax = np.array(sc.cfg.boxMax[:], np.float32)
    dpos, dvel, drb = Dev(sc.pos), Dev(sc.vel), Dev(sc.rb)
    kt = np.uint32 if bits == 30 else np.uint64
    dk, di, ak, ai = zeros(n, kt), zeros(n, np.uint32), zeros(n, kt), zeros(n, np.uint32)
    fn = L.arsim_morton30 if bits == 30 else L.arsim_morton63
    capi.check(fn(dpos.p, dk.p, di.p, n, capi.ptr(bmin), capi.ptr(bmax), None, None))
    kb = 4 if bits == 30 else 8
    ws_bytes = L.arsim_sort_workspace_bytes(n, kb)
    ws = Dev(shape=(ws_bytes,), dtype=np.uint8)
    pk, pv, pak, pav = (C.c_void_p(x.p.value) for x in (dk, di, ak, ai))
    sf = L.arsim_sort_pairs_u32 if kb == 4 else L.arsim_sort_pairs_u64
    capi.check(sf(C.byref(pk), C.byref(pv), C.byref(pak), C.byref(pav), n, 0, bits, ws.p, ws_bytes, None))
    sk = dk if pk.value == dk.p.value else ak
    si = di if pv.value == di.p.value else ai
    m = 2 * n
    leaf = zeros(m, np.uint8)
    left, right, parent = zeros(m, np.int32), zeros(m, np.int32), zeros(m, np.int32)
    minr, maxr, cnt = zeros(m, np.int32), zeros(m, np.int32), zeros(n, np.int32)
    capi.check(L.arsim_lbvh_build(sk.p, kb, n, leaf.p, left.p, right.p, parent.p, minr.p, maxr.p, cnt.p, None))
    cms, svel, bounds = zeros((n, 4), np.float32), zeros((n, 4), np.float32), zeros((m, 8), np.float32)
    capi.check(L.arsim_lbvh_refit(si.p, dpos.p, dvel.p, drb.p, sc.params.particleRadius, n, leaf.p, minr.p, maxr.p,
                                  left.p, right.p, parent.p, cms.p, svel.p, bounds.p, cnt.p, None))
    sync()
    return dict(pos=dpos, vel=dvel, rb=drb, sk=sk, si=si, leaf=leaf, left=left, right=right, parent=parent,
                minr=minr, maxr=maxr, cms=cms, svel=svel, bounds=bounds, keep=(ws, dk, di, ak, ai, cnt))


@pytest.mark.parametrize("bits", [30, 63])
def test_lbvh_build_refit_bit_exact(bits):
    print(f"stage: lbvh build/refit bits={bits}", flush=True)
    sc = jittered_scene()
    n = sc.n
    g = lbvh_pipeline(sc, bits)
    bmin, bmax = np.array(sc.cfg.boxMin[:], np.float32), np.array(sc.cfg.boxMax[:], np.float32)
    codes = oracle.morton30(sc.pos, bmin, bmax) if bits == 30 else oracle.morton63(sc.pos, bmin, bmax)
    sk, si = oracle.sort_pairs(codes, np.arange(n, dtype=np.uint32))
    assert np.array_equal(g["sk"].get(), sk) and np.array_equal(g["si"].get(), si)
    t = oracle.karras(sk)
    inner = slice(n, 2 * n - 1)
    for k in ("left", "right", "min", "max"):
        gk = {"left": "left", "right": "right", "min": "minr", "max": "maxr"}[k]
        assert np.array_equal(g[gk].get()[inner], t[k][inner]), k
    assert np.array_equal(g["parent"].get()[:2 * n - 1], t["parent"][:2 * n - 1])
    assert np.array_equal(g["leaf"].get()[:2 * n - 1], t["leaf"][:2 * n - 1])
    cms, bounds = oracle.refit(si.astype(np.int32), sc.pos, sc.params.particleRadius, t["left"], t["right"])
    gb = g["bounds"].get()
    assert np.array_equal(gb[:2 * n - 1, [0, 1, 2, 4, 5, 6]], bounds[:2 * n - 1, [0, 1, 2, 4, 5, 6]])
    gc = g["cms"].get()
    assert np.array_equal(gc[:, :3], cms[:, :3]) and np.array_equal(gc[:, 3].view(np.int32), sc.rb[si])


@pytest.mark.parametrize("bits", [30, 63])
def test_dem_lbvh_forces(bits):
    print(f"stage: dem_lbvh bits={bits}", flush=True)
    sc = jittered_scene(20000, seed=4)
    n = sc.n
    g = lbvh_pipeline(sc, bits)
    rel = np.zeros((n, 4), np.float32)
    body = sc.rb >= 0
    rel[body, :3] = sc.pos[body, :3] - sc.rb_pos[sc.rb[body], :3]
    drel = Dev(rel)
    dF, dT, ov = zeros((n, 4), np.float32), zeros((n, 4), np.float32), zeros(1, np.int32)
    capi.check(L.arsim_dem_lbvh(g["rb"].p, dF.p, dT.p, g["vel"].p, g["svel"].p, drel.p, g["si"].p, g["left"].p,
                                g["right"].p, g["cms"].p, g["bounds"].p, n, C.byref(sc.params), C.byref(sc.cfg), ov.p,
                                None))
    sync()
    assert ov.get()[0] == 0
    si = g["si"].get().astype(np.int32)
    left, right, bounds = g["left"].get(), g["right"].get(), g["bounds"].get()
    ref = {}
    for dt in (np.float32, np.float64):
        st = oracle.State(sc, dt)
        st.rel[:] = rel
        st.dem_lbvh(si, left, right, bounds, sc.pos[si].astype(dt), sc.vel[si].astype(dt), sc.rb[si])
        ref[dt] = st
    F, T, V = dF.get(), dT.get(), g["vel"].get()
    assert np.array_equal(F[body], ref[np.float32].force[body])
    assert np.array_equal(T[body], ref[np.float32].torque[body])
    assert np.array_equal(V[~body], ref[np.float32].vel[~body])
    kr = sc.params.spring * sc.params.particleRadius
    assert rel_err_vs_scale(F[body, :3], ref[np.float64].force[body, :3], kr) < 1e-5


def test_contact_sets_grid_lbvh_brute_equal():
    print("stage: contact sets", flush=True)
    sc = jittered_scene(seed=5)
    n = sc.n
    maxp = 64
    g = grid_pipeline(sc)
    pairs, counts = zeros((n, maxp), np.int32), zeros(n, np.int32)
    capi.check(L.arsim_contact_pairs_grid(g["rb"].p, g["spos"].p, g["si"].p, g["cs"].p, g["ce"].p, n,
                                          C.byref(sc.params), pairs.p, counts.p, maxp, None))
    sync()
    from diploma_thesis_amd import contact_set
    grid_set = contact_set(pairs.get(), counts.get())
    for bits in (30, 63):
        t = lbvh_pipeline(sc, bits)
        lp, lc, ov = zeros((n, maxp), np.int32), zeros(n, np.int32), zeros(1, np.int32)
        capi.check(L.arsim_contact_pairs_lbvh(t["rb"].p, t["si"].p, t["left"].p, t["right"].p, t["cms"].p,
                                              t["bounds"].p, n, C.byref(sc.params), ov.p, lp.p, lc.p, maxp, None))
        sync()
        assert ov.get()[0] == 0
        assert contact_set(lp.get(), lc.get()) == grid_set
    brute = oracle.contacts_brute(sc.pos, sc.rb, sc.params.particleRadius)
    assert grid_set == brute and len(brute) > 1000
    # each contact is reported from both sides
    c = counts.get()
    assert c.sum() == 2 * len(brute)


def test_body_integrate_pose_reduce_stages():
    print("stage: integrate/pose/reduce", flush=True)
    sc = scenes.config0()
    rng = np.random.default_rng(8)
    sc.rb_ang[:, :3] = rng.normal(scale=3.0, size=(sc.b, 3)).astype(np.float32)
    n, b = sc.n, sc.b
    dpos, dvel, dang = Dev(sc.rb_pos), Dev(sc.rb_vel), Dev(sc.rb_ang)
    qinc, qcum = zeros((b, 4), np.float32), Dev(sc.rb_quat)
    Ib, Iw, model = Dev(sc.rb_inertia), zeros((b, 9), np.float32), zeros((b, 16), np.float32)
    capi.check(L.arsim_integrate_bodies(dpos.p, dvel.p, dang.p, qinc.p, qcum.p, Ib.p, Iw.p, model.p, b,
                                        C.byref(sc.params), C.byref(sc.cfg), None))
    ppos, pvel, prel = Dev(sc.pos), Dev(sc.vel), zeros((n, 4), np.float32)
    dbf, drb = Dev(sc.body_frame), Dev(sc.rb)
    capi.check(L.arsim_particle_pose(dpos.p, dvel.p, dang.p, qcum.p, dbf.p, drb.p, prel.p, ppos.p, pvel.p, n, None))
    sync()
    st = oracle.State(sc, np.float32)
    st.integrate_bodies()
    st.pose_grains()
    for name, dev, host in [("rb_pos", dpos, st.rb_pos), ("rb_vel", dvel, st.rb_vel), ("rb_ang", dang, st.rb_ang),
                            ("qinc", qinc, st.qinc), ("qcum", qcum, st.qcum), ("Iw", Iw, st.Iw),
                            ("pos", ppos, st.pos), ("vel", pvel, st.vel), ("rel", prel, st.rel)]:
        assert np.array_equal(dev.get(), host), name
    mm = model.get().reshape(b, 4, 4)
    assert np.allclose(mm[:, 3, :3], st.rb_pos[:, :3]) and np.all(mm[:, 3, 3] == 1)
    # reduce + apply with random per-particle forces; torque from relativePos x F
    F = rng.normal(scale=100.0, size=(n, 4)).astype(np.float32)
    F[:, 3] = 0
    dF = Dev(F)
    ps, pc = Dev(sc.pstart), Dev(sc.ppo)
    dm = Dev(sc.rb_mass)
    capi.check(L.arsim_segmented_reduce_apply(dF.p, None, prel.p, dvel.p, dang.p, dm.p, Iw.p, ps.p, pc.p, b,
                                              C.byref(sc.params), C.byref(sc.cfg), None))
    sync()
    st.force[:] = F
    st.torque[:, :3] = np.stack([st.rel[:, 1] * F[:, 2] - st.rel[:, 2] * F[:, 1],
                                 st.rel[:, 2] * F[:, 0] - st.rel[:, 0] * F[:, 2],
                                 st.rel[:, 0] * F[:, 1] - st.rel[:, 1] * F[:, 0]], 1)
    st.reduce_apply()
    assert np.array_equal(dvel.get(), st.rb_vel)
    assert np.array_equal(dang.get(), st.rb_ang)
    # FP64 model of the same reduction, from the same inputs
    s64 = oracle.State(sc, np.float64)
    s64.rb_vel[:] = st.rb_vel.astype(np.float64)
    s64.rb_vel[:] = 0
    s64.rb_ang[:] = 0
    s64.Iw[:] = st.Iw
    s64.force[:] = F
    s64.torque[:] = st.torque
    s64.reduce_apply()
    s32 = oracle.State(sc, np.float32)
    s32.rb_vel[:] = 0
    s32.rb_ang[:] = 0
    s32.Iw[:] = st.Iw
    s32.force[:] = F
    s32.torque[:] = st.torque
    s32.reduce_apply()
    assert rel_err_vs_scale(s32.rb_vel[:, :3], s64.rb_vel[:, :3], 1e-12) < 1e-5
