"""Seeded synthetic scenes (host arrays) for the configurations the pipeline is measured on.

The generator itself is C++ (arsim_generate_scene / arsim_generate_mix_scene); this module
wraps it and copies the result into numpy arrays in the reference layout.
"""
import ctypes as C
import math

import numpy as np

from . import _capi as capi


class HostScene:
    """Host-side scene arrays (numpy) plus SimParams / step config; owns its memory."""

    def __init__(self, view, keep=None):
        n, b = view.numParticles, view.numBodies

        def arr(p, count, dtype):
            if count == 0 or not p:
                return np.zeros(0, dtype)
            buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(p)
            return np.frombuffer(buf, dtype=dtype, count=count).copy()

        self.pos = arr(view.pos, 4 * n, np.float32).reshape(n, 4)
        self.vel = arr(view.vel, 4 * n, np.float32).reshape(n, 4)
        self.rb = arr(view.rbIndices, n, np.int32)
        self.body_frame = arr(view.bodyFramePos, 4 * n, np.float32).reshape(n, 4)
        self.rb_pos = arr(view.rbPositions, 4 * b, np.float32).reshape(b, 4)
        self.rb_vel = arr(view.rbVelocities, 4 * b, np.float32).reshape(b, 4)
        self.rb_ang = arr(view.rbAngularVelocity, 4 * b, np.float32).reshape(b, 4)
        self.rb_quat = arr(view.rbQuaternion, 4 * b, np.float32).reshape(b, 4)
        self.rb_inertia = arr(view.rbInertia, 9 * b, np.float32).reshape(b, 9)
        self.rb_mass = arr(view.rbMass, b, np.float32)
        self.ppo = arr(view.particlesPerObject, b, np.int32)
        self.params = capi.ArSimParams.from_buffer_copy(view.params)
        self.cfg = capi.ArSimStepConfig.from_buffer_copy(view.cfg)

    @property
    def n(self):
        return int(self.pos.shape[0])

    @property
    def b(self):
        return int(self.rb_pos.shape[0])

    @property
    def pstart(self):
        return (np.cumsum(self.ppo) - self.ppo).astype(np.int32)

    def view(self):
        """An ArSimSceneHost pointing at this object's arrays (valid while self lives)."""
        v = capi.ArSimSceneHost()
        v.numParticles, v.numBodies = self.n, self.b
        v.pos, v.vel = capi.ptr(self.pos), capi.ptr(self.vel)
        v.rbIndices, v.bodyFramePos = capi.ptr(self.rb), capi.ptr(self.body_frame)
        v.rbPositions, v.rbVelocities = capi.ptr(self.rb_pos), capi.ptr(self.rb_vel)
        v.rbAngularVelocity, v.rbQuaternion = capi.ptr(self.rb_ang), capi.ptr(self.rb_quat)
        v.rbInertia, v.rbMass = capi.ptr(self.rb_inertia), capi.ptr(self.rb_mass)
        v.particlesPerObject = capi.ptr(self.ppo)
        v.params, v.cfg = self.params, self.cfg
        return v

    ARRAYS = ("pos", "vel", "rb", "body_frame", "rb_pos", "rb_vel", "rb_ang", "rb_quat", "rb_inertia", "rb_mass", "ppo")

    def to_arrays(self):
        """Plain arrays (and the raw SimParams / step-config bytes) for storing a scene."""
        out = {k: getattr(self, k) for k in self.ARRAYS}
        out["params_bytes"] = np.frombuffer(bytes(self.params), np.uint8).copy()
        out["cfg_bytes"] = np.frombuffer(bytes(self.cfg), np.uint8).copy()
        return out

    @classmethod
    def from_arrays(cls, d):
        s = object.__new__(cls)
        for k in cls.ARRAYS:
            setattr(s, k, np.ascontiguousarray(d[k]))
        s.params = capi.ArSimParams.from_buffer_copy(bytes(np.asarray(d["params_bytes"], np.uint8)))
        s.cfg = capi.ArSimStepConfig.from_buffer_copy(bytes(np.asarray(d["cfg_bytes"], np.uint8)))
        return s

    def copy(self):
        other = object.__new__(HostScene)
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(other, k, v.copy())
            elif isinstance(v, (capi.ArSimParams, capi.ArSimStepConfig)):
                setattr(other, k, type(v).from_buffer_copy(v))
            else:
                setattr(other, k, v)
        return other


def _from_owned(owned):
    view = capi.ArSimSceneHost()
    capi.check(capi.lib.arsim_scene_view(owned, C.byref(view)), "scene_view")
    s = HostScene(view)
    capi.lib.arsim_scene_free(owned)
    return s


def generate(spec):
    owned = C.c_void_p()
    capi.check(capi.lib.arsim_generate_scene(C.byref(spec), C.byref(owned)), "generate_scene")
    return _from_owned(owned)


def mix_scene(n, seed, method=capi.METHOD_GRID):
    owned = C.c_void_p()
    capi.check(capi.lib.arsim_generate_mix_scene(int(n), C.c_uint64(seed), int(method), C.byref(owned)),
               "generate_mix_scene")
    return _from_owned(owned)


def _spec(seed=0, cubes=0, grains=0, box=(64.0, 64.0, 64.0), grid_log2=(6, 6, 6), method=capi.METHOD_GRID,
          body_sigma=0.5, grain_sigma=0.5, ellipsoids=0, semi=(1.0, 4.0), drop=0.0):
    g = capi.ArSimGenSpec()
    g.seed = seed
    g.numCubes, g.cubeEdge, g.numEllipsoids = cubes, 4, ellipsoids
    g.ellipsoidMinSemi, g.ellipsoidMaxSemi = semi
    g.numGrains = grains
    for a in range(3):
        g.box[a] = box[a]
        g.gridLog2[a] = grid_log2[a]
    g.radius = 0.5
    g.bodyVelSigma, g.grainVelSigma = body_sigma, grain_sigma
    g.dropHeightMax = drop
    g.method = method
    g.dt, g.spring, g.damping, g.shear, g.gravityY = 1e-4, 5e4, 50.0, 10.0, -9.81
    return g


def config0(method=capi.METHOD_GRID, seed=0):
    """256 cubes of 4x4x4 (16,384 particles), r = 0.5, [0,64]^3, grid 64^3, v ~ N(0, 0.5)."""
    return generate(_spec(seed=seed, cubes=256, method=method))


def config1(method=capi.METHOD_GRID, seed=0):
    """8,192 cubes (524,288 particles) + 524,288 grains in [0,512]^3, grid 256^3 hashed, v ~ N(0,1)."""
    return generate(_spec(seed=seed, cubes=8192, grains=524288, box=(512.0, 512.0, 512.0), grid_log2=(8, 8, 8),
                          method=method, body_sigma=1.0, grain_sigma=1.0))


def config2(method=capi.METHOD_GRID, seed=0, scale=1.0, gap=0.98, surface=0.9):
    """Dense settled pile: 16,384 ellipsoids (semi-axes on [2r, 8r], voxelised at spacing r) and
    grains on a compressed FCC lattice, 4,194,304 particles in a container 160 wide, grid 256^3.
    scale < 1 shrinks the counts and the container's floor area (same pile height) for tests."""
    sp = capi.ArSimPileSpec()
    sp.seed = seed
    sp.numEllipsoids = max(1, int(round(16384 * scale)))
    sp.targetParticles = int(round(4194304 * scale))
    sp.edge = 160.0 * math.sqrt(scale)
    sp.radius = 0.5
    sp.latticeGap = gap
    sp.surfaceGap = surface
    sp.gridLog2 = 8
    sp.method = method
    owned = C.c_void_p()
    capi.check(capi.lib.arsim_generate_packed_pile(C.byref(sp), C.byref(owned)), "generate_packed_pile")
    return _from_owned(owned)


def batch_mix(num_scenes=512, lo=1024, hi=65536, seed=0, method=capi.METHOD_GRID):
    """Scenes with particle counts log-uniform in [lo, hi], each from the small-scene generator."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(lo), math.log(hi), size=num_scenes)).round().astype(int)
    return [mix_scene(int(n), seed * 100003 + k, method) for k, n in enumerate(counts)]


def stacked_cubes(nx=32, nz=32, layers=2, seed=0, k=4, pitch=None, floor_pitch=0.5):
    """Sequential-impulse workload: `layers` layers of nx x nz cubes of k^3 particles (spacing 2r), each
    resting on the static floor or on the cube below (0.1 r penetration), with a random yaw and small
    random velocities; returns (scene, floor points, floor normals).  Bodies only, ARSIM_CONTACT_SIS set."""
    rng = np.random.default_rng(seed)
    base = mix_scene(1024, seed)
    r = float(base.params.particleRadius)
    g = (np.arange(k, dtype=np.float32) - (k - 1) / 2) * (2 * r)
    offs = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    m = offs.shape[0]
    half = (k - 1) * r + r                        # half extent of a cube, particle radius included
    pitch = pitch or 2 * half + 0.4 * r
    floor_y = 1.0
    cms = []
    for layer in range(layers):
        y = floor_y + r + half - 0.1 * r + layer * (2 * half - 0.1 * r)
        for i in range(nx):
            for j in range(nz):
                cms.append((2 * half + i * pitch, y, 2 * half + j * pitch))
    cms = np.asarray(cms, np.float32)
    nb = len(cms)
    yaw = rng.uniform(-0.05, 0.05, nb)
    q = np.zeros((nb, 4), np.float32)
    q[:, 1], q[:, 3] = np.sin(yaw / 2), np.cos(yaw / 2)
    inertia = (offs.astype(np.float64) ** 2).sum() * np.eye(3) - offs.T.astype(np.float64) @ offs
    inv = np.linalg.inv(inertia).astype(np.float32)
    sc = base.copy()
    sc.rb = np.repeat(np.arange(nb, dtype=np.int32), m)
    sc.body_frame = np.zeros((nb * m, 4), np.float32)
    sc.body_frame[:, :3] = np.tile(offs, (nb, 1))
    sc.pos = np.zeros((nb * m, 4), np.float32)
    c, sn = np.cos(yaw), np.sin(yaw)
    for b in range(nb):
        R = np.array([[c[b], 0, sn[b]], [0, 1, 0], [-sn[b], 0, c[b]]], np.float32)
        sc.pos[b * m:(b + 1) * m, :3] = cms[b] + offs @ R.T
    sc.vel = np.zeros_like(sc.pos)
    sc.rb_pos = np.zeros((nb, 4), np.float32)
    sc.rb_pos[:, :3] = cms
    sc.rb_vel = np.zeros((nb, 4), np.float32)
    sc.rb_vel[:, :3] = rng.normal(scale=0.05, size=(nb, 3))
    sc.rb_ang = np.zeros((nb, 4), np.float32)
    sc.rb_quat = q
    sc.rb_inertia = np.tile(inv.T.reshape(1, 9), (nb, 1))
    sc.rb_mass = np.full(nb, float(m), np.float32)
    sc.ppo = np.full(nb, m, np.int32)
    ext = 4 * half + max(nx, nz) * pitch
    size = 1 << max(6, int(math.ceil(math.log2(ext))))
    p = sc.params
    p.gridSize.x = p.gridSize.y = p.gridSize.z = size
    p.numCells = size ** 3
    p.numBodies = nb
    p.ARrestitution, p.RBrestitution, p.ARfriction, p.RBfriction = 0.5, 0.5, 0.3, 0.1
    sc.cfg.boxMax[:] = [ext, ext, ext]
    sc.cfg.flags = capi.CONTACT_SIS
    sc.cfg.dt = 1e-3
    xs = np.arange(0.0, ext, floor_pitch, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    fp = np.zeros((X.size, 4), np.float32)
    fp[:, 0], fp[:, 1], fp[:, 2] = X.ravel(), floor_y, Z.ravel()
    fn = np.zeros_like(fp)
    fn[:, 1] = 1.0
    return sc, fp, fn


class Template:
    """A particle template loaded as the reference's initObject does (arsim_template_load)."""

    def __init__(self, data_dir, model, scale, radius):
        self.h = C.c_void_p()
        capi.check(capi.lib.arsim_template_load(str(data_dir).encode(), model.encode(), float(scale), float(radius),
                                                C.byref(self.h)), "template_load")
        n = C.c_int(0)
        capi.check(capi.lib.arsim_template_info(self.h, C.byref(n), None, None, None, None), "template_info")
        self.n = n.value
        self.rel = np.zeros((self.n, 4), np.float32)
        self.inv_inertia = np.zeros(9, np.float32)
        rad, mass = C.c_float(0), C.c_float(0)
        capi.check(capi.lib.arsim_template_info(self.h, None, capi.ptr(self.rel), capi.ptr(self.inv_inertia),
                                                C.byref(rad), C.byref(mass)), "template_info")
        self.radius, self.mass = rad.value, mass.value

    def close(self):
        if self.h:
            capi.lib.arsim_template_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def template_suffix(scale):
    return capi.lib.arsim_template_suffix(float(scale)).decode()


class SceneBuilder:
    """An empty scene filled with template bodies and free grains (arsim_scene_create / add_*)."""

    def __init__(self, params, cfg):
        self.h = C.c_void_p()
        capi.check(capi.lib.arsim_scene_create(C.byref(params), C.byref(cfg), C.byref(self.h)), "scene_create")

    def add_body(self, template, pos, vel=(0.0, 0.0, 0.0), ang=(0.0, 0.0, 0.0)):
        f3 = lambda v: np.ascontiguousarray(v, np.float32)
        p, v, w = f3(pos), f3(vel), f3(ang)
        capi.check(capi.lib.arsim_scene_add_body(self.h, template.h, capi.ptr(p), capi.ptr(v), capi.ptr(w)),
                   "scene_add_body")

    def add_grains(self, pos4, vel4):
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        vel4 = np.ascontiguousarray(vel4, np.float32).reshape(-1, 4)
        assert pos4.shape == vel4.shape
        capi.check(capi.lib.arsim_scene_add_grains(self.h, pos4.shape[0], capi.ptr(pos4), capi.ptr(vel4)),
                   "scene_add_grains")

    def build(self):
        """A HostScene copy of the current contents (the builder stays usable)."""
        view = capi.ArSimSceneHost()
        capi.check(capi.lib.arsim_scene_view(self.h, C.byref(view)), "scene_view")
        return HostScene(view)

    def close(self):
        if self.h:
            capi.lib.arsim_scene_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def template_pile(data_dir, models=("bunny", "teapot", "banana", "cube"), scale=1.0, per_model=4, grains=0,
                  method=capi.METHOD_GRID, seed=0, spacing=0.6):
    """Template bodies of the reference's models (radius 1/64, the templates' own scale) on a jittered
    lattice with small random velocities, plus free grains above them, in a walled box."""
    r = 1.0 / 64.0
    temps = [Template(data_dir, m, scale, r) for m in models]
    ext = max(float(np.max(np.abs(t.rel[:, :3]))) for t in temps) + r
    pitch = 2.0 * ext * spacing
    nb = len(temps) * per_model
    side = int(math.ceil(nb ** (1.0 / 3.0)))
    box = side * pitch + 4 * r
    g = _spec(seed=seed, cubes=1, box=(box, box * 2.0, box), grid_log2=(6, 7, 6), method=method)
    g.radius = r
    g.spring, g.damping, g.shear, g.dt = 2000.0, 2.0, 0.5, 2e-5
    base = generate(g)
    params, cfg = base.params, base.cfg
    params.particleRadius = r
    params.cellSize.x = params.cellSize.y = params.cellSize.z = 2.0 * r
    b = SceneBuilder(params, cfg)
    rng = np.random.default_rng(seed)
    k = 0
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                if k >= nb:
                    break
                c = (np.array([ix, iy, iz]) + 0.5) * pitch + 2 * r + rng.uniform(-0.05, 0.05, 3) * pitch
                b.add_body(temps[k % len(temps)], c, rng.normal(scale=0.2, size=3), rng.normal(scale=2.0, size=3))
                k += 1
    if grains:
        gp = np.zeros((grains, 4), np.float32)
        gp[:, 0] = rng.uniform(2 * r, box - 2 * r, grains)
        gp[:, 1] = side * pitch + 4 * r + np.arange(grains) * (2.5 * r) / max(1, int(grains ** 0.5))
        gp[:, 2] = rng.uniform(2 * r, box - 2 * r, grains)
        gv = np.zeros((grains, 4), np.float32)
        gv[:, :3] = rng.normal(scale=0.1, size=(grains, 3))
        b.add_grains(gp, gv)
    sc = b.build()
    b.close()
    for t in temps:
        t.close()
    return sc
