"""TEST INFRASTRUCTURE ONLY: numpy-facing loader of the host oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package.
It is the checker of the gfx950 kernels, never part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "-B", "liboracle.so"], check=True, capture_output=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()
VP = C.c_void_p


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data_as(VP)


for name, res, args in [
    ("orc_calc_hash", None, [VP, C.c_int, VP, VP, VP]),
    ("orc_morton30", None, [VP, C.c_int, VP, VP, VP]),
    ("orc_morton63", None, [VP, C.c_int, VP, VP, VP]),
    ("orc_morton30_unit", None, [VP, C.c_int, VP]),
    ("orc_morton63_unit", None, [VP, C.c_int, VP]),
    ("orc_expand10", None, [VP, C.c_int, VP]),
    ("orc_expand21", None, [VP, C.c_int, VP]),
    ("orc_sort_pairs_u32", None, [VP, VP, C.c_int, C.c_int, C.c_int, VP, VP]),
    ("orc_sort_pairs_u64", None, [VP, VP, C.c_int, C.c_int, C.c_int, VP, VP]),
    ("orc_cell_bounds", None, [VP, C.c_int, C.c_uint, VP, VP]),
    ("orc_karras_u32", None, [VP, C.c_int, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_karras_u64", None, [VP, C.c_int, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_refit", None, [VP, VP, C.c_float, C.c_int, VP, VP, VP, VP]),
    ("orc_contacts_brute", C.c_long, [VP, VP, C.c_int, C.c_float, VP, C.c_long]),
    ("orc_step_f32", C.c_int, [VP, VP, VP, C.c_int]),
    ("orc_step_f64", C.c_int, [VP, VP, VP, C.c_int]),
    ("orc_integrate_bodies_f32", C.c_int, [VP, VP, VP]),
    ("orc_integrate_bodies_f64", C.c_int, [VP, VP, VP]),
    ("orc_pose_grains_f32", C.c_int, [VP, VP, VP]),
    ("orc_pose_grains_f64", C.c_int, [VP, VP, VP]),
    ("orc_reduce_apply_f32", C.c_int, [VP, VP, VP]),
    ("orc_reduce_apply_f64", C.c_int, [VP, VP, VP]),
    ("orc_dem_grid_f32", C.c_int, [VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_dem_grid_f64", C.c_int, [VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_dem_lbvh_f32", C.c_int, [VP, VP, VP, C.c_int, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_dem_lbvh_f64", C.c_int, [VP, VP, VP, C.c_int, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_pair_force_f32", C.c_int, [VP, VP, VP, VP, VP, VP, VP]),
    ("orc_wall_force_f64", None, [VP, VP, VP, VP, VP]),
    ("orc_sincos", None, [VP, C.c_int, VP, VP]),
    ("orc_range_image", None, [VP, C.c_int, C.c_int, VP, C.c_float, VP, VP, VP, VP]),
    ("orc_dem_static_f32", C.c_int, [VP, VP, VP, VP, VP, VP, VP, C.c_int]),
    ("orc_dem_static_f64", C.c_int, [VP, VP, VP, VP, VP, VP, VP, C.c_int]),
    ("orc_sis_find_deepest_f32", None, [C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_sis_find_static_f32", None, [C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    ("orc_sis_gather_f32", C.c_long, [VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, C.c_long]),
    ("orc_sis_solve_f32", C.c_int, [VP, VP, VP, VP, VP, VP, VP, C.c_long, C.c_int]),
    ("orc_sis_solve_f64", C.c_int, [VP, VP, VP, VP, VP, VP, VP, C.c_long, C.c_int]),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


# ----------------------------------------------------------------------------- stage wrappers
def range_image(depth, cam, radius):
    """Static scene points / normals (float4 per pixel) from a uint16 depth image (H, W) and a
    column-major camera matrix; pixels the reference leaves unwritten stay NaN here."""
    depth = np.ascontiguousarray(depth, np.uint16)
    H, W = depth.shape
    pos = np.full((H * W, 4), np.nan, np.float32)
    norm = np.full((H * W, 4), np.nan, np.float32)
    radii = np.full(H * W, np.nan, np.float32)
    vao = np.full((H * W, 5), np.nan, np.float32)
    lib.orc_range_image(_p(depth), W, H, _p(np.ascontiguousarray(cam, np.float32).reshape(16)), float(radius),
                        _p(pos), _p(norm), _p(radii), _p(vao))
    return pos, norm, radii, vao


def calc_hash(pos, params):
    n = pos.shape[0]
    h = np.zeros(n, np.uint32)
    i = np.zeros(n, np.uint32)
    lib.orc_calc_hash(_p(np.ascontiguousarray(pos, np.float32)), n, C.byref(params), _p(h), _p(i))
    return h, i


def morton30(pos, bmin, bmax):
    n = pos.shape[0]
    out = np.zeros(n, np.uint32)
    lib.orc_morton30(_p(np.ascontiguousarray(pos, np.float32)), n, _p(np.asarray(bmin, np.float32)),
                     _p(np.asarray(bmax, np.float32)), _p(out))
    return out


def morton63(pos, bmin, bmax):
    n = pos.shape[0]
    out = np.zeros(n, np.uint64)
    lib.orc_morton63(_p(np.ascontiguousarray(pos, np.float32)), n, _p(np.asarray(bmin, np.float32)),
                     _p(np.asarray(bmax, np.float32)), _p(out))
    return out


def morton30_unit(xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros(xyz.shape[0], np.uint32)
    lib.orc_morton30_unit(_p(xyz), xyz.shape[0], _p(out))
    return out


def morton63_unit(xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros(xyz.shape[0], np.uint64)
    lib.orc_morton63_unit(_p(xyz), xyz.shape[0], _p(out))
    return out


def expand10(v):
    v = np.ascontiguousarray(v, np.uint32)
    out = np.zeros_like(v)
    lib.orc_expand10(_p(v), v.size, _p(out))
    return out


def expand21(v):
    v = np.ascontiguousarray(v, np.uint64)
    out = np.zeros_like(v)
    lib.orc_expand21(_p(v), v.size, _p(out))
    return out


def sort_pairs(keys, vals, begin_bit=0, end_bit=None):
    keys = np.ascontiguousarray(keys)
    vals = np.ascontiguousarray(vals, np.uint32)
    n = keys.size
    ok, ov = np.zeros_like(keys), np.zeros_like(vals)
    if keys.dtype == np.uint32:
        lib.orc_sort_pairs_u32(_p(keys), _p(vals), n, begin_bit, 32 if end_bit is None else end_bit, _p(ok), _p(ov))
    else:
        lib.orc_sort_pairs_u64(_p(keys), _p(vals), n, begin_bit, 64 if end_bit is None else end_bit, _p(ok), _p(ov))
    return ok, ov


def cell_bounds(sorted_hash, num_cells):
    cs = np.zeros(num_cells, np.uint32)
    ce = np.zeros(num_cells, np.uint32)
    sorted_hash = np.ascontiguousarray(sorted_hash, np.uint32)
    lib.orc_cell_bounds(_p(sorted_hash), sorted_hash.size, num_cells, _p(cs), _p(ce))
    return cs, ce


def karras(codes, base=None):
    codes = np.ascontiguousarray(codes)
    n = codes.size
    m = max(2 * n, 2)
    left = np.full(m, -1, np.int32)
    right = np.full(m, -1, np.int32)
    parent = np.full(m, -1, np.int32)
    minr = np.zeros(m, np.int32)
    maxr = np.zeros(m, np.int32)
    leaf = np.zeros(m, np.uint8)
    b = None if base is None else np.ascontiguousarray(base, np.int32)
    f = lib.orc_karras_u32 if codes.dtype == np.uint32 else lib.orc_karras_u64
    f(_p(codes), n, _p(b), _p(left), _p(right), _p(parent), _p(minr), _p(maxr), _p(leaf))
    return dict(left=left, right=right, parent=parent, min=minr, max=maxr, leaf=leaf)


def refit(sorted_idx, pos, r, left, right):
    n = sorted_idx.size
    cms = np.zeros((n, 4), np.float32)
    bounds = np.zeros((max(2 * n, 2), 8), np.float32)
    lib.orc_refit(_p(np.ascontiguousarray(sorted_idx, np.int32)), _p(np.ascontiguousarray(pos, np.float32)), r, n,
                  _p(left), _p(right), _p(cms), _p(bounds))
    return cms, bounds


def contacts_brute(pos, rb, r, cap=None):
    n = pos.shape[0]
    cap = cap or max(16 * n, 1024)
    pairs = np.zeros((cap, 2), np.int32)
    cnt = lib.orc_contacts_brute(_p(np.ascontiguousarray(pos, np.float32)), _p(np.ascontiguousarray(rb, np.int32)),
                                 n, r, _p(pairs), cap)
    if cnt < 0:
        raise RuntimeError("contact capacity exceeded")
    return set(map(tuple, pairs[:cnt].tolist()))


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    lib.orc_sincos(_p(x), x.size, _p(s), _p(c))
    return s, c


def pair_force_f32(pa, pb, va, vb, params, cfg):
    out = np.zeros(3, np.float32)
    a = [np.ascontiguousarray(v, np.float32) for v in (pa, pb, va, vb)]
    hit = lib.orc_pair_force_f32(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), C.byref(params), C.byref(cfg), _p(out))
    return bool(hit), out


def wall_force_f64(p, v, params, cfg):
    out = np.zeros(3, np.float64)
    lib.orc_wall_force_f64(_p(np.ascontiguousarray(p, np.float64)), _p(np.ascontiguousarray(v, np.float64)),
                           C.byref(params), C.byref(cfg), _p(out))
    return out


def sis_find_deepest(sorted_idx, cell_start, cell_end, sorted_pos, rb, params):
    """Deepest contact per particle against other bodies' particles (reference layout cell tables);
    returns (dist, partner body, partner particle) per original index (dist 0 = no contact)."""
    n = int(sorted_idx.size)
    dist = np.zeros(n, np.float32)
    body = np.full(n, -1, np.int32)
    part = np.full(n, -1, np.int32)
    lib.orc_sis_find_deepest_f32(n, C.byref(params), _p(np.ascontiguousarray(sorted_idx, np.uint32)),
                                 _p(np.ascontiguousarray(cell_start, np.uint32)),
                                 _p(np.ascontiguousarray(cell_end, np.uint32)),
                                 _p(np.ascontiguousarray(sorted_pos, np.float32)),
                                 _p(np.ascontiguousarray(rb, np.int32)), _p(dist), _p(body), _p(part))
    return dist, body, part


def sis_find_static(sorted_idx, sorted_pos, static_sorted_idx, cell_start, cell_end, static_sorted_pos, params):
    """Deepest contact per particle against the static points; returns (dist, static point index)."""
    n = int(sorted_idx.size)
    dist = np.zeros(n, np.float32)
    part = np.full(n, -1, np.int32)
    lib.orc_sis_find_static_f32(n, C.byref(params), _p(np.ascontiguousarray(sorted_idx, np.uint32)),
                                _p(np.ascontiguousarray(sorted_pos, np.float32)),
                                _p(np.ascontiguousarray(static_sorted_idx, np.uint32)),
                                _p(np.ascontiguousarray(cell_start, np.uint32)),
                                _p(np.ascontiguousarray(cell_end, np.uint32)),
                                _p(np.ascontiguousarray(static_sorted_pos, np.float32)), _p(dist), _p(part))
    return dist, part


# ----------------------------------------------------------------------------- full state
class _StateF(C.Structure):
    _fields_ = [("N", C.c_int), ("B", C.c_int)] + [(k, VP) for k in (
        "pos", "vel", "relPos", "bodyFrame", "pForce", "pTorque", "rb", "rbPos", "rbVel", "rbAng", "qinc", "qcum",
        "Ib", "Iw", "mass", "pstart", "pcount", "model", "staticPos", "staticNorm")] + [("numStatic", C.c_int)]


class State:
    """Oracle state in float32 (mirror) or float64; arrays in the reference layout."""

    def __init__(self, scene, dtype=np.float32):
        self.dtype = dtype
        t = dtype
        self.N, self.B = scene.n, scene.b
        self.pos = scene.pos.astype(t).copy()
        self.vel = scene.vel.astype(t).copy()
        self.rel = np.zeros((self.N, 4), t)
        self.body_frame = scene.body_frame.astype(t).copy()
        self.force = np.zeros((self.N, 4), t)
        self.torque = np.zeros((self.N, 4), t)
        self.rb = scene.rb.astype(np.int32).copy()
        self.rb_pos = scene.rb_pos.astype(t).copy()
        self.rb_vel = scene.rb_vel.astype(t).copy()
        self.rb_ang = scene.rb_ang.astype(t).copy()
        self.qinc = np.tile(np.array([0, 0, 0, 1], t), (max(self.B, 1), 1))[:self.B].copy()
        self.qcum = scene.rb_quat.astype(t).copy()
        self.Ib = scene.rb_inertia.astype(t).copy()
        self.Iw = scene.rb_inertia.astype(t).copy()
        self.mass = scene.rb_mass.astype(t).copy()
        self.pstart = scene.pstart.copy()
        self.pcount = scene.ppo.astype(np.int32).copy()
        self.model = np.zeros((self.B, 16), t)     # written by each body integration
        self.static_pos = self.static_norm = None  # optional static scene points (float32, float4 each)
        self.params = scene.params
        self.cfg = scene.cfg

    def _struct(self):
        s = _StateF()
        s.N, s.B = self.N, self.B
        for k, a in (("pos", self.pos), ("vel", self.vel), ("relPos", self.rel), ("bodyFrame", self.body_frame),
                     ("pForce", self.force), ("pTorque", self.torque), ("rb", self.rb), ("rbPos", self.rb_pos),
                     ("rbVel", self.rb_vel), ("rbAng", self.rb_ang), ("qinc", self.qinc), ("qcum", self.qcum),
                     ("Ib", self.Ib), ("Iw", self.Iw), ("mass", self.mass), ("pstart", self.pstart),
                     ("pcount", self.pcount), ("model", self.model), ("staticPos", self.static_pos),
                     ("staticNorm", self.static_norm)):
            setattr(s, k, _p(a))
        s.numStatic = 0 if self.static_pos is None else int(self.static_pos.shape[0])
        return s

    def set_static(self, pos4, norm4):
        """Static scene points and normals; each step adds their contact forces to the body particles
        after the DEM (rebuilding their grid index every step, as the reference does)."""
        if pos4 is None or len(pos4) == 0:
            self.static_pos = self.static_norm = None
            return
        self.static_pos = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        self.static_norm = np.ascontiguousarray(norm4, np.float32).reshape(-1, 4)
        assert self.static_pos.shape == self.static_norm.shape

    def dem_static(self, cell_start, cell_end, sorted_static_pos, sorted_static_norm, ref=False):
        self._call("orc_dem_static", _p(np.ascontiguousarray(cell_start, np.uint32)),
                   _p(np.ascontiguousarray(cell_end, np.uint32)),
                   _p(np.ascontiguousarray(sorted_static_pos, np.float32)),
                   _p(np.ascontiguousarray(sorted_static_norm, np.float32)), 1 if ref else 0)

    def _call(self, base, *extra):
        s = self._struct()
        f = getattr(lib, base + ("_f32" if self.dtype == np.float32 else "_f64"))
        return f(C.byref(s), C.byref(self.params), C.byref(self.cfg), *extra)

    def step(self, n=1):
        s = self._struct()
        f = lib.orc_step_f32 if self.dtype == np.float32 else lib.orc_step_f64
        rc = f(C.byref(s), C.byref(self.params), C.byref(self.cfg), n)
        assert rc == 0, "oracle step refused the scene (the sequential-impulse model takes bodies only, grid method)"

    def integrate_bodies(self):
        self._call("orc_integrate_bodies")

    def sis_gather(self, dist_static, part_static, static_norm, dist, body, part):
        """Contact list of the sequential-impulse model (static contacts first); float32 state only."""
        assert self.dtype == np.float32
        cap = 2 * self.N + 1
        n4 = np.zeros((cap, 4), np.float32)
        pt4 = np.zeros((cap, 4), np.float32)
        b12 = np.zeros((cap, 2), np.int32)
        bias = np.zeros(cap, np.float32)
        s = self._struct()
        arr = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
        cnt = lib.orc_sis_gather_f32(C.byref(s), C.byref(self.params), _p(arr(dist_static, np.float32)),
                                     _p(arr(part_static, np.int32)), _p(arr(static_norm, np.float32)),
                                     _p(arr(dist, np.float32)), _p(arr(body, np.int32)), _p(arr(part, np.int32)),
                                     _p(n4), _p(pt4), _p(b12), _p(bias), cap)
        assert cnt >= 0
        return {"normal": n4[:cnt].copy(), "point": pt4[:cnt].copy(), "bodies": b12[:cnt].copy(),
                "bias": bias[:cnt].copy(), "impulse": np.zeros((cnt, 4), np.float32)}

    def sis_solve(self, contacts, iterations=18):
        """`iterations` sequential sweeps over the contacts; updates body velocities and
        contacts["impulse"] (normal, friction, friction 2 accumulators)."""
        s = self._struct()
        f = lib.orc_sis_solve_f32 if self.dtype == np.float32 else lib.orc_sis_solve_f64
        acc = contacts["impulse"]
        rc = f(C.byref(s), C.byref(self.params), _p(np.ascontiguousarray(contacts["normal"], np.float32)),
               _p(np.ascontiguousarray(contacts["point"], np.float32)),
               _p(np.ascontiguousarray(contacts["bodies"], np.int32)),
               _p(np.ascontiguousarray(contacts["bias"], np.float32)), _p(acc), int(acc.shape[0]), int(iterations))
        assert rc == 0

    def pose_grains(self):
        self._call("orc_pose_grains")

    def reduce_apply(self):
        self._call("orc_reduce_apply")

    def dem_grid(self, sorted_idx, cell_start, cell_end, sorted_pos, sorted_vel, sorted_rb):
        t = self.dtype
        self._call("orc_dem_grid", _p(np.ascontiguousarray(sorted_idx, np.uint32)), _p(cell_start), _p(cell_end),
                   _p(np.ascontiguousarray(sorted_pos, t)), _p(np.ascontiguousarray(sorted_vel, t)),
                   _p(np.ascontiguousarray(sorted_rb, np.int32)))

    def dem_lbvh(self, sorted_idx, left, right, bounds, cms, sorted_vel, sorted_rb):
        t = self.dtype
        self._call("orc_dem_lbvh", int(sorted_idx.size), _p(np.ascontiguousarray(sorted_idx, np.int32)), _p(left),
                   _p(right), _p(np.ascontiguousarray(bounds, np.float32)), _p(np.ascontiguousarray(cms, t)),
                   _p(np.ascontiguousarray(sorted_vel, t)), _p(np.ascontiguousarray(sorted_rb, np.int32)))
