"""Python handle on the C++ step engine (arsim_engine_*), the scene-level ParticleSystem::update API."""
import ctypes as C

import numpy as np

from . import _capi as capi


class Engine:
    """S >= 1 scenes resident on the GPU; step() advances all of them (one launch per stage)."""

    def __init__(self, scenes, dem_block=128, sort_ipt=16, sort_radix_bits=8, use_graph=True, device=0,
                 fuse_reduce=False, bucket_split=0):
        if not isinstance(scenes, (list, tuple)):
            scenes = [scenes]
        self.scenes = list(scenes)
        self._views = (capi.ArSimSceneHost * len(scenes))(*[s.view() for s in scenes])
        self.tuning = capi.ArSimTuning(dem_block, sort_ipt, sort_radix_bits, 1 if use_graph else 0,
                                       1 if fuse_reduce else 0, bucket_split)
        self.h = C.c_void_p()
        capi.check(capi.lib.arsim_engine_create(self._views, len(scenes), C.byref(self.tuning), device,
                                                C.byref(self.h)), "engine_create")

    def close(self):
        if self.h:
            capi.lib.arsim_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n=1):
        capi.check(capi.lib.arsim_engine_step(self.h, int(n)), "engine_step")

    def update(self, dt):
        """ParticleSystem::update(dt) (Particles/particleSystem.h:288): one step of length dt."""
        capi.check(capi.lib.arsim_engine_set_dt(self.h, float(dt)), "engine_set_dt")
        self.step(1)

    def sync(self):
        capi.check(capi.lib.arsim_engine_synchronize(self.h), "engine_synchronize")

    @property
    def stream(self):
        s = C.c_void_p()
        capi.check(capi.lib.arsim_engine_stream(self.h, C.byref(s)))
        return s

    @property
    def num_particles(self):
        return capi.lib.arsim_engine_num_particles(self.h, -1)

    @property
    def device_bytes(self):
        return capi.lib.arsim_engine_device_bytes(self.h)

    def download(self, scene=0):
        s = self.scenes[scene]
        n, b = s.n, s.b
        out = dict(pos=np.zeros((n, 4), np.float32), vel=np.zeros((n, 4), np.float32),
                   rel=np.zeros((n, 4), np.float32), force=np.zeros((n, 4), np.float32),
                   torque=np.zeros((n, 4), np.float32), rb_pos=np.zeros((b, 4), np.float32),
                   rb_vel=np.zeros((b, 4), np.float32), rb_ang=np.zeros((b, 4), np.float32),
                   rb_quat=np.zeros((b, 4), np.float32), rb_iw=np.zeros((b, 9), np.float32))
        P = capi.ptr
        capi.check(capi.lib.arsim_engine_download(self.h, scene, P(out["pos"]), P(out["vel"]), P(out["rel"]),
                                                  P(out["force"]), P(out["torque"]), P(out["rb_pos"]),
                                                  P(out["rb_vel"]), P(out["rb_ang"]), P(out["rb_quat"]),
                                                  P(out["rb_iw"])), "engine_download")
        return out

    def checksum(self, scene=0):
        """FNV-1a over the scene's state (arsim_engine_state_checksum): pos, vel, body pos, vel,
        angular velocity, orientation."""
        h = C.c_uint64(0)
        capi.check(capi.lib.arsim_engine_state_checksum(self.h, int(scene), C.byref(h)), "engine_state_checksum")
        return h.value

    def snapshot(self, scene=0):
        """A HostScene that rebuilds scene `scene` in its current state (arsim_engine_snapshot)."""
        from .scenes import _from_owned
        owned = C.c_void_p()
        capi.check(capi.lib.arsim_engine_snapshot(self.h, int(scene), C.byref(owned)), "engine_snapshot")
        return _from_owned(owned)

    def rebuild(self, source, keep=(), added=()):
        """Holds scenes `keep` of engine `source` (state carried over on the device) followed by the
        host scenes `added`; this engine's allocations are reused where they fit."""
        keep = np.ascontiguousarray(list(keep), np.int32)
        added = list(added)
        views = (capi.ArSimSceneHost * max(1, len(added)))(*[s.view() for s in added])
        capi.check(capi.lib.arsim_engine_rebuild(self.h, source.h if source is not None else None,
                                                 capi.ptr(keep) if keep.size else None, int(keep.size),
                                                 views if added else None, len(added)), "engine_rebuild")
        self._views = views
        self.scenes = [source.scenes[int(k)] for k in keep] + added

    def set_static(self, pos4, norm4):
        """Static collider points and normals (float4 each) for a one-scene engine; None clears them."""
        if pos4 is None:
            capi.check(capi.lib.arsim_engine_set_static(self.h, None, None, 0), "engine_set_static")
            return
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        norm4 = np.ascontiguousarray(norm4, np.float32).reshape(-1, 4)
        assert pos4.shape == norm4.shape
        capi.check(capi.lib.arsim_engine_set_static(self.h, capi.ptr(pos4), capi.ptr(norm4), pos4.shape[0]),
                   "engine_set_static")

    def render_state(self, scene=0):
        """Per-body model matrices (B, 4, 4) in column-major storage order and cumulative quaternions."""
        b = self.scenes[scene].b
        m = np.zeros((b, 16), np.float32)
        q = np.zeros((b, 4), np.float32)
        capi.check(capi.lib.arsim_engine_render_state(self.h, int(scene), capi.ptr(m), capi.ptr(q)),
                   "engine_render_state")
        return m, q

    def profile(self):
        ms = (C.c_float * 16)()
        names = (C.c_char_p * 16)()
        k = capi.lib.arsim_engine_profile_step(self.h, ms, names, 16)
        if k < 0 or k > 16:
            capi.check(k, "engine_profile_step")
        return [(names[i].decode(), ms[i]) for i in range(k)]

    def contact_pairs(self, max_per=64):
        n = self.num_particles
        pairs = np.zeros((n, max_per), np.int32)
        counts = np.zeros(n, np.int32)
        capi.check(capi.lib.arsim_engine_contact_pairs(self.h, capi.ptr(pairs), capi.ptr(counts), max_per),
                   "engine_contact_pairs")
        return pairs, counts

    def check_finite(self):
        flag = C.c_int(0)
        rc = capi.lib.arsim_engine_check_finite(self.h, C.byref(flag))
        return rc == capi.ARSIM_OK and flag.value == 0

    def autotune(self, steps_per_trial=5):
        best = capi.ArSimTuning()
        ms = C.c_float(0)
        capi.check(capi.lib.arsim_engine_autotune(self.h, steps_per_trial, C.byref(best), C.byref(ms)),
                   "engine_autotune")
        self.tuning = best
        return (best.demBlock, best.sortItemsPerThread, best.sortRadixBits, best.bucketSplit), ms.value


def contact_set(pairs, counts, offset=0):
    """Set of unordered pairs (i < j) from an (N, max_per) contact list."""
    out = set()
    for i in np.nonzero(counts)[0]:
        for j in pairs[i, :counts[i]]:
            a, b = int(i) + offset, int(j) + offset
            out.add((a, b) if a < b else (b, a))
    return out
