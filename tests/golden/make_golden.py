"""Writes the golden fixtures in this directory (run from the repo root: python tests/golden/make_golden.py).

Each fixture holds a small seeded scene (inputs) and the oracle's state after a few steps, in FP32
(the mirror the device must match bit for bit) and FP64 (the reference-semantics model).  The
scenes are squeezed so that bodies, grains and walls are in contact.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from diploma_thesis_amd import capi, scenes  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = 5
OUT_KEYS = ("pos", "vel", "rel", "rb_pos", "rb_vel", "rb_ang", "qcum", "Iw")


def squeezed(n, seed, method, factor=0.6):
    sc = scenes.mix_scene(n, seed, method)
    c = sc.rb_pos[:, :3].mean(0)
    sc.rb_pos[:, :3] = c + (sc.rb_pos[:, :3] - c) * np.float32(factor)
    g = sc.rb < 0
    sc.pos[g, :3] = c + (sc.pos[g, :3] - c) * np.float32(factor)
    sc.rb_ang[:, :3] = np.random.default_rng(seed).normal(scale=1.5, size=(sc.b, 3)).astype(np.float32)
    return sc


def make(name, sc):
    out = {f"in_{k}": v for k, v in sc.to_arrays().items()}
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        st = oracle.State(sc, dt)
        st.step(STEPS)
        for k in OUT_KEYS:
            out[f"{tag}_{k}"] = getattr(st, k)
    out["steps"] = np.array(STEPS)
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name, sc.n, "particles", sc.b, "bodies")


if __name__ == "__main__":
    make("grid_1536.npz", squeezed(1536, 11, capi.METHOD_GRID))
    make("lbvh30_1536.npz", squeezed(1536, 12, capi.METHOD_LBVH30))
    make("lbvh63_1024.npz", squeezed(1024, 13, capi.METHOD_LBVH63))
