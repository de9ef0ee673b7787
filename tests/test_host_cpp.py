"""A C++ program that uses the library only through include/arsim.h and the ParticleSystemLite
wrapper (include/particle_system_lite.hpp, the reference's update(dt) / getArray shape): it compiles
and links with g++ against libarsim.so on any host, and on the GPU its state after ParticleSystemLite
updates equals the oracle's bit for bit."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from diploma_thesis_amd import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.dirname(capi.LIB_PATH)


def build(tmp_path):
    exe = str(tmp_path / "lite_update")
    res = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                          os.path.join(ROOT, "tests", "host", "lite_update_main.cpp"), "-L", LIBDIR, "-larsim",
                          f"-Wl,-rpath,{LIBDIR}", "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cpp_caller_compiles_and_links(tmp_path):
    exe = build(tmp_path)
    assert os.path.exists(exe)
    # without arguments it prints its usage and touches no device
    res = subprocess.run([exe], capture_output=True, text=True)
    assert res.returncode == 2 and "usage" in res.stderr


@pytest.mark.gpu
def test_cpp_caller_update_matches_oracle(tmp_path):
    import oracle
    print("host C++: ParticleSystemLite::update on the small-scene mix, 5 steps", flush=True)
    exe = build(tmp_path)
    out = str(tmp_path / "state.bin")
    n, seed, steps = 6000, 17, 5
    res = subprocess.run([exe, out, str(n), str(seed), str(steps)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    sc = scenes.mix_scene(n, seed, capi.METHOD_GRID)
    raw = np.fromfile(out, np.float32)
    N, B = sc.n, sc.b
    assert raw.size == 8 * N + 16 * B
    pos, vel = raw[:4 * N].reshape(N, 4), raw[4 * N:8 * N].reshape(N, 4)
    model = raw[8 * N:].reshape(B, 16)
    st = oracle.State(sc, np.float32)
    st.step(steps)
    assert np.array_equal(pos, st.pos) and np.array_equal(vel, st.vel)
    assert np.array_equal(model, st.model)
