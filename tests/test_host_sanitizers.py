"""The pure host C++ -- scene generators (scenegen.cpp), template loader and scene builder
(sceneio.cpp) and the oracle -- built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer
and run on small scenes of every kind (lattice scene with cubes, ellipsoids and grains under the grid
and both LBVH broadphases, the packed pile, the reference's own templates, a range image).  Any
sanitizer report makes the driver exit non-zero."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = [os.path.join(ROOT, "tests", "host", "sanitize_driver.cpp"), os.path.join(ROOT, "oracle", "arsim_oracle.cpp"),
           os.path.join(ROOT, "diploma_thesis_amd", "csrc", "scenegen.cpp"),
           os.path.join(ROOT, "diploma_thesis_amd", "csrc", "sceneio.cpp")]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "hostsan")
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                            "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-ffp-contract=off",
                            *SOURCES, "-o", exe], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "OBJparticles")], capture_output=True, text=True,
                         env=env, timeout=600)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-6000:])
    assert "runtime error" not in run.stderr
    assert "host code clean under the sanitizers" in run.stdout
