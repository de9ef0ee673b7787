"""Builds libarsim.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Every source is compiled on every call (the caller wants a fresh build); objects go to
diploma_thesis_amd/_build/ and the shared library to diploma_thesis_amd/libarsim.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libarsim.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-reproducible FP32: no a*b+c contraction, IEEE division and square root
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-value",
    "-Wno-unused-result",
]

SOURCES = ["sort.hip", "body.hip", "grid.hip", "lbvh.hip", "static.hip", "sis.hip", "bucket.hip", "capi.cpp", "engine.cpp", "serve.cpp", "scenegen.cpp",
           "sceneio.cpp"]


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if cand.startswith("/") and os.path.exists(cand):
            return cand
    return "hipcc"


def _compile(src, build_dir=BUILD, defines=()):
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), "-x", "hip", *HIP_FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err


def build(verbose=False, out=LIB, build_dir=BUILD, defines=()):
    """Compiles every source and links `out`; `defines` are extra -D options (build variants)."""
    os.makedirs(build_dir, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, build_dir, defines), SOURCES))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = out + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, out)
    return out


# Counting build of the LBVH packet traversal (nodes visited, leaf tests, stack depth), loaded through
# ARSIM_LIB by `bench.py --traversal-stats`; never the product library.
STATS_LIB = os.path.join(HERE, "libarsim_traversal_stats.so")


def build_traversal_stats(verbose=False):
    return build(verbose=verbose, out=STATS_LIB, build_dir=os.path.join(HERE, "_build_stats"),
                 defines=("ARSIM_TRAVERSAL_STATS",))


if __name__ == "__main__":
    print(build(verbose=True))
    if "--traversal-stats" in sys.argv:
        print(build_traversal_stats(verbose=True))
