"""ctypes binding of include/arsim.h (libarsim.so, built in-tree by diploma_thesis_amd/build.py).

The library is loaded from this directory only; there is no fallback implementation: if the
shared library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ARSIM_LIB selects another build of the same library (for A/B timing of build options)
LIB_PATH = os.environ.get("ARSIM_LIB", os.path.join(HERE, "libarsim.so"))

ARSIM_OK = 0
ARSIM_ERR_INVALID_ARG = 1
ARSIM_ERR_NOT_POW2 = 2
ARSIM_ERR_HIP = 3
ARSIM_ERR_OUT_OF_MEMORY = 4
ARSIM_ERR_WORKSPACE = 5
ARSIM_ERR_NAN = 6
ARSIM_ERR_OVERFLOW = 7
ARSIM_ERR_BUDGET = 8
ARSIM_ERR_NO_DEVICE = 9
ARSIM_ERR_UNSUPPORTED = 10

METHOD_GRID, METHOD_LBVH30, METHOD_LBVH63 = 0, 1, 2
WALLS_OPEN_TOP = 0x37
COMPAT_REF_UNITS, COMPAT_DROP_FIRST, COMPAT_GRAIN_CLAMP, DYNAMIC_AABB = 0x1, 0x2, 0x4, 0x8
CONTACT_SIS = 0x10


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Uint3(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint)]


class ArSimParams(C.Structure):
    """Byte-identical to the reference SimParams (Kernels/particleSystem.cuh:27-55)."""
    _fields_ = [
        ("colliderPos", Float3), ("colliderRadius", C.c_float),
        ("gravity", Float3), ("globalDamping", C.c_float), ("particleRadius", C.c_float),
        ("gridSize", Uint3), ("numCells", C.c_uint),
        ("worldOrigin", Float3), ("cellSize", Float3),
        ("numBodies", C.c_uint), ("maxParticlesPerCell", C.c_uint),
        ("spring", C.c_float), ("damping", C.c_float), ("shear", C.c_float), ("attraction", C.c_float),
        ("boundaryDamping", C.c_float), ("ARrestitution", C.c_float), ("RBrestitution", C.c_float),
        ("ARfriction", C.c_float), ("RBfriction", C.c_float), ("Wdamping", C.c_float), ("Vdamping", C.c_float),
    ]


class ArSimStepConfig(C.Structure):
    _fields_ = [("dt", C.c_float), ("boxMin", C.c_float * 3), ("boxMax", C.c_float * 3),
                ("wallMask", C.c_uint), ("method", C.c_int), ("flags", C.c_uint)]


class ArSimSceneHost(C.Structure):
    _fields_ = [
        ("numParticles", C.c_int), ("numBodies", C.c_int),
        ("pos", C.c_void_p), ("vel", C.c_void_p), ("rbIndices", C.c_void_p), ("bodyFramePos", C.c_void_p),
        ("rbPositions", C.c_void_p), ("rbVelocities", C.c_void_p), ("rbAngularVelocity", C.c_void_p),
        ("rbQuaternion", C.c_void_p), ("rbInertia", C.c_void_p), ("rbMass", C.c_void_p),
        ("particlesPerObject", C.c_void_p),
        ("params", ArSimParams), ("cfg", ArSimStepConfig),
    ]


class ArSimTuning(C.Structure):
    _fields_ = [("demBlock", C.c_int), ("sortItemsPerThread", C.c_int), ("sortRadixBits", C.c_int),
                ("useGraph", C.c_int), ("fuseReduce", C.c_int), ("bucketSplit", C.c_int)]


class ArSimGenSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("numCubes", C.c_int), ("cubeEdge", C.c_int), ("numEllipsoids", C.c_int),
        ("ellipsoidMinSemi", C.c_float), ("ellipsoidMaxSemi", C.c_float), ("numGrains", C.c_int),
        ("box", C.c_float * 3), ("radius", C.c_float), ("bodyVelSigma", C.c_float), ("grainVelSigma", C.c_float),
        ("dropHeightMax", C.c_float), ("gridLog2", C.c_int * 3), ("method", C.c_int),
        ("dt", C.c_float), ("spring", C.c_float), ("damping", C.c_float), ("shear", C.c_float),
        ("gravityY", C.c_float),
    ]


class ArSimPileSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("numEllipsoids", C.c_int), ("targetParticles", C.c_int), ("edge", C.c_float),
        ("radius", C.c_float), ("latticeGap", C.c_float), ("surfaceGap", C.c_float),
        ("gridLog2", C.c_int), ("method", C.c_int),
    ]


class ArSimServeConfig(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("numRequests", C.c_int), ("arrivalRate", C.c_double),
        ("minParticles", C.c_int), ("maxParticles", C.c_int), ("minSteps", C.c_int), ("maxSteps", C.c_int),
        ("hbmBudgetBytes", C.c_double), ("maxResident", C.c_int), ("oneAtATime", C.c_int),
        ("timeScale", C.c_double), ("method", C.c_int), ("resultChecksums", C.c_void_p),
    ]


class ArSimServeReport(C.Structure):
    _fields_ = [
        ("completed", C.c_int), ("wallSeconds", C.c_double), ("particleUpdates", C.c_double),
        ("updatesPerSecond", C.c_double), ("stepLatencyP50Ms", C.c_double), ("stepLatencyP99Ms", C.c_double),
        ("completionLatencyP50Ms", C.c_double), ("completionLatencyP99Ms", C.c_double),
        ("queueWaitP50Ms", C.c_double), ("queueWaitP99Ms", C.c_double), ("maxResidentSeen", C.c_int),
        ("rejectedForBudget", C.c_int), ("peakDeviceBytes", C.c_double), ("deferredForBudget", C.c_int),
    ]


assert C.sizeof(ArSimParams) == 128

VP, I, U, F, SZ = C.c_void_p, C.c_int, C.c_uint, C.c_float, C.c_size_t
PVP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every name is declared in include/arsim.h
SIGNATURES = {
    "arsim_device_count": (I, [C.POINTER(I)]),
    "arsim_set_device": (I, [I]),
    "arsim_malloc": (I, [PVP, SZ]),
    "arsim_free": (I, [VP]),
    "arsim_memcpy_htod": (I, [VP, VP, SZ, VP]),
    "arsim_memcpy_dtoh": (I, [VP, VP, SZ, VP]),
    "arsim_memcpy_dtod": (I, [VP, VP, SZ, VP]),
    "arsim_memset": (I, [VP, I, SZ, VP]),
    "arsim_stream_create": (I, [PVP]),
    "arsim_stream_destroy": (I, [VP]),
    "arsim_stream_synchronize": (I, [VP]),
    "arsim_last_error": (C.c_char_p, []),
    "arsim_abi_version": (I, []),
    "arsim_integrate_bodies": (I, [VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP, VP]),
    "arsim_particle_pose": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP]),
    "arsim_integrate_grains": (I, [VP, VP, VP, I, VP, VP, VP]),
    "arsim_calc_hash": (I, [VP, VP, VP, I, VP, VP]),
    "arsim_sort_workspace_bytes": (SZ, [I, I]),
    "arsim_sort_pairs_u32": (I, [PVP, PVP, PVP, PVP, I, I, I, VP, SZ, VP]),
    "arsim_sort_pairs_u64": (I, [PVP, PVP, PVP, PVP, I, I, I, VP, SZ, VP]),
    "arsim_sort_pairs_u32_noalloc": (I, [PVP, PVP, PVP, PVP, I, I, I, VP]),
    "arsim_reorder_find_cell_start": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, I, U, VP]),
    "arsim_dem_grid": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP, VP]),
    "arsim_segmented_reduce_apply": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP, VP]),
    "arsim_aabb_workspace_bytes": (SZ, [I]),
    "arsim_aabb": (I, [VP, I, VP, VP, VP, SZ, VP]),
    "arsim_morton30": (I, [VP, VP, VP, I, VP, VP, VP, VP]),
    "arsim_morton63": (I, [VP, VP, VP, I, VP, VP, VP, VP]),
    "arsim_lbvh_build": (I, [VP, I, I, VP, VP, VP, VP, VP, VP, VP, VP]),
    "arsim_lbvh_refit": (I, [VP, VP, VP, VP, F, I, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "arsim_dem_lbvh": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP, VP, VP]),
    "arsim_rebase_indices": (I, [VP, VP, I, I, VP]),
    "arsim_contact_pairs_grid": (I, [VP, VP, VP, VP, VP, I, VP, VP, VP, I, VP]),
    "arsim_contact_pairs_lbvh": (I, [VP, VP, VP, VP, VP, VP, I, VP, VP, VP, VP, I, VP]),
    "arsim_engine_create": (I, [VP, I, VP, I, PVP]),
    "arsim_engine_destroy": (I, [VP]),
    "arsim_engine_step": (I, [VP, I]),
    "arsim_engine_synchronize": (I, [VP]),
    "arsim_engine_stream": (I, [VP, PVP]),
    "arsim_engine_num_particles": (I, [VP, I]),
    "arsim_engine_device_bytes": (SZ, [VP]),
    "arsim_engine_download": (I, [VP, I, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "arsim_engine_profile_step": (I, [VP, VP, VP, I]),
    "arsim_engine_debug_buffers": (I, [VP, PVP, PVP, PVP, PVP, PVP, PVP, PVP, PVP, PVP, PVP]),
    "arsim_engine_contact_pairs": (I, [VP, VP, VP, I]),
    "arsim_engine_check_finite": (I, [VP, C.POINTER(I)]),
    "arsim_engine_autotune": (I, [VP, I, VP, C.POINTER(F)]),
    "arsim_engine_set_dt": (I, [VP, F]),
    "arsim_generate_scene": (I, [VP, PVP]),
    "arsim_generate_mix_scene": (I, [I, C.c_uint64, I, PVP]),
    "arsim_generate_packed_pile": (I, [VP, PVP]),
    "arsim_scene_view": (I, [VP, VP]),
    "arsim_scene_free": (I, [VP]),
    "arsim_serve_run": (I, [VP, VP]),
    "arsim_serve_request": (I, [VP, I, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "arsim_engine_state_checksum": (I, [VP, I, C.POINTER(C.c_uint64)]),
    "arsim_template_suffix": (C.c_char_p, [F]),
    "arsim_template_load": (I, [C.c_char_p, C.c_char_p, F, F, PVP]),
    "arsim_template_info": (I, [VP, C.POINTER(I), VP, VP, C.POINTER(F), C.POINTER(F)]),
    "arsim_template_free": (I, [VP]),
    "arsim_scene_create": (I, [VP, VP, PVP]),
    "arsim_scene_add_body": (I, [VP, VP, VP, VP, VP]),
    "arsim_scene_add_grains": (I, [VP, I, VP, VP]),
    "arsim_engine_snapshot": (I, [VP, I, PVP]),
    "arsim_engine_render_state": (I, [VP, I, VP, VP]),
    "arsim_engine_rebuild": (I, [VP, VP, VP, I, VP, I]),
    "arsim_engine_set_static": (I, [VP, VP, VP, I]),
    "arsim_range_image_load": (I, [VP, I, I, VP, F, VP, VP, VP, VP, VP]),
    "arsim_dem_static_grid": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP, VP]),
    "arsim_find_deepest_contacts_grid": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP]),
    "arsim_find_static_contacts_grid": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP]),
    "arsim_sis_workspace_bytes": (SZ, [I, I]),
    "arsim_sis_gather": (I, [VP, VP, VP, VP, VP, I, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, I, VP, VP, SZ,
                             VP]),
    "arsim_sis_solve": (I, [VP, VP, VP, VP, VP, VP, I, VP, VP, VP, VP, VP, I, I, VP, VP, SZ, VP]),
}


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(f"libarsim.so not built at {path}: run diploma_thesis_amd/build.py "
                          "(there is no CPU fallback for the HIP path)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


class ArSimError(RuntimeError):
    def __init__(self, code, what):
        msg = lib.arsim_last_error()
        super().__init__(f"{what} failed with status {code}: {msg.decode() if msg else ''}")
        self.code = code


def check(code, what="arsim call"):
    if code != ARSIM_OK:
        raise ArSimError(code, what)
    return code


def ptr(a):
    """Host pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)
