"""MI355X-native (gfx950 HIP) per-step DEM pipeline of AR_Sim: Morton / grid keys, onesweep radix
sort, Karras LBVH and uniform grid broadphase, spring-dashpot-shear contacts, segmented per-body
reduction and quaternion integration, with batched and served forms.

The compute path is libarsim.so (hand-written HIP kernels behind the C ABI of include/arsim.h);
this package is the Python plumbing around it.  Importing it loads the library and fails if it
has not been built (python diploma_thesis_amd/build.py).
"""
from . import _capi as capi
from ._capi import ArSimError, lib
from .engine import Engine, contact_set
from . import scenes

__all__ = ["capi", "lib", "ArSimError", "Engine", "contact_set", "scenes"]
