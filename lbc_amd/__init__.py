"""MI355X-native lattice-Boltzmann hot path of lbc (D2Q9 BGK step + boundaries).

The product is the shared library ``lbc_amd/_lib/liblbm_hip.so`` built from
``lbc_amd/csrc`` (gfx950 HIP kernels + a C drop-in layer).  This package only
loads it:

* :func:`lib` -- the loaded library (``ctypes.CDLL``); raises if it is missing.
* :mod:`lbc_amd.native` -- Python view of the native C ABI (include/lbmx.h).
* :mod:`lbc_amd.capi`   -- ctypes binding of the reference's own C API
  (include/lbm_core.h, lbm_bc.h) as the library exports it.

There is no CPU fallback: every entry point runs on the GPU or fails.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblbm_hip.so")
_lib = None


def lib():
    """Load (once) and return the HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m lbc_amd.build` "
                "(there is no CPU fallback)")
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    return _lib
