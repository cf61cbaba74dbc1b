"""Build recipe of the MI355X library (``lbc_amd/_lib/liblbm_hip.so``).

Compiles every HIP source for gfx950 with hipcc and the C drop-in layer with gcc,
then links one shared library exporting both the reference's C API
(include/lbm_core.h, lbm_bc.h, lbm_utils.h) and the native C ABI (include/lbmx.h).

    python -m lbc_amd.build            # or __graft_entry__.build()
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lbc_amd")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "liblbm_hip.so")
ARCH = os.environ.get("LBMX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

# -ffp-contract=off: FP64 kernels keep the reference's separate multiply and add
# roundings (its x86-64 build has no FMA); the FP32 collision writes its fused
# multiply-adds explicitly (__fmaf_rn).  Packed-FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{ROOT}/include", f"-I{CSRC}/hip",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-parameter",
           "-ffp-contract=off", f"-I{ROOT}/include", f"-I{CSRC}/capi"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({' '.join(cmd[:3])} ...):\n{out}")
    if out:
        print(out, file=sys.stderr)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "hip", "*.hip")))
    c_srcs = sorted(glob.glob(os.path.join(CSRC, "capi", "*.c")))
    for src in hip_srcs:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        if verbose:
            print(f"[hipcc {ARCH}] {os.path.relpath(src, ROOT)}", file=sys.stderr)
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", o])
        objs.append(o)
    for src in c_srcs:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        if verbose:
            print(f"[gcc] {os.path.relpath(src, ROOT)}", file=sys.stderr)
        _run([CC, *C_FLAGS, "-c", src, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,--no-undefined", "-lm"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {os.path.relpath(LIB, ROOT)}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
