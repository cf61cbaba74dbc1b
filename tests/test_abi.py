"""CPU checks of the drop-in boundary (no GPU calls).

* liblbm_hip.so loads and exports every function declared in include/*.h.
* The public headers declare the same prototypes as the reference's
  include/*.h (compared after preprocessing, whitespace-normalised), so the
  reference's drivers compile against them unchanged.
* The reference's drivers that build against its own headers (cases/nit/nit.c,
  cases/testbc/testbc.c) compile and link against the library.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "lbc_amd", "_lib", "liblbm_hip.so")
REF = "/root/reference"


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INC)):
        if not h.endswith(".h"):
            continue
        pre = subprocess.run(["gcc", "-E", "-P", "-I", INC, os.path.join(INC, h)],
                             capture_output=True, text=True, check=True).stdout
        pre = pre.replace('__attribute__((visibility("default")))', "")
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*"
                             r"([A-Za-z][A-Za-z0-9_]*)\s*\(", pre, re.M):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof", "void", "int"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB), "build the library first (python -m lbc_amd.build)"
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert {"main_lbm", "mei_initialization", "velocity_inlet_regularized",
            "write_vtk_binary_2D", "lbmx_step", "lbmx_create"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def _prototypes(include_dir, header):
    src = f'#include "{header}"\n'
    pre = subprocess.run(["gcc", "-E", "-P", "-I", include_dir, "-x", "c", "-"], input=src,
                         capture_output=True, text=True, check=True).stdout
    text = re.sub(r"\s+", " ", pre)
    protos = {}
    for m in re.finditer(r"(void|int) ([a-z_0-9]+)\s*\(([^;{]*)\)\s*;", text):
        params = re.sub(r"\s*([,\[\]\(\)*])\s*", r"\1", m.group(3)).strip()
        protos[m.group(2)] = (m.group(1), params)
    return protos


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")),
                    reason="reference tree absent")
@pytest.mark.parametrize("header", ["lbm_types.h", "lbm_core.h", "lbm_bc.h", "lbm_utils.h"])
def test_headers_identical_to_reference(header):
    ours = _prototypes(INC, header)
    theirs = _prototypes(os.path.join(REF, "include"), header)
    for name, proto in theirs.items():
        assert ours.get(name) == proto, name
    extra = set(ours) - set(theirs)
    assert extra <= {"convective_top", "convective_bottom"}, extra
    if header == "lbm_types.h":
        strip = lambda p: re.sub(r"\s+", " ", subprocess.run(   # noqa: E731
            ["gcc", "-E", "-P", "-I", p, "-x", "c", "-"], input=f'#include "{header}"\n',
            capture_output=True, text=True, check=True).stdout).strip()
        assert strip(INC) == strip(os.path.join(REF, "include"))


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "cases")), reason="reference tree absent")
@pytest.mark.parametrize("driver", ["nit/nit.c", "testbc/testbc.c"])
def test_reference_drivers_link_unchanged(driver, tmp_path):
    exe = tmp_path / "driver"
    r = subprocess.run(["gcc", "-Wall", "-O3", "-I", INC, "-o", str(exe),
                        os.path.join(REF, "cases", driver), "-L", os.path.dirname(LIB),
                        "-llbm_hip", f"-Wl,-rpath,{os.path.dirname(LIB)}", "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()
