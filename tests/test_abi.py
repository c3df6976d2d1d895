"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and rejects bad arguments before touching the device."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from duct_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "duct_hip.h")


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\bint\s+(duct_[a-z0-9_]+_)\s*\(", text)))


def test_header_and_binding_agree():
    syms = header_symbols()
    assert syms, "no declarations found"
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for s in header_symbols():
        assert getattr(L, s) is not None


@pytest.mark.skipif(not shutil.which("gcc"), reason="no C compiler")
def test_header_is_plain_c(tmp_path):
    """C callers (the reference's fortran.c style) include the header as C99."""
    src = tmp_path / "use.c"
    src.write_text('#include "duct_hip.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                    "-I", os.path.dirname(HEADER), str(src)], check=True)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _lib.open_library(str(tmp_path / "nope.so"))


def test_setup_rejects_bad_arguments_on_host():
    L = _lib.load()
    m = ctypes.c_int(9)
    z = np.zeros(100)
    iz = ctypes.c_int(1)
    w = ctypes.c_double(0.0)
    # bad bound character: rejected before any HIP call (pseudomat.f90:300-303 stops)
    st = L.duct_setup_(b"x", ctypes.byref(m), z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                       z.ctypes.data, ctypes.byref(iz), ctypes.byref(w), 1)
    assert st == 1 and "bound" in _lib.last_error()
    # size outside the supported range
    big = ctypes.c_int(100000)
    st = L.duct_setup_(b"d", ctypes.byref(big), z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                       None, None, None, 1)
    assert st == 5
    # null pointers
    st = L.duct_setup_(b"d", ctypes.byref(m), None, None, None, None, None, None, None, 1)
    assert st == 1


def test_solver_rejects_bad_bound_before_the_library():
    import duct_amd
    with pytest.raises(ValueError):
        duct_amd.solvediag(np.zeros((9, 9)), "x", 1.0)
