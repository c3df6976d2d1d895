"""CPU oracle for the duct cross-plane Helmholtz/Poisson solve stage.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the checker
(or as the timed CPU baseline).  The duct_amd product path never imports it.

Contents
  pseudomat.py   numpy restatement of pseudomat.f90 (one-time eigenbases)
  solvediag.c    C restatement of solvediag.f90 / int2d3d.f90 / helm3d.f90
  shen.py        textbook Shen-Legendre Galerkin solver (cross-check of the
                 helm_shen.f90 chain; agrees to truncation error, not 1e-12)
  polynomials.py exact Chebyshev-polynomial helpers for manufactured solutions

Field layout everywhere: complex128 arrays of shape (M, m, m) indexed
[mode, z, y] in C order, which is the reference's complex(8) (my, mz, M)
column-major storage (helm3d.f90:29).  A real plane is (m, m) indexed [z, y].

Parity status of each function is listed in the header of tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .pseudomat import Basis, cgl_nodes, colloc_deriv, pseudomat  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_duct.so")


class _CBasis(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int),
        ("ED2", ctypes.c_void_p), ("PDT", ctypes.c_void_p), ("PDM1T", ctypes.c_void_p),
        ("EN", ctypes.c_void_p), ("PN", ctypes.c_void_p), ("PNT", ctypes.c_void_p),
        ("PNM1T", ctypes.c_void_p), ("AM1", ctypes.c_void_p), ("wanted", ctypes.c_double),
        ("iofzero", ctypes.c_int), ("DC", ctypes.c_void_p), ("D2bx", ctypes.c_void_p),
        ("D2by", ctypes.c_void_p),
    ]


def build(force: bool = False) -> str:
    """Compile solvediag.c into liboracle_duct.so with the Makefile beside it."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []) + ["liboracle_duct.so"],
                       check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, d, c = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char
        L.oracle_solvediag.argtypes = [ctypes.POINTER(_CBasis), vp, c, d, vp]
        L.oracle_solvediag.restype = i
        L.oracle_int2d3d.argtypes = [ctypes.POINTER(_CBasis), vp, vp, c, d, i, vp]
        L.oracle_int2d3d.restype = i
        L.oracle_helm3d.argtypes = [ctypes.POINTER(_CBasis), vp, vp, i, i, d, d, c, i]
        L.oracle_helm3d.restype = i
        L.oracle_work_doubles.argtypes = [i]
        L.oracle_work_doubles.restype = ctypes.c_size_t
        L.oracle_set_dgemm.argtypes = [vp]
        L.oracle_set_dgemm.restype = None
        _lib = L
    return _lib


def use_blas(enable: bool = True) -> str:
    """Point the oracle's gemm at scipy's bundled OpenBLAS dgemm (single call,
    Fortran ABI), or back at the built-in loop.  Returns a description."""
    L = lib()
    if not enable:
        L.oracle_set_dgemm(None)
        return "built-in triple loop"
    import glob
    import scipy
    libdir = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
    cands = sorted(glob.glob(os.path.join(libdir, "libscipy_openblas-*.so")))
    for path in cands:
        try:
            bl = ctypes.CDLL(path)
            fn = bl.scipy_dgemm_
        except (OSError, AttributeError):
            continue
        L.oracle_set_dgemm(ctypes.cast(fn, ctypes.c_void_p))
        # one BLAS thread per call: the oracle parallelises over modes instead
        for name in ("scipy_openblas_set_num_threads", "openblas_set_num_threads"):
            if hasattr(bl, name):
                getattr(bl, name)(1)
                break
        return "OpenBLAS dgemm (" + os.path.basename(path) + ", 1 thread per call)"
    L.oracle_set_dgemm(None)
    return "built-in triple loop"


class OracleBasis:
    """ctypes view of a pseudomat Basis; keeps the Fortran-ordered arrays alive."""

    def __init__(self, b: Basis):
        self.b = b
        keep = {}

        def f(name):
            a = np.asfortranarray(getattr(b, name), dtype=np.float64)
            keep[name] = a
            return a.ctypes.data

        self.c = _CBasis(b.m, f("ED2"), f("PDT"), f("PDM1T"), f("EN"), f("PN"), f("PNT"),
                         f("PNM1T"), f("AM1"), float(b.wanted), int(b.iofzero), f("DC"),
                         f("D2bx"), f("D2by"))
        self._keep = keep


def solvediag(ob: OracleBasis, U: np.ndarray, bound: str, xlambda: float) -> np.ndarray:
    """One real plane U indexed [z, y] (C order == Fortran U(y,z)); returns a new array."""
    m = ob.b.m
    out = np.ascontiguousarray(U, dtype=np.float64).copy()
    assert out.shape == (m, m)
    work = np.zeros(lib().oracle_work_doubles(m))
    st = lib().oracle_solvediag(ctypes.byref(ob.c), out.ctypes.data, bound.encode(), float(xlambda),
                                work.ctypes.data)
    if st:
        raise ValueError("oracle solvediag: bad bound %r" % bound)
    return out


def helm3d(ob: OracleBasis, rhs: np.ndarray, bound: str, alambda: float, alp: float = 1.0,
           ib: int = 0, nthreads: int = 1, inplace_rhs: bool = False):
    """helm3d over a complex field (M, m, m) [mode, z, y].  Returns (sol, rhs_after);
    rhs_after carries the in-place boundary zeroing helm3d.f90:48-60 performs."""
    m = ob.b.m
    r = rhs if inplace_rhs else np.array(rhs, dtype=np.complex128, copy=True, order="C")
    assert r.dtype == np.complex128 and r.flags.c_contiguous and r.shape[1:] == (m, m)
    sol = np.zeros_like(r)
    st = lib().oracle_helm3d(ctypes.byref(ob.c), r.ctypes.data, sol.ctypes.data, r.shape[0], int(ib),
                             float(alp), float(alambda), bound.encode(), int(nthreads))
    if st:
        raise ValueError("oracle helm3d failed (status %d)" % st)
    return sol, r
