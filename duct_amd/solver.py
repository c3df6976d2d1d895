"""Host-side mirror of the reference's solve-stage interface over the HIP library.

Reference interface (src/ of ando2800/duct) -> here
  pseudomat(bound, m)                          -> setup(m) / upload_basis(...)
  helm3d(rhs, sol, work, ipro, bound, ...)     -> helm3d(rhs, sol, bound, alambda, ...)
  int2d3d(sol, rhs, work, bound, xlambda, im)  -> int2d3d(rhs, bound, xlambda, imode)
  solvediag(U, F, FT, bound, xlambda)          -> solvediag(U, bound, xlambda)

Fields are complex128 tensors of shape (nmodes, m, m) indexed [mode, z, y]
(C order), the reference's complex(8) rhs(my, mz, ib:ie) storage.  Device
fields are torch CUDA tensors; torch only provides the memory and the stream,
every operation on them is a launch of libduct_hip.so.  A bad bound character
raises (the reference prints and stops, solvediag.f90:324-326).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .basis import Setup, setup as build_setup

_state = {"m": None, "setup": None}


def _bound_bytes(bound: str) -> bytes:
    if bound not in ("d", "n"):
        raise ValueError("in solvediag error in BC: bound %r is neither 'd' nor 'n'" % (bound,))
    return bound.encode()


def init(device: int = 0) -> None:
    _lib.call("duct_init_", _lib.byref_int(device))


def shutdown() -> None:
    _lib.call("duct_shutdown_")
    _state["m"] = None
    _state["setup"] = None


def set_stream(stream_handle: int) -> None:
    _lib.call("duct_set_stream_", _lib.byref_ptr(stream_handle))


def synchronize() -> None:
    _lib.call("duct_synchronize_")


def upload_basis(bound: str, m: int, P, Pinv, lam, DC, AM1=None, iofzero: int = 0, wanted: float = 0.0):
    """duct_setup_ with caller-supplied arrays (what a Fortran pseudomat hands over)."""
    def f64(a):
        return np.asfortranarray(np.asarray(a, dtype=np.float64))
    P, Pinv, lam, DC = f64(P), f64(Pinv), f64(lam), f64(DC)
    am1 = f64(AM1) if AM1 is not None else np.zeros((4, 4), order="F")
    _lib.call("duct_setup_", _bound_bytes(bound), _lib.byref_int(m), P.ctypes.data, Pinv.ctypes.data,
              lam.ctypes.data, DC.ctypes.data, am1.ctypes.data, _lib.byref_int(iofzero),
              _lib.byref_double(wanted), ctypes.c_size_t(1))


def setup(m: int) -> Setup:
    """pseudomat('d', m) + pseudomat('n', m): host FP64 eigensolve, then upload."""
    s = build_setup(m)
    for b in (s.dirichlet, s.neumann):
        upload_basis(b.bound, m, b.P, b.Pinv, b.lam, s.DC, b.AM1, b.iofzero, b.wanted)
    _state["m"] = m
    _state["setup"] = s
    return s


def _check_field(t, m: int):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex128):
        raise TypeError("fields must be complex128 CUDA tensors")
    if t.dim() != 3 or tuple(t.shape[1:]) != (m, m) or not t.is_contiguous():
        raise ValueError("field must be contiguous (nmodes, m, m) = [mode, z, y], got %s" % (tuple(t.shape),))


def _use_current_stream():
    import torch
    set_stream(torch.cuda.current_stream().cuda_stream)


def helm3d(rhs, sol, bound: str, alambda: float, alp: float = 1.0, ib: int = 0) -> None:
    """helm3d.f90:2 with ipro >= 1 on device fields.  rhs's boundary is zeroed
    in place for bound 'd' (helm3d.f90:48-60)."""
    m = rhs.shape[-1]
    _check_field(rhs, m)
    _check_field(sol, m)
    if sol.shape != rhs.shape:
        raise ValueError("rhs and sol shapes differ")
    _use_current_stream()
    _lib.call("duct_helm3d_dev_", _lib.byref_ptr(rhs.data_ptr()), _lib.byref_ptr(sol.data_ptr()),
              _lib.byref_int(m), _lib.byref_int(rhs.shape[0]), _lib.byref_int(ib),
              _lib.byref_double(alp), _lib.byref_double(alambda), _bound_bytes(bound), ctypes.c_size_t(1))


def helm3d_batch(rhs_list, sol_list, bounds: str, alambdas, alp: float = 1.0, ib: int = 0) -> None:
    """Several fields (e.g. u, v, w Dirichlet + p Neumann) in one pass sequence."""
    nf = len(rhs_list)
    if nf != len(sol_list) or nf != len(bounds) or nf != len(alambdas):
        raise ValueError("rhs, sol, bounds, alambdas lengths differ")
    m = rhs_list[0].shape[-1]
    for t in list(rhs_list) + list(sol_list):
        _check_field(t, m)
        if t.shape != rhs_list[0].shape:
            raise ValueError("all fields must have the same shape")
    for b in bounds:
        _bound_bytes(b)
    _use_current_stream()
    R = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in rhs_list])
    S = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in sol_list])
    A = (ctypes.c_double * nf)(*[float(a) for a in alambdas])
    _lib.call("duct_helm3d_batch_dev_", _lib.byref_int(nf), R, S, bounds.encode(), A, _lib.byref_int(m),
              _lib.byref_int(rhs_list[0].shape[0]), _lib.byref_int(ib), _lib.byref_double(alp),
              ctypes.c_size_t(nf))


def helm3d_host(rhs: np.ndarray, bound: str, alambda: float, alp: float = 1.0, ib: int = 0):
    """duct_helm3d_ on host arrays; returns (sol, rhs_after)."""
    r = np.array(rhs, dtype=np.complex128, order="C", copy=True)
    if r.ndim != 3 or r.shape[1] != r.shape[2]:
        raise ValueError("rhs must be (nmodes, m, m)")
    sol = np.zeros_like(r)
    m = r.shape[-1]
    _lib.call("duct_helm3d_", r.ctypes.data, sol.ctypes.data, _lib.byref_int(m), _lib.byref_int(r.shape[0]),
              _lib.byref_int(ib), _lib.byref_double(alp), _lib.byref_double(alambda), _bound_bytes(bound),
              ctypes.c_size_t(1))
    return sol, r


def int2d3d(rhs_plane: np.ndarray, bound: str, xlambda: float, imode: int) -> np.ndarray:
    """int2d3d.f90:2 on one complex host plane (m, m) [z, y]; returns sol."""
    r = np.ascontiguousarray(rhs_plane, dtype=np.complex128)
    sol = np.zeros_like(r)
    _lib.call("duct_int2d3d_", sol.ctypes.data, r.ctypes.data, _bound_bytes(bound), _lib.byref_double(xlambda),
              _lib.byref_int(imode), ctypes.c_size_t(1))
    return sol


def solvediag(U: np.ndarray, bound: str, xlambda: float) -> np.ndarray:
    """solvediag.f90:2 on one real host plane (m, m) [z, y]; returns the solved plane."""
    out = np.array(U, dtype=np.float64, order="C", copy=True)
    _lib.call("duct_solvediag_", out.ctypes.data, _bound_bytes(bound), _lib.byref_double(xlambda),
              ctypes.c_size_t(1))
    return out


def set_fold(mode) -> None:
    """Even/odd folded eigenbasis for the FP64 solves (duct_set_fold_):
    0/False off, 1/True Dirichlet, 2 Dirichlet and Neumann."""
    _lib.call("duct_set_fold_", _lib.byref_int(int(mode)))


def set_overlap(on: bool = True) -> None:
    """Run the Neumann group of a batched solve on a second stream, concurrently with the Dirichlet
    group (duct_set_overlap_; on by default, not while kernel timing is enabled)."""
    _lib.call("duct_set_overlap_", _lib.byref_int(1 if on else 0))


def set_fused_quarters(on: bool = True) -> None:
    """Folded passes 2 and 3 of each 128 x 128 quarter in one kernel with the quarter in LDS
    (duct_set_fused_quarters_; on by default); off: two planar launches.  Same results to the bit."""
    _lib.call("duct_set_fused_quarters_", _lib.byref_int(1 if on else 0))


def fold_size(bound: str) -> int:
    """Folded size NH of the basis set up for bound (0: the basis does not fold)."""
    v = ctypes.c_int(0)
    _lib.call("duct_fold_size_", _bound_bytes(bound), ctypes.byref(v), ctypes.c_size_t(1))
    return v.value


class KernelTimer:
    """Per-kernel-class device time from hipEvents the library records on its stream."""

    def __enter__(self):
        _lib.call("duct_prof_enable_", _lib.byref_int(1))
        return self

    def __exit__(self, *exc):
        n = len(_lib.PROF_CLASSES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        _lib.call("duct_prof_read_", _lib.byref_int(n), ms, cnt)
        self.ms = {k: ms[i] for i, k in enumerate(_lib.PROF_CLASSES)}
        self.count = {k: cnt[i] for i, k in enumerate(_lib.PROF_CLASSES)}
        _lib.call("duct_prof_enable_", _lib.byref_int(0))
        return False


SCHEMES = {"f32": 1, "bf16x3": 2, "bf16x6": 3, "bf16x1": 4, "fold16": 5, "fold16q": 6}


def helm3d_mixed_batch(rhs_list, sol_list, bounds: str, alambdas, alp: float = 1.0, ib: int = 0,
                       scheme0: str = "bf16x3", scheme1: str = "bf16x6", niter: int = 1, stats: bool = False):
    """Mixed-precision refined solve of several fields (duct_helm3d_mixed_batch_dev_).
    Returns None, or with stats=True a numpy array (niter+1, nfields, 2) of
    [max|r_it|, max|f|] per refinement step (the last after the final correction)."""
    import numpy as np
    nf = len(rhs_list)
    if nf != len(sol_list) or nf != len(bounds) or nf != len(alambdas):
        raise ValueError("rhs, sol, bounds, alambdas lengths differ")
    if scheme0 not in SCHEMES or scheme1 not in SCHEMES:
        raise ValueError("scheme must be one of %s" % sorted(SCHEMES))
    m = rhs_list[0].shape[-1]
    for t in list(rhs_list) + list(sol_list):
        _check_field(t, m)
        if t.shape != rhs_list[0].shape:
            raise ValueError("all fields must have the same shape")
    for b in bounds:
        _bound_bytes(b)
    _use_current_stream()
    R = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in rhs_list])
    S = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in sol_list])
    A = (ctypes.c_double * nf)(*[float(a) for a in alambdas])
    st = np.zeros((niter + 1, nf, 2)) if stats else None
    _lib.call("duct_helm3d_mixed_batch_dev_", _lib.byref_int(nf), R, S, bounds.encode(), A, _lib.byref_int(m),
              _lib.byref_int(rhs_list[0].shape[0]), _lib.byref_int(ib), _lib.byref_double(alp),
              _lib.byref_int(SCHEMES[scheme0]), _lib.byref_int(SCHEMES[scheme1]), _lib.byref_int(niter),
              st.ctypes.data if stats else None, ctypes.c_size_t(nf))
    return st
