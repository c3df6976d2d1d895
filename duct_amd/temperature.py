"""Mixed-boundary (temperature) Helmholtz solve on the device.

The host mirror of pseudomaty_temp / pseudomatz_temp (src/pseudomat_temp.f90:2-321) and
helm3d_temp (src/helm3d_temp.f90:2-44):

  setup_temp(m, bounds, sy, sz, mz)
                                  builds both directions' bases in FP64 on the host (the
                                  reference's operation order for the wall elimination,
                                  LAPACK through numpy), y of size my = m and z of size mz
                                  (default m), and uploads them (duct_setup_temp_)
  helm3d_temp(rhs, sol, bcy, bcz, alambda, alp, ib)
                                  duct_helm3d_temp_dev_: rhs, sol (nmodes, mz, my), the wall
                                  data bcy (nmodes, 2, mz) over z and bcz (nmodes, 2, my) over y

bounds is four characters 'd' / 'n' for the walls y=-1, y=+1, z=-1, z=+1 (the module
variables boundy1, boundy2, boundz1, boundz2 of the reference).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .solver import _use_current_stream


def cheb_derivative(m: int, s: float) -> np.ndarray:
    """The CHQZ collocation derivative matrix of pseudomat_temp.f90:60-94 scaled by 1/s (the
    textbook entries the reference's elimination and edges use)."""
    N = m - 1
    y = -np.cos(np.pi * np.arange(m) / N)
    dd = np.ones(m)
    dd[0] = dd[-1] = 2.0
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    with np.errstate(divide="ignore", invalid="ignore"):
        D = (dd[i] / dd[j]) * np.where((i + j) % 2, -1.0, 1.0) / (y[i] - y[j])
    D[np.arange(m), np.arange(m)] = 0.0
    D[0, 0] = -(2.0 * N * N + 1.0) / 6.0
    D[-1, -1] = -D[0, 0]
    for k in range(1, N):
        D[k, k] = -0.5 * y[k] / (1.0 - y[k] ** 2)
    return D / s


def pseudomat_temp(bound1: str, bound2: str, m: int, s: float):
    """(D, P, Pinv, E) of one direction: the interior operator with the walls eliminated per
    type (pseudomat_temp.f90:106-150), its eigenvectors, their inverse and the eigenvalues."""
    if bound1 not in "dn" or bound2 not in "dn" or len(bound1) != 1 or len(bound2) != 1:
        raise ValueError("bound must be 'd' or 'n'")
    D = cheb_derivative(m, s)
    D2 = D @ D
    n = m - 1
    if bound1 == "n" and bound2 == "n":
        det = D[0, 0] * D[-1, -1] - D[0, -1] * D[-1, 0]
        ter1 = D2[1:n, :1] * (D[0, -1] * D[-1, 1:n] - D[-1, -1] * D[0, 1:n])[None, :] / det
        ter3 = D2[1:n, -1:] * (D[-1, 0] * D[0, 1:n] - D[0, 0] * D[-1, 1:n])[None, :] / det
        D2T = ter1 + D2[1:n, 1:n] + ter3
    elif bound1 == "d" and bound2 == "d":
        D2T = D2[1:n, 1:n].copy()
    elif bound1 == "d":
        D2T = D2[1:n, 1:n] - D2[1:n, -1:] * D[-1, 1:n][None, :] / D[-1, -1]
    else:
        D2T = D2[1:n, 1:n] - D2[1:n, :1] * D[0, 1:n][None, :] / D[0, 0]
    E, P = np.linalg.eig(D2T)
    E, P = np.ascontiguousarray(E.real), np.ascontiguousarray(P.real)
    return D, P, np.linalg.inv(P), E


_state = {"temp": None}


def setup_temp(m: int, bounds: str, sy: float = 1.0, sz: float = 1.0, mz: int | None = None) -> None:
    """pseudomaty_temp(boundy1, boundy2, my) and pseudomatz_temp(boundz1, boundz2, mz), sy = a,
    sz = b of helm3d_temp's ipro <= 0 branch, uploaded with duct_setup_temp_; my = m."""
    if len(bounds) != 4 or any(c not in "dn" for c in bounds):
        raise ValueError("bounds = four of 'd' / 'n' for y=-1, y=+1, z=-1, z=+1")
    mz = m if mz is None else int(mz)
    Dy, Py, Pyi, Ey = pseudomat_temp(bounds[0], bounds[1], m, sy)
    Dz, Pz, Pzi, Ez = pseudomat_temp(bounds[2], bounds[3], mz, sz)
    cm = [np.asfortranarray(x) for x in (Py, Pyi, Dy, Pz, Pzi, Dz)]
    ptr = [x.ctypes.data_as(ctypes.c_void_p) for x in cm]
    _lib.call("duct_setup_temp_", bounds.encode(), _lib.byref_int(m), _lib.byref_int(mz), ptr[0], ptr[1],
              Ey.ctypes.data_as(ctypes.c_void_p), ptr[2], ptr[3], ptr[4], Ez.ctypes.data_as(ctypes.c_void_p),
              ptr[5], ctypes.c_size_t(4))
    _state["temp"] = dict(my=m, mz=mz, bounds=bounds, sy=sy, sz=sz)


def helm3d_temp(rhs, sol, bcy, bcz, alambda: float, alp: float = 1.0, ib: int = 0) -> None:
    """sol = helm3d_temp(rhs) on device tensors: rhs, sol complex128 (nmodes, mz, my) [mode, z, y];
    bcy float64 (nmodes, 2, mz) [mode, wall, z], bcz (nmodes, 2, my) [mode, wall, y]."""
    import torch
    for t in (rhs, sol):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex128):
            raise TypeError("fields must be complex128 CUDA tensors")
        if t.dim() != 3 or not t.is_contiguous():
            raise ValueError("field must be contiguous (nmodes, mz, my) = [mode, z, y], got %s" % (tuple(t.shape),))
    if sol.shape != rhs.shape:
        raise ValueError("rhs and sol shapes differ")
    M, mz, my = rhs.shape
    for b, n, name in ((bcy, mz, "bcy"), (bcz, my, "bcz")):
        if tuple(b.shape) != (M, 2, n) or not b.is_contiguous() or b.dtype != torch.float64 or not b.is_cuda:
            raise ValueError("%s must be contiguous float64 CUDA (nmodes, 2, %d)" % (name, n))
    _use_current_stream()
    _lib.call("duct_helm3d_temp_dev_", *[_lib.byref_ptr(t.data_ptr()) for t in (rhs, sol, bcy, bcz)],
              _lib.byref_int(my), _lib.byref_int(mz), _lib.byref_int(M), _lib.byref_int(ib),
              _lib.byref_double(alp), _lib.byref_double(alambda))
