"""Mixed-boundary (temperature) Helmholtz solve on the device.

The host mirror of pseudomaty_temp / pseudomatz_temp (src/pseudomat_temp.f90:2-321) and
helm3d_temp (src/helm3d_temp.f90:2-44):

  setup_temp(m, bounds, sy, sz)   builds both directions' bases in FP64 on the host (the
                                  reference's operation order for the wall elimination,
                                  LAPACK through numpy) and uploads them (duct_setup_temp_)
  helm3d_temp(rhs, sol, bcy, bcz, alambda, alp, ib)
                                  duct_helm3d_temp_dev_: sol = the solution with the wall
                                  data bcy (nmodes, 2, m) over z and bcz (nmodes, 2, m) over y

bounds is four characters 'd' / 'n' for the walls y=-1, y=+1, z=-1, z=+1 (the module
variables boundy1, boundy2, boundz1, boundz2 of the reference).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .solver import _check_field, _use_current_stream


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


def setup_temp(m: int, bounds: str, sy: float = 1.0, sz: float = 1.0) -> None:
    """pseudomaty_temp(boundy1, boundy2, my) and pseudomatz_temp(boundz1, boundz2, mz), sy = a,
    sz = b of helm3d_temp's ipro <= 0 branch, uploaded with duct_setup_temp_."""
    if len(bounds) != 4 or any(c not in "dn" for c in bounds):
        raise ValueError("bounds = four of 'd' / 'n' for y=-1, y=+1, z=-1, z=+1")
    Dy, Py, Pyi, Ey = pseudomat_temp(bounds[0], bounds[1], m, sy)
    Dz, Pz, Pzi, Ez = pseudomat_temp(bounds[2], bounds[3], m, sz)
    cm = [np.asfortranarray(x) for x in (Py, Pyi, Dy, Pz, Pzi, Dz)]
    ptr = [x.ctypes.data_as(ctypes.c_void_p) for x in cm]
    _lib.call("duct_setup_temp_", bounds.encode(), _lib.byref_int(m), ptr[0], ptr[1],
              Ey.ctypes.data_as(ctypes.c_void_p), ptr[2], ptr[3], ptr[4], Ez.ctypes.data_as(ctypes.c_void_p),
              ptr[5], ctypes.c_size_t(4))
    _state["temp"] = dict(m=m, bounds=bounds, sy=sy, sz=sz)


def helm3d_temp(rhs, sol, bcy, bcz, alambda: float, alp: float = 1.0, ib: int = 0) -> None:
    """sol = helm3d_temp(rhs) on device tensors: rhs, sol complex128 (nmodes, m, m) [mode, z, y];
    bcy, bcz float64 (nmodes, 2, m) [mode, wall, z] and [mode, wall, y]."""
    m = rhs.shape[-1]
    _check_field(rhs, m)
    _check_field(sol, m)
    M = rhs.shape[0]
    for b in (bcy, bcz):
        if tuple(b.shape) != (M, 2, m) or not b.is_contiguous() or str(b.dtype) != "torch.float64":
            raise ValueError("bcy / bcz must be contiguous float64 (nmodes, 2, m)")
    _use_current_stream()
    _lib.call("duct_helm3d_temp_dev_", *[_lib.byref_ptr(t.data_ptr()) for t in (rhs, sol, bcy, bcz)],
              _lib.byref_int(m), _lib.byref_int(M), _lib.byref_int(ib), _lib.byref_double(alp),
              _lib.byref_double(alambda))
