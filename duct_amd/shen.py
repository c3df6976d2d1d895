"""Shen-Legendre spectral-Galerkin Helmholtz solve on the device (the -DHELM_SHEN path).

helm3d_shen (src/helm_shen.f90:2-85) -> solveshen (src/solveshen.f90:5-94) solves
    -lap(u) + xlambda u = f,   xlambda = alambda + (alp k)^2
per Fourier mode in the Shen basis phi_k = L_k + b_k L_{k+2} (b_k = -1 Dirichlet,
-k(k+1)/((k+2)(k+3)) Neumann; makeMS, preshen.f90:332-396), with the mass matrix
M (diagonal + second off-diagonal) and the diagonal stiffness S of that basis.  The
Galerkin system S U M + M U S + xlambda M U M = F (F = (f, phi_p phi_q)) is
diagonalised along y by the generalised eigenproblem M x = lam S x (precomeigen,
preshen.f90:398-535; with x^T S x = 1 the reference's inv(P) inv(S) is P^T) and
solved along z, for every y-eigenmode lam_p, by the symmetric tridiagonal systems
    (lam_p (xlambda M + S) + M) v = h     (one per parity of the z index; solV,
                                            solveshen.f90:335-513)
The Neumann y-constant mode (b_0 = 0, S_00 = 0) decouples: (S + xlambda M) v = F_0 / M_00;
at xlambda = 0 its z-constant coefficient is pinned to zero (zero mean).

On the device (duct_helm3d_shen_dev_): F = A_y f A_z^T (A_y = P^T W, A_z = W, W the
nodal-value -> Galerkin-inner-product map of Gauss-Legendre quadrature, exact for the
degree-N interpolant), the batched tridiagonal solves, then u = B_y V B_z^T (B_y = R P,
B_z = R, R the Shen-coefficient -> nodal-value map) - four whole-plane products and one
batched solve.  The matrices are built here on the host in FP64 (scipy's LAPACK for the
generalised eigenproblem), the reference's Chebyshev -> Legendre -> Shen sweeps and
their inverses folded into W and R.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.linalg
import scipy.special
from numpy.polynomial import legendre as Lg

from . import _lib
from .solver import _check_field, _use_current_stream


def shen_b(N: int, bound: str) -> np.ndarray:
    """b_k, k = 0 .. N-2 (makeMS, preshen.f90:360-370)."""
    k = np.arange(N - 1, dtype=np.float64)
    if bound == "d":
        return -np.ones(N - 1)
    if bound == "n":
        return -k * (k + 1.0) / ((k + 2.0) * (k + 3.0))
    raise ValueError("bound must be 'd' or 'n'")


def mass_stiffness(N: int, bound: str):
    """Diagonal of M, M_{k,k+2}, diagonal of S (preshen.f90:375-392), all of length N-1."""
    b = shen_b(N, bound)
    k = np.arange(N - 1, dtype=np.float64)
    md = 2.0 / (2 * k + 1) + 2.0 * b * b / (2 * k + 5)
    mu = b * 2.0 / (2 * k + 5)
    mu[-2:] = 0.0                                      # M_{k,k+2} exists for k <= N-4
    sd = -b * (4 * k + 6)
    return md, mu, sd


def _cgl(m: int) -> np.ndarray:
    return -np.cos(np.pi * np.arange(m) / (m - 1))


def _phi(x: np.ndarray, N: int, bound: str) -> np.ndarray:
    """phi_k(x), k = 0 .. N-2: [point, k]."""
    L = Lg.legvander(x, N)                             # L_i(x), i <= N
    b = shen_b(N, bound)
    return L[:, :N - 1] + L[:, 2:N + 1] * b[None, :]


def _interp_cgl(m: int, xq: np.ndarray) -> np.ndarray:
    """Barycentric interpolation matrix from the m CGL nodes to the points xq."""
    x = _cgl(m)
    w = (-1.0) ** np.arange(m)
    w[0] *= 0.5
    w[-1] *= 0.5
    d = xq[:, None] - x[None, :]
    exact = np.abs(d) < 1e-300
    d[exact] = 1.0
    T = w[None, :] / d
    T /= T.sum(1, keepdims=True)
    rows = exact.any(1)
    T[rows] = exact[rows].astype(float)
    return T


def shen_matrices(m: int, bound: str):
    """(A_y, A_z, B_y, B_z, md, mu, sd, lam) for duct_setup_shen_."""
    N = m - 1
    xg, wg = scipy.special.roots_legendre(N + 1)        # exact for degree <= 2N + 1
    W = _phi(xg, N, bound).T @ (wg[:, None] * _interp_cgl(m, xg))   # (N-1, m): f -> (f, phi_k)
    R = _phi(_cgl(m), N, bound)                                      # (m, N-1): U -> u(x_j)
    md, mu, sd = mass_stiffness(N, bound)
    n = N - 1
    M = np.diag(md) + np.diag(mu[:n - 2], 2) + np.diag(mu[:n - 2], -2)
    S = np.diag(sd)
    if bound == "d":
        lam, P = scipy.linalg.eigh(M, S)                # M x = lam S x, x^T S x = 1
        C = P.T                                         # inv(P) inv(S)
    else:
        lam1, P1 = scipy.linalg.eigh(M[1:, 1:], S[1:, 1:])
        P = np.zeros((n, n))
        P[0, 0] = 1.0
        P[1:, 1:] = P1
        lam = np.concatenate([[np.nan], lam1])          # row 0: the decoupled y-constant mode
        C = np.zeros((n, n))
        C[0, 0] = 1.0 / md[0]
        C[1:, 1:] = P1.T
    return C @ W, W, R @ P, R, md, mu, sd, lam


_state = {"shen": None}


def setup_shen(m: int, bound: str) -> None:
    """preshen(bound) (preshen.f90:2-330): build the matrices and upload them (duct_setup_shen_)."""
    Ay, Az, By, Bz, md, mu, sd, lam = shen_matrices(m, bound)
    lam = np.where(np.isnan(lam), 0.0, lam)
    mats = [np.asfortranarray(x) for x in (Ay, Az, By, Bz)]
    vecs = [np.ascontiguousarray(x) for x in (md, mu, sd, lam)]
    ptr = [x.ctypes.data_as(ctypes.c_void_p) for x in mats + vecs]
    _lib.call("duct_setup_shen_", bound.encode(), _lib.byref_int(m), *ptr, ctypes.c_size_t(1))
    _state["shen"] = dict(m=m, bound=bound)


def helm3d_shen(rhs, sol, bound: str, alambda: float, alp: float = 1.0, ib: int = 0) -> None:
    """sol = the Shen-Galerkin solution of -lap(u) + xlambda u = rhs on device tensors
    (complex128 (nmodes, m, m) [mode, z, y], nodal values at the CGL nodes; rhs kept).
    The reference's callers pass -rhs of the collocation solve (main1.f90:613)."""
    m = rhs.shape[-1]
    _check_field(rhs, m)
    _check_field(sol, m)
    _use_current_stream()
    _lib.call("duct_helm3d_shen_dev_", _lib.byref_ptr(rhs.data_ptr()), _lib.byref_ptr(sol.data_ptr()),
              _lib.byref_int(m), _lib.byref_int(rhs.shape[0]), _lib.byref_int(ib), _lib.byref_double(alp),
              _lib.byref_double(alambda), bound.encode(), ctypes.c_size_t(1))
