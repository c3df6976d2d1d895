"""Host-side one-time setup: the FP64 eigenbases the device solve consumes.

This is the Python host's counterpart of the reference's pseudomat
(src/pseudomat.f90:2-312): it builds the CGL collocation derivative matrix,
eliminates the boundary rows for the Dirichlet and Neumann problems, and
diagonalises the interior second-derivative operator once in FP64 with LAPACK
dgeev (through numpy.linalg.eig, as pseudomat calls dgeev) and dgetrf/dgetri
(numpy.linalg.inv).  Fortran callers keep their own pseudomat and hand the
same arrays to duct_setup_ (include/duct_hip.h).

The result is independent of eigenvector scaling and ordering except for the
Neumann gauge node, which is pinned the way the reference pins it
(first index of min |lambda|, pseudomat.f90:146-152).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class BoundaryBasis:
    bound: str           # 'd' or 'n'
    m: int
    P: np.ndarray        # (n, n) right eigenvectors, Fortran order
    Pinv: np.ndarray     # (n, n) inverse
    lam: np.ndarray      # (n,) eigenvalues
    op: np.ndarray       # (n, n) the interior operator that was diagonalised
    iofzero: int = 0     # 1-based eigen-index of the null mode (Neumann)
    wanted: float = 0.0
    AM1: np.ndarray | None = None


@dataclass
class Setup:
    m: int
    x: np.ndarray        # CGL nodes, increasing (pseudomat.f90:83-85)
    DC: np.ndarray       # (m, m) first-derivative collocation matrix
    D2: np.ndarray       # DC @ DC
    dirichlet: BoundaryBasis
    neumann: BoundaryBasis


def collocation_matrix(m: int):
    """Nodes x_j = -cos(pi j / N) and the CHQZ (2.4.31) derivative matrix with
    the reference's corner and diagonal entries (pseudomat.f90:72-107)."""
    N = m - 1
    j = np.arange(m)
    x = -np.cos(np.pi * j / N)
    cbar = np.ones(m)
    cbar[0] = cbar[-1] = 2.0
    sign = np.where((j[:, None] + j[None, :]) % 2 == 0, 1.0, -1.0)
    diff = x[:, None] - x[None, :]
    np.fill_diagonal(diff, 1.0)
    DC = (cbar[:, None] / cbar[None, :]) * sign / diff
    inner = np.arange(1, N)
    DC[inner, inner] = -0.5 * x[inner] / (1.0 - x[inner] ** 2)
    DC[0, 0] = -(2.0 * N * N + 1.0) / 6.0
    DC[N, N] = -DC[0, 0]
    return x, DC


def _eig(op: np.ndarray):
    w, v = np.linalg.eig(op)
    if np.iscomplexobj(w) and np.max(np.abs(np.imag(w))) > 1e-8 * np.max(np.abs(w)):
        raise ValueError("interior operator has complex eigenvalues")
    return np.real(w).copy(), np.asfortranarray(np.real(v))


def setup(m: int) -> Setup:
    """Both bases helm3d builds on its first call (helm3d.f90:43-44)."""
    if m < 4:
        raise ValueError("m must be >= 4")
    x, DC = collocation_matrix(m)
    D2 = DC @ DC
    N = m - 1
    s = slice(1, N)

    # Dirichlet: interior block of D2 (pseudomat.f90:232-240)
    opd = np.ascontiguousarray(D2[s, s])
    lam_d, Pd = _eig(opd)
    dirichlet = BoundaryBasis("d", m, Pd, np.asfortranarray(np.linalg.inv(Pd)), lam_d, opd)

    # Neumann: boundary values eliminated through DC rows 1 and m (:114-128)
    # Evaluated in the reference's operation order, (D2(i,1)*c)/Det etc., so
    # that the matrix handed to dgeev is the same to the last bit and the
    # eigenvalue order (hence the gauge node iofzero) agrees with it.
    det = DC[0, 0] * DC[N, N] - DC[0, N] * DC[N, 0]
    c1 = DC[0, N] * DC[N, s] - DC[N, N] * DC[0, s]              # u_1 = (c1 . u_int)/det
    cm = DC[N, 0] * DC[0, s] - DC[0, 0] * DC[N, s]              # u_m = (cm . u_int)/det
    ter1 = (D2[s, 0][:, None] * c1[None, :]) / det
    ter3 = (D2[s, N][:, None] * cm[None, :]) / det
    opn = np.ascontiguousarray((ter1 + D2[s, s]) + ter3)
    lam_n, Pn = _eig(opn)
    iofzero = int(np.argmin(np.abs(lam_n))) + 1
    AM = np.array([
        [2.0 * DC[0, 0], DC[0, N], 0.0, DC[0, N]],
        [-DC[N, 0], DC[0, 0] - DC[N, N], DC[0, N], 0.0],
        [0.0, DC[N, 0], 2.0 * DC[N, N], DC[N, 0]],
        [-DC[N, 0], 0.0, DC[0, N], DC[0, 0] - DC[N, N]],
    ])
    neumann = BoundaryBasis("n", m, Pn, np.asfortranarray(np.linalg.inv(Pn)), lam_n, opn,
                            iofzero=iofzero, wanted=0.0, AM1=np.asfortranarray(np.linalg.inv(AM)))
    return Setup(m, x, np.asfortranarray(DC), D2, dirichlet, neumann)
