"""Oracle restatement of the reference's one-time eigenbasis setup.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the duct_amd package.

Follows src/pseudomat.f90 of ando2800/duct loop by loop (1-based Fortran
indices are shifted to 0-based).  LAPACK dgeev / dgetrf+dgetri are reached
through numpy.linalg.eig / numpy.linalg.inv (numpy's bundled LAPACK), which is
the same LAPACK interface the reference calls.
"""
from __future__ import annotations

import numpy as np


class Basis:
    """Holds the module state that pseudomat fills (mod.f90:277-343)."""


def cgl_nodes(m: int) -> np.ndarray:
    """pseudomat.f90:83-85: x_j = -cos(pi (j-1)/(m-1)), increasing."""
    n = m - 1
    x = np.zeros(m)
    for j in range(m):
        x[j] = -np.cos(np.pi * float(j) / float(n))
    return x


def colloc_deriv(m: int) -> np.ndarray:
    """pseudomat.f90:72-107 (CHQZ p.69): first-derivative collocation matrix DC."""
    n = m - 1
    x = cgl_nodes(m)
    dd = np.ones(m)
    dd[0] = 2.0
    dd[m - 1] = 2.0
    DC = np.zeros((m, m))
    DC[0, 0] = -(2.0 * float(n) ** 2 + 1.0) / 6.0
    DC[m - 1, m - 1] = -DC[0, 0]
    for i in range(1, n):
        DC[i, i] = -0.5 * x[i] / (1.0 - x[i] ** 2)
    for l in range(m):
        for j in range(l):
            coef = dd[l] / dd[j]
            DC[l, j] = coef * (-1.0) ** (l + j) / (x[l] - x[j])
        for j in range(l + 1, m):
            coef = dd[l] / dd[j]
            DC[l, j] = coef * (-1.0) ** (l + j) / (x[l] - x[j])
    return DC


def _pad(a: np.ndarray) -> np.ndarray:
    """(m-2)^2 -> (m-1)^2 with a zero last row/column (pseudomat.f90:45-57)."""
    k = a.shape[0]
    out = np.zeros((k + 1, k + 1))
    out[:k, :k] = a
    return out


def pseudomat(m: int) -> Basis:
    """Both calls helm3d makes (helm3d.f90:43-44): pseudomat('d') and ('n')."""
    n = m - 1
    mm2 = m - 2
    b = Basis()
    b.m = m
    b.x = cgl_nodes(m)
    b.DC = colloc_deriv(m)
    DC = b.DC
    D2 = DC @ DC                                                    # :110
    b.D2 = D2

    # ---- Neumann (:114-221)
    Det = DC[0, 0] * DC[m - 1, m - 1] - DC[0, m - 1] * DC[m - 1, 0]
    D2B = np.zeros((mm2, mm2))
    for ii, i in enumerate(range(1, n)):
        for jj, j in enumerate(range(1, n)):
            ter1 = D2[i, 0] * (DC[0, m - 1] * DC[m - 1, j] - DC[m - 1, m - 1] * DC[0, j]) / Det
            ter3 = D2[i, m - 1] * (DC[m - 1, 0] * DC[0, j] - DC[0, 0] * DC[m - 1, j]) / Det
            D2B[ii, jj] = ter1 + D2[i, j] + ter3
    b.D2B_N = D2B.copy()
    w, v = np.linalg.eig(D2B)                                       # dgeev('n','v') :131
    b.EN_imag = np.asarray(w).imag.copy()
    b.EN = np.asarray(w).real.copy()                                # wr; wi discarded
    b.PN = np.asarray(v).real.copy()
    b.EN2 = np.zeros((mm2 + 1, mm2 + 1))
    for j in range(mm2):
        for i in range(mm2):
            b.EN2[i, j] = b.EN[i] + b.EN[j]
    b.PNM1 = np.linalg.inv(b.PN)                                    # dgetrf/dgetri :140-142
    eigmin = 1.0e6                                                  # :146-152
    iofzero = 0
    for i in range(mm2):
        eig = min(eigmin, abs(b.EN[i]))
        if eig != eigmin:
            iofzero = i + 1
            eigmin = eig
    b.iofzero = iofzero                                             # 1-based, as in Fortran
    b.wanted = 0.0                                                  # :156
    AM = np.zeros((4, 4))                                           # :158-176
    AM[0, :] = [2.0 * DC[0, 0], DC[0, m - 1], 0.0, DC[0, m - 1]]
    AM[1, :] = [-DC[m - 1, 0], DC[0, 0] - DC[m - 1, m - 1], DC[0, m - 1], 0.0]
    AM[2, :] = [0.0, DC[m - 1, 0], 2.0 * DC[m - 1, m - 1], DC[m - 1, 0]]
    AM[3, :] = [-DC[m - 1, 0], 0.0, DC[0, m - 1], DC[0, 0] - DC[m - 1, m - 1]]
    b.AM = AM
    b.AM1 = np.linalg.inv(AM)                                       # :177-178
    b.PNT = _pad(b.PN.T)                                            # :181-186
    b.PNM1T = _pad(b.PNM1.T)

    # ---- Dirichlet (:223-263)
    b.D2bx = np.zeros((m, 2))
    b.D2by = np.zeros((m, 2))
    for i in range(m):
        b.D2bx[i, 0] = D2[i, 0]
        b.D2bx[i, 1] = D2[i, m - 1]
        b.D2by[i, 0] = D2[i, 0]
        b.D2by[i, 1] = D2[i, m - 1]
    D2B = np.zeros((mm2, mm2))
    for ii, i in enumerate(range(1, n)):
        for jj, j in enumerate(range(1, n)):
            D2B[ii, jj] = D2[i, j]
    b.D2B_D = D2B.copy()
    w, v = np.linalg.eig(D2B)                                       # :244
    b.ED_imag = np.asarray(w).imag.copy()
    b.ED = np.asarray(w).real.copy()
    b.PD = np.asarray(v).real.copy()
    b.ED2 = np.zeros((mm2 + 1, mm2 + 1))
    for j in range(mm2):
        for i in range(mm2):
            b.ED2[i, j] = b.ED[i] + b.ED[j]
    b.PDM1 = np.linalg.inv(b.PD)                                    # :253-255
    b.PDT = _pad(b.PD.T)                                            # :258-263
    b.PDM1T = _pad(b.PDM1.T)
    return b
