"""Oracle restatement of the mixed-boundary (temperature) solve.

TEST INFRASTRUCTURE ONLY: imported by tests/ (never by the duct_amd package).
Follows, in numpy (LAPACK dgeev / getrf / getri through numpy, dense products):

  pseudomat_temp    src/pseudomat_temp.f90:2-162 (y) and :164-321 (z): CGL nodes, the CHQZ
                    derivative matrix scaled by 1/s (domain [-s, s]), D2 = D D, the interior
                    operator with the boundary eliminated per wall type (d-d, d-n, n-d, n-n),
                    its eigenvectors P, P^-1 and eigenvalues E
  mod_term          mod_term_FTy / mod_term_FTz, solvediag_temp.f90:136-247: the right-hand side
                    terms of the wall data bcy / bcz per wall type
  edge_values       edge_value_y / edge_value_z, solvediag_temp.f90:249-352
  solvediag_temp    solvediag_temp.f90:2-134: F + bcontv + bconth, F = P_y^-1 F P_z^-T,
                    ./ (E_y(i) + E_z(j) - xlambda), P_y F P_z^T, edges; the corners keep U's values
  helm3d_temp       helm3d_temp.f90:2-44 with int2d3d_temp.f90:1-40: per mode xlambda =
                    alambda + (alp k)^2, Re and Im solved with the same (real) wall data, Im not
                    solved at k = 0 (zero), the right-hand side not modified

Arrays: a plane U is [z, y] (C order) = the reference's U(my, mz) with y fastest; bcy is
[wall, z] (the reference's bcy(mz, 2)), bcz is [wall, y]; per mode [mode, wall, .].
my and mz may differ (helm3d_temp.f90:35-36 builds the y basis of size my and the z basis of
size mz): a plane is then [mz, my], bcy [2, mz] and bcz [2, my].
"""
from __future__ import annotations

import numpy as np


def cheb_D(m: int, s: float) -> np.ndarray:
    """pseudomat_temp.f90:60-94: the CHQZ matrix (textbook entries), scaled by 1/s."""
    N = m - 1
    y = -np.cos(np.pi * np.arange(m) / N)
    dd = np.ones(m)
    dd[0] = dd[-1] = 2.0
    D = np.zeros((m, m))
    D[0, 0] = -(2.0 * N * N + 1.0) / 6.0
    D[-1, -1] = -D[0, 0]
    for i in range(1, N):
        D[i, i] = -0.5 * y[i] / (1.0 - y[i] ** 2)
    for l in range(m):
        for j in range(m):
            if j != l:
                D[l, j] = dd[l] / dd[j] * (-1.0) ** (l + j) / (y[l] - y[j])
    return D / s


def pseudomat_temp(bound1: str, bound2: str, m: int, s: float):
    """Returns (D, D2, P, Pinv, E) of one direction (D, D2 m x m; P, Pinv, E interior)."""
    D = cheb_D(m, s)
    D2 = D @ D
    n = m - 1
    if bound1 == "n" and bound2 == "n":
        det = D[0, 0] * D[-1, -1] - D[0, -1] * D[-1, 0]
        i, j = np.meshgrid(np.arange(1, n), np.arange(1, n), indexing="ij")
        ter1 = D2[i, 0] * (D[0, -1] * D[-1, j] - D[-1, -1] * D[0, j]) / det
        ter3 = D2[i, -1] * (D[-1, 0] * D[0, j] - D[0, 0] * D[-1, j]) / det
        D2T = ter1 + D2[i, j] + ter3
    elif bound1 == "d" and bound2 == "d":
        D2T = D2[1:n, 1:n].copy()
    elif bound1 == "d" and bound2 == "n":
        D2T = D2[1:n, 1:n] - np.outer(D2[1:n, -1], D[-1, 1:n]) / D[-1, -1]
    elif bound1 == "n" and bound2 == "d":
        D2T = D2[1:n, 1:n] - np.outer(D2[1:n, 0], D[0, 1:n]) / D[0, 0]
    else:
        raise ValueError("bound must be 'd' or 'n'")
    E, P = np.linalg.eig(D2T)
    E, P = E.real, P.real                    # dgeev's real parts, as the reference keeps them
    return D, D2, P, np.linalg.inv(P), E


class TempBasis:
    """The two directions' bases of helm3d_temp (pseudomaty_temp / pseudomatz_temp)."""

    def __init__(self, m: int, bounds: str, sy: float = 1.0, sz: float = 1.0, mz: int | None = None):
        """m = my, the y points; mz the z points (default my)."""
        if len(bounds) != 4 or any(c not in "dn" for c in bounds):
            raise ValueError("bounds = 4 characters of 'd' / 'n': y=-1, y=+1, z=-1, z=+1")
        mz = m if mz is None else mz
        self.m, self.my, self.mz, self.bounds, self.sy, self.sz = m, m, mz, bounds, sy, sz
        self.Dy, self.D2y, self.Py, self.Pyi, self.Ey = pseudomat_temp(bounds[0], bounds[1], m, sy)   # pseudomaty_temp
        self.Dz, self.D2z, self.Pz, self.Pzi, self.Ez = pseudomat_temp(bounds[2], bounds[3], mz, sz)  # pseudomatz_temp


def wall_terms(b1: str, b2: str, D: np.ndarray, bc: np.ndarray):
    """The effective wall values (g1, g2) of mod_term_FT*: the interior correction is
    -D2(i,1) g1 - D2(i,m) g2 (solvediag_temp.f90:136-247, written in that common form)."""
    if b1 == "d" and b2 == "d":
        return bc[0], bc[1]
    if b1 == "d" and b2 == "n":
        return bc[0], (bc[1] - D[-1, 0] * bc[0]) / D[-1, -1]
    if b1 == "n" and b2 == "d":
        return (bc[0] - D[0, -1] * bc[1]) / D[0, 0], bc[1]
    det = D[0, 0] * D[-1, -1] - D[0, -1] * D[-1, 0]
    return (D[-1, -1] * bc[0] - D[0, -1] * bc[1]) / det, (D[0, 0] * bc[1] - D[-1, 0] * bc[0]) / det


def edge_line(b1: str, b2: str, D: np.ndarray, bc: np.ndarray, U: np.ndarray):
    """edge_value_* along the last axis of U ([line, point]) for the interior lines: returns
    (u at point 0, u at point m-1) per line."""
    a = U[:, 1:-1] @ D[0, 1:-1]
    b = U[:, 1:-1] @ D[-1, 1:-1]
    if b1 == "d" and b2 == "d":
        return bc[0], bc[1]
    if b1 == "d" and b2 == "n":
        u0 = bc[0]
        return u0, (bc[1] - D[-1, 0] * u0 - b) / D[-1, -1]
    if b1 == "n" and b2 == "d":
        um = bc[1]
        return (bc[0] - D[0, -1] * um - a) / D[0, 0], um
    det = D[0, 0] * D[-1, -1] - D[0, -1] * D[-1, 0]
    return ((D[-1, -1] * (bc[0] - a) - D[0, -1] * (bc[1] - b)) / det,
            (-D[-1, 0] * (bc[0] - a) + D[0, 0] * (bc[1] - b)) / det)


def solvediag_temp(tb: TempBasis, U: np.ndarray, xlambda: float, bcy: np.ndarray, bcz: np.ndarray):
    """One real plane U [z, y] (rhs in, solution out, corners kept); bcy [2, mz] over z, bcz [2, my]
    over y.  Returns the new plane."""
    if U.shape != (tb.mz, tb.my):
        raise ValueError("plane must be [mz, my]")
    by1, by2, bz1, bz2 = tb.bounds
    F = U[1:-1, 1:-1].copy()                                    # [z, y] interior
    gy1, gy2 = wall_terms(by1, by2, tb.Dy, bcy)                 # over z
    gz1, gz2 = wall_terms(bz1, bz2, tb.Dz, bcz)                 # over y
    # bcontv(i, j) = -D2by(i,1) gy1(j) - D2by(i,2) gy2(j);  bconth(i, j) = -D2bz(j,1) gz1(i) - D2bz(j,2) gz2(i)
    F += -(np.outer(gy1[1:-1], tb.D2y[1:-1, 0]) + np.outer(gy2[1:-1], tb.D2y[1:-1, -1]))
    F += -(np.outer(tb.D2z[1:-1, 0], gz1[1:-1]) + np.outer(tb.D2z[1:-1, -1], gz2[1:-1]))
    # F = Py^-1 F Pz^-T in [y, z] terms: here F is [z, y] = (F_yz)^T
    G = tb.Pzi @ F @ tb.Pyi.T                                  # [z-eigen, y-eigen]
    G = G / (tb.Ey[None, :] + tb.Ez[:, None] - xlambda)
    S = tb.Pz @ G @ tb.Py.T                                    # [z, y]
    out = U.copy()
    out[1:-1, 1:-1] = S
    # edge_value_y: U(1,j), U(my,j) for interior j (z); lines are z, points y
    e0, e1 = edge_line(by1, by2, tb.Dy, bcy[:, 1:-1], out[1:-1, :])
    out[1:-1, 0], out[1:-1, -1] = e0, e1
    # edge_value_z: U(i,1), U(i,mz) for interior i (y); lines are y, points z
    e0, e1 = edge_line(bz1, bz2, tb.Dz, bcz[:, 1:-1], out[:, 1:-1].T)
    out[0, 1:-1], out[-1, 1:-1] = e0, e1
    return out


def helm3d_temp(tb: TempBasis, rhs: np.ndarray, alambda: float, bcy: np.ndarray, bcz: np.ndarray,
                alp: float = 1.0, ib: int = 0) -> np.ndarray:
    """rhs [mode, z, y] complex; bcy [mode, 2, mz] (over z), bcz [mode, 2, my] (over y), real.
    Returns sol; rhs is not modified."""
    sol = np.zeros_like(rhs)
    for k in range(rhs.shape[0]):
        xlambda = alambda + (alp * (ib + k)) ** 2
        re = solvediag_temp(tb, rhs[k].real, xlambda, bcy[k], bcz[k])
        if ib + k == 0:
            sol[k] = re
        else:
            im = solvediag_temp(tb, rhs[k].imag, xlambda, bcy[k], bcz[k])
            sol[k] = re + 1j * im
    return sol
