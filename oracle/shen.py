"""Shen-Legendre spectral-Galerkin solver of -lap(u) + lam u = g on [-1,1]^2.

TEST INFRASTRUCTURE ONLY (an independent cross-check of the collocation
solve, like the reference's -DHELM_SHEN path: helm_shen.f90:2-85 ->
solveshen.f90:5-94, 335-513, preshen.f90:332-535, legendre.f90).

Textbook method (J. Shen, SIAM J. Sci. Comput. 15 (1994) 1489-1505):
  Dirichlet basis  phi_k = L_k - L_{k+2}
  Neumann basis    phi_k = L_k - k(k+1)/((k+2)(k+3)) L_{k+2}
  k = 0 .. N-2 (N+1 = m CGL points), mass M = (phi_j, phi_k), stiffness
  S = (phi_j', phi_k'); the weak form S U M + M U S + lam M U M = G with
  G = (g, phi_p phi_q) is diagonalised by the generalised eigenproblem
  S X = M X Lambda (X^T M X = I):  U = X ((X^T G X) ./ (Lam_i + Lam_j + lam)) X^T.
The reference sends -rhs to this solver because the collocation solve is of
lap(u) - xlambda u = f (main1.f90:613); shen_solve(m, -f, xlambda) is the
counterpart of solvediag on the same data.  g is represented by its Chebyshev
interpolant on the CGL nodes (helm_shen.f90:53-59 transforms it to Chebyshev
coefficients), and all inner products use Gauss-Legendre quadrature that is
exact for those polynomials.  The two discretisations agree to truncation
error, and exactly when u lies in the polynomial space of both.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg
from numpy.polynomial import chebyshev as C
from numpy.polynomial import legendre as Lg

from .pseudomat import cgl_nodes


def shen_basis(N: int, bound: str):
    """Legendre coefficient matrix B (N+1, N-1): phi_k = sum_i B[i, k] L_i."""
    B = np.zeros((N + 1, N - 1))
    for k in range(N - 1):
        B[k, k] = 1.0
        if bound == "d":
            B[k + 2, k] = -1.0
        elif bound == "n":
            B[k + 2, k] = -k * (k + 1.0) / ((k + 2.0) * (k + 3.0))
        else:
            raise ValueError("bound must be 'd' or 'n'")
    return B


def _eval_basis(B, x, der=0):
    out = np.zeros((len(x), B.shape[1]))
    for k in range(B.shape[1]):
        c = B[:, k]
        out[:, k] = Lg.legval(x, Lg.legder(c, der) if der else c)
    return out


def shen_solve(m: int, g_nodes: np.ndarray, lam: float, bound: str) -> np.ndarray:
    """Solve -lap(u) + lam u = g; g and the result are real (m, m) arrays [z, y]
    of values at the CGL nodes.  For bound 'n' with lam = 0 the mean of the
    constant mode is set to zero (the solution is defined up to a constant)."""
    N = m - 1
    x = cgl_nodes(m)
    B = shen_basis(N, bound)
    q, w = Lg.leggauss(N + 2)                         # exact to degree 2N+3
    Pq, dPq = _eval_basis(B, q), _eval_basis(B, q, 1)
    M = Pq.T @ (w[:, None] * Pq)
    S = dPq.T @ (w[:, None] * dPq)
    # Chebyshev interpolant of g (exact: square Vandermonde on the CGL nodes)
    V = C.chebvander(x, N)
    coef = np.linalg.solve(V, np.linalg.solve(V, g_nodes.T).T).T   # coef[p (y), q (z)]: V c V^T = g^T
    Tq = C.chebvander(q, N)
    gq = Tq @ coef @ Tq.T                              # g at quadrature points [y, z]
    G = Pq.T @ (w[:, None] * gq * w[None, :]) @ Pq     # (g, phi_p(y) phi_q(z))
    lamS, X = scipy.linalg.eigh(S, M)
    Gt = X.T @ G @ X
    den = lamS[:, None] + lamS[None, :] + lam
    if bound == "n" and abs(lam) <= 1e-10:
        i0 = int(np.argmin(np.abs(lamS)))
        den[i0, i0] = np.inf
    U = X @ (Gt / den) @ X.T                           # Shen coefficients [y-mode, z-mode]
    Pn = _eval_basis(B, x)
    u = Pn @ U @ Pn.T                                  # [y, z] at the CGL nodes
    return u.T.copy()                                  # [z, y]
