"""Exact polynomial helpers used to pin the oracle (test infrastructure only).

A polynomial of degree <= m-1 in y and in z is reproduced exactly by the
collocation operators built on the m CGL nodes, so a field u sampled at the
nodes with f = u_yy + u_zz - xlambda*u computed by exact Chebyshev-series
differentiation (numpy.polynomial.chebyshev, independent of the DC matrix)
must be recovered by solvediag up to rounding.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import chebyshev as C

from .pseudomat import cgl_nodes


def cheb_vander(m: int, deg: int) -> np.ndarray:
    """T_p(y_j) for the reference's increasing nodes, shape (m, deg+1)."""
    return C.chebvander(cgl_nodes(m), deg)


def separable_terms(m: int, rng: np.random.Generator, nterms: int, bc: str, deg: int | None = None):
    """Random sums of products g(y) h(z) of Chebyshev series of degree <= m-1.
    bc='d': g(+-1)=h(+-1)=0;  bc='n': g'(+-1)=h'(+-1)=0;  bc='free': no condition.
    Returns a list of (cy, cz) Chebyshev coefficient pairs."""
    N = m - 1
    if deg is None:
        deg = N
    terms = []
    for _ in range(nterms):
        pair = []
        for _axis in range(2):
            if bc == "d":
                p = rng.standard_normal(deg - 1) * np.exp(-0.3 * np.arange(deg - 1))
                c = C.chebmul([1.0, 0.0, -1.0], p)            # (1-x^2)/... times p, deg <= deg
                c = C.chebmul([0.5, 0.0, -0.5], p)            # 1 - x^2 = 0.5 T0 - 0.5 T2
            elif bc == "n":
                p = rng.standard_normal(deg - 2) * np.exp(-0.3 * np.arange(deg - 2))
                d1 = C.chebmul([0.5, 0.0, -0.5], p)           # g' = (1-x^2) p
                c = C.chebint(d1)
                c[0] += rng.standard_normal()
            else:
                c = rng.standard_normal(deg + 1) * np.exp(-0.3 * np.arange(deg + 1))
            pair.append(np.asarray(c, dtype=np.float64))
        terms.append(tuple(pair))
    return terms


def evaluate(m: int, terms, dy: int = 0, dz: int = 0) -> np.ndarray:
    """Sum of g^(dy)(y) h^(dz)(z) at the nodes, returned as [z, y]."""
    x = cgl_nodes(m)
    out = np.zeros((m, m))
    for cy, cz in terms:
        gy = C.chebval(x, C.chebder(cy, dy) if dy else cy)
        hz = C.chebval(x, C.chebder(cz, dz) if dz else cz)
        out += np.outer(hz, gy)
    return out


def helmholtz_rhs(m: int, terms, xlambda: float) -> np.ndarray:
    """f = u_yy + u_zz - xlambda u at all nodes, [z, y]."""
    return evaluate(m, terms, 2, 0) + evaluate(m, terms, 0, 2) - xlambda * evaluate(m, terms)
