"""Shen-Legendre Galerkin cross-check of the collocation solve (CPU only).

The Galerkin solver (oracle/shen.py) is pinned on its own by polynomial
manufactured solutions (exact in its space); it then agrees with the
collocation oracle to truncation error, and the disagreement shrinks with m,
as both discretisations converge to the same PDE solution.
"""
import numpy as np
import pytest

from oracle import polynomials as P
from oracle import shen


@pytest.mark.parametrize("m", [9, 17, 33])
@pytest.mark.parametrize("bound,lam", [("d", 1003.0), ("d", 2.0), ("n", 5.0), ("n", 0.0)])
def test_shen_exact_on_manufactured(m, bound, lam):
    rng = np.random.default_rng(3 * m)
    terms = P.separable_terms(m, rng, 3, "d" if bound == "d" else "n")
    u = P.evaluate(m, terms)
    f = P.helmholtz_rhs(m, terms, lam)          # lap(u) - lam u
    us = shen.shen_solve(m, -f, lam, bound)     # -lap(u) + lam u = -f
    d = us - u
    if bound == "n" and lam == 0.0:
        d -= d.mean()
    assert np.max(np.abs(d)) <= 1e-11 * np.max(np.abs(u))


def _smooth_rhs(m, decay, seed):
    rng = np.random.default_rng(seed)
    k = np.arange(m)
    g = rng.standard_normal((m, m)) * np.exp(-decay * (k[:, None] + k[None, :]))
    T = P.cheb_vander(m, m - 1)
    return T @ g @ T.T


def test_shen_vs_collocation_neumann_converges(oracle_mod):
    diffs = []
    for m in (17, 33, 65):
        ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
        f = _smooth_rhs(m, 0.8, m)
        uc = oracle_mod.solvediag(ob, f, "n", 3.0)
        us = shen.shen_solve(m, -f, 3.0, "n")
        diffs.append(np.max(np.abs(us - uc)) / np.max(np.abs(uc)))
    assert diffs[0] > diffs[1] > diffs[2]
    assert diffs[2] <= 1e-10


@pytest.mark.parametrize("lam", [1000.0, 1.0])
def test_shen_vs_collocation_dirichlet_converges(oracle_mod, lam):
    diffs = []
    for m in (17, 33, 65):
        ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
        f = _smooth_rhs(m, 0.8, m)
        F = f.copy()
        F[0, :] = F[-1, :] = F[:, 0] = F[:, -1] = 0.0   # collocation reads the interior only
        uc = oracle_mod.solvediag(ob, F, "d", lam)
        us = shen.shen_solve(m, -f, lam, "d")
        diffs.append(np.max(np.abs(us - uc)) / np.max(np.abs(uc)))
    assert diffs[0] > diffs[1] > diffs[2]
    assert diffs[2] <= 1e-4


def test_shen_vs_collocation_config0(oracle_mod):
    """The CPU reference case: 33^2, alpha = 1000, 16 modes of the seeded spectrum."""
    from duct_amd import workloads
    cfg = workloads.CPU_33
    m = cfg.m
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    f = workloads.rhs_field(cfg, 0)
    sol, _ = oracle_mod.helm3d(ob, f, "d", 1000.0)
    worst = 0.0
    for k in range(cfg.nmodes):
        xl = 1000.0 + k * k
        # Galerkin on the unzeroed Chebyshev series (the collocation solve reads f's interior only)
        N = m - 1
        rng = np.random.Generator(np.random.PCG64(workloads.seed_for(0, 0)))
        g1 = rng.standard_normal((cfg.nmodes, N + 1, N + 1))
        g2 = rng.standard_normal((cfg.nmodes, N + 1, N + 1))
        q, p = np.arange(N + 1)[:, None], np.arange(N + 1)[None, :]
        a = (g1[k] + 1j * g2[k]) * np.exp(-0.15 * (p + q))
        T = workloads.cheb_nodes_matrix(m, N)
        fk = T @ a @ T.T
        if k == 0:
            fk = fk.real
        us = shen.shen_solve(m, -np.real(fk), xl, "d") + 1j * shen.shen_solve(m, -np.imag(fk), xl, "d")
        worst = max(worst, np.max(np.abs(us - sol[k])) / np.max(np.abs(sol[k])))
    print("33^2x16: Shen-Legendre vs collocation max rel. difference %.2e" % worst)
    assert worst <= 5e-3
