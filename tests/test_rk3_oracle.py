"""Pins of the Runge-Kutta substep oracle (oracle/rk3.py, the operators in oracle/spectral.py)
on the CPU:

- the low-storage coefficients (mod.f90:51-58) satisfy alpha_k + beta_k = gamma_k + zeta_k and
  sum (alpha_k + beta_k) = 1, the consistency conditions of the scheme;
- laplacian3d / grad3d / div3d (operators.F:278-501), restated through the reference's
  transform-recurrence route, against exact differentiation of polynomial fields (numpy
  chebder), aspect ratio included;
- the laminar duct profile, -(d_yy + d_zz) U = dpval with the reference's own dpval table
  (operators.F:2494-2507), is a steady state of the full substep with dpdx = get_dpdx: every
  term the substep adds must cancel exactly (a wrong sign or coefficient in any of items
  (i)-(xi) breaks it);
- the projection (items ix-xi) removes the interior divergence of the predicted velocity.
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from oracle import rk3 as R
from oracle import spectral as S


def test_rk_coefficients_are_consistent():
    for a, b, g, r in zip(R.AKKC, R.BKKC, R.GKKC, R.RKKC):
        assert abs((a + b) - (g + r)) < 1e-15
    assert abs(sum(a + b for a, b in zip(R.AKKC, R.BKKC)) - 1.0) < 1e-15
    assert R.get_dpdx(1.0, 0.01) == pytest.approx(-0.0339344926982230, rel=1e-15)
    with pytest.raises(ValueError):
        R.get_dpdx(1.2, 0.01)


def _poly_field(m, M, deg, rng):
    """Field [mode, z, y] = sum_{q,p <= deg} A[k, q, p] T_q(z) T_p(y) at the CGL nodes, and A."""
    x = -np.cos(np.pi * np.arange(m) / (m - 1))
    A = rng.standard_normal((M, deg + 1, deg + 1)) + 1j * rng.standard_normal((M, deg + 1, deg + 1))
    T = np.stack([C.chebval(x, np.eye(deg + 1)[p]) for p in range(deg + 1)], axis=1)   # [node, p]
    return np.einsum("zq,kqp,yp->kzy", T, A, T), A, T


def _eval(A, T):
    n = A.shape[-1]
    Tn = T[:, :n]
    return np.einsum("zq,kqp,yp->kzy", T[:, :A.shape[1]], A, Tn)


@pytest.mark.parametrize("m,aspect", [(17, 1.0), (24, 1.5)])
def test_operators_differentiate_polynomials_exactly(m, aspect):
    rng = np.random.default_rng(m)
    M, deg, alp = 3, m - 3, 1.3
    f, A, T = _poly_field(m, M, deg, rng)
    k = np.arange(M)[:, None, None]
    fy = _eval(C.chebder(A, 1, axis=2), T)
    fz = _eval(C.chebder(A, 1, axis=1), T) / aspect
    fyy = _eval(C.chebder(A, 2, axis=2), T)
    fzz = _eval(C.chebder(A, 2, axis=1), T) / aspect ** 2
    scale = np.abs(fyy).max() + np.abs(fzz).max()
    lap = S.laplacian3d(f, 0.7, alp, aspect)
    want = 0.7 * (-(alp * k) ** 2 * f + fyy + fzz)
    assert np.abs(lap - want).max() <= 1e-12 * scale
    gx, gy, gz = S.grad3d(f, -2.0, alp, aspect)
    assert np.abs(gx - (-2.0) * 1j * alp * k * f).max() <= 1e-14 * np.abs(f).max() * alp * M
    assert np.abs(gy + 2.0 * fy).max() <= 1e-12 * np.abs(fy).max()
    assert np.abs(gz + 2.0 * fz).max() <= 1e-12 * np.abs(fz).max()
    g2, _, _ = _poly_field(m, M, deg, rng)
    g3, _, _ = _poly_field(m, M, deg, rng)
    d = S.div3d(f, g2, g3, 0.5, alp, aspect)
    want = 0.5 * (1j * alp * k * f + S.grad3d(g2, 1.0, alp, aspect)[1] + S.grad3d(g3, 1.0, alp, aspect)[2])
    assert np.abs(d - want).max() <= 1e-13 * np.abs(want).max()


def _helm(oracle_mod, m, alp):
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))

    def helm3d(rhs, bound, alam):
        return oracle_mod.helm3d(ob, rhs, bound, alam, alp=alp)[0]
    return helm3d


def laminar_state(helm3d, m, M):
    """U [mode 0] solving (d_yy + d_zz) U = -dpval(1), zero on the walls."""
    rhs = np.zeros((M, m, m), dtype=np.complex128)
    rhs[0] = -R.DPVAL[0]
    return helm3d(rhs, "d", 0.0)


def test_laminar_duct_is_a_steady_state_of_the_substep(oracle_mod):
    m, M, mgalx, fnu, dt, alp = 17, 16, 48, 0.01, 0.05, 1.0
    helm3d = _helm(oracle_mod, m, alp)
    U = laminar_state(helm3d, m, M)
    assert abs(U.real.max() - 1.0) < 2e-3            # max U = 1 of the continuous problem
    u, v, w, p = U.copy(), np.zeros_like(U), np.zeros_like(U), np.zeros_like(U)
    adv = (np.zeros_like(U),) * 3
    for krk in (1, 2, 3):
        u, v, w, p, adv, phi = R.rk3_substep(u, v, w, p, adv, krk, fnu, dt, mgalx, helm3d, alp)
    assert np.abs(u - U).max() <= 1e-12
    assert max(np.abs(v).max(), np.abs(w).max(), np.abs(p).max()) <= 1e-12
    # the driving term is what holds it: without dpdx the flow decays
    u2, *_ = R.rk3_substep(U.copy(), v, w, p, adv, 1, fnu, dt, mgalx, helm3d, alp, dpdx=0.0)
    assert np.abs(u2 - U).max() > 1e-4


def test_projection_removes_interior_divergence(oracle_mod):
    from duct_amd import workloads
    m, M, mgalx, fnu, dt, alp = 17, 16, 48, 0.01, 0.05, 1.0
    helm3d = _helm(oracle_mod, m, alp)
    cfg = workloads.STEP_129
    u, v, w = workloads.duct_velocity(cfg, m=m, nmodes=M)
    p = np.zeros_like(u)
    adv = (np.zeros_like(u),) * 3
    un, vn, wn, pn, adv, phi = R.rk3_substep(u, v, w, p, adv, 1, fnu, dt, mgalx, helm3d, alp)
    su = -dt * (R.AKKC[0] + R.BKKC[0])
    gx, gy, gz = S.grad3d(phi, su, alp)
    div_star = S.div3d(un - gx, vn - gy, wn - gz, 1.0, alp)          # the predicted u*
    div_new = S.div3d(un, vn, wn, 1.0, alp)
    ratio = np.abs(div_new[:, 1:-1, 1:-1]).max() / np.abs(div_star[:, 1:-1, 1:-1]).max()
    print("interior divergence after / before the projection: %.2e" % ratio)
    assert ratio <= 1e-9
