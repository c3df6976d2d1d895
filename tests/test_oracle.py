"""Pins of the CPU oracle (oracle/) - CPU only.

The reference holds no tests, fixtures or golden vectors for this path; its
only known answer is the laminar square-duct table (operators.F:2496-2507).
Every oracle function is therefore pinned by mathematics that a plausible
mistake (dropped term, wrong sign or index, transposed operand) would break:

  function                      pinned by
  pseudomat: nodes, DC          closed form: DC T_k(x_j) = T_k'(x_j) for k <= N;
                                anti-centrosymmetry; zero row sums
  pseudomat: eigenbases         A P = P diag(lam) to 1e-12 ||A||; real spectrum;
                                Neumann: one null eigenvalue, constant eigenvector
  solvediag 'd' (homog. BC)     polynomial manufactured solutions (exact series
                                differentiation, independent of DC) to 1e-12;
                                dpval known answer: max U = 1 (operators.F:2486-2498)
  solvediag 'd' BC correction   manufactured solutions with NONZERO boundary data
  solvediag 'n', edges, corners manufactured solutions with g'(+-1) = 0, boundary
                                values compared too
  solvediag 'n' gauge (k=0)     manufactured solution modulo a constant, and
                                U(iofzero+1, iofzero+1) = wanted = 0
  int2d3d / helm3d              per-mode xlambda = alambda + (alp k)^2, Im skipped
                                at k=0, boundary zeroing of rhs, complex modes
  bad bound                     raises (reference stops, solvediag.f90:324-326)
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from oracle import polynomials as P

SIZES = [5, 9, 17, 33]


@pytest.fixture(scope="module")
def bases(oracle_mod):
    return {m: oracle_mod.OracleBasis(oracle_mod.pseudomat(m)) for m in SIZES + [65]}


def _zero_boundary(a):
    a = a.copy()
    a[0, :] = 0
    a[-1, :] = 0
    a[:, 0] = 0
    a[:, -1] = 0
    return a


@pytest.mark.parametrize("m", SIZES)
def test_dc_exact_on_chebyshev(oracle_mod, m):
    x = oracle_mod.cgl_nodes(m)
    DC = oracle_mod.colloc_deriv(m)
    N = m - 1
    for k in range(N + 1):
        c = np.zeros(k + 1)
        c[k] = 1.0
        tk = C.chebval(x, c)
        dtk = C.chebval(x, C.chebder(c))
        assert np.max(np.abs(DC @ tk - dtk)) <= 1e-12 * max(1.0, N * N) * max(1.0, np.max(np.abs(dtk)))
    assert np.allclose(DC.sum(axis=1), 0.0, atol=1e-10 * N * N)
    assert np.allclose(DC[::-1, ::-1], -DC, atol=1e-12 * N * N)
    assert np.all(np.diff(x) > 0) and x[0] == -1.0 and abs(x[-1] - 1.0) < 1e-15


@pytest.mark.parametrize("m", SIZES + [65])
def test_eigenbases(bases, m):
    b = bases[m].b
    for A, lam, Pm, Pinv, imag in ((b.D2B_D, b.ED, b.PD, b.PDM1, b.ED_imag), (b.D2B_N, b.EN, b.PN, b.PNM1, b.EN_imag)):
        scale = np.linalg.norm(A, 2) * np.linalg.norm(Pm, 2)
        assert np.linalg.norm(A @ Pm - Pm * lam[None, :], 2) <= 1e-12 * scale
        assert np.max(np.abs(imag)) == 0.0
        assert np.allclose(Pinv @ Pm, np.eye(m - 2), atol=1e-11)
    # Dirichlet: all negative, smallest magnitude -> -(pi/2)^2
    assert np.all(b.ED < 0)
    if m >= 17:
        assert abs(np.max(b.ED) + (np.pi / 2) ** 2) < 1e-8
    # Neumann: exactly one null eigenvalue; its eigenvector is constant
    small = np.abs(b.EN) < 1e-8 * np.max(np.abs(b.EN))
    assert small.sum() == 1
    io = b.iofzero - 1
    assert small[io]
    v = b.PN[:, io]
    assert np.max(np.abs(v - v.mean())) < 1e-10 * np.max(np.abs(v))
    # padded transposes
    assert np.array_equal(b.PDT[: m - 2, : m - 2], b.PD.T) and not b.PDT[-1].any() and not b.PDT[:, -1].any()
    assert np.array_equal(b.PNM1T[: m - 2, : m - 2], b.PNM1.T)


@pytest.mark.parametrize("m", SIZES)
@pytest.mark.parametrize("xl", [0.0, 3.7, 1000.0 + 25.0])
def test_dirichlet_manufactured_homogeneous(oracle_mod, bases, m, xl):
    rng = np.random.default_rng(100 + m)
    terms = P.separable_terms(m, rng, 3, "d")
    u = P.evaluate(m, terms)
    f = _zero_boundary(P.helmholtz_rhs(m, terms, xl))
    sol = oracle_mod.solvediag(bases[m], f, "d", xl)
    assert np.max(np.abs(sol - u)) <= 1e-11 * np.max(np.abs(u)) * max(1, m / 8)
    assert not sol[0].any() and not sol[-1].any() and not sol[:, 0].any() and not sol[:, -1].any()


@pytest.mark.parametrize("m", SIZES)
def test_dirichlet_manufactured_with_boundary_data(oracle_mod, bases, m):
    """Nonzero wall data exercises the correction solvediag.f90:37-45."""
    rng = np.random.default_rng(7 + m)
    terms = P.separable_terms(m, rng, 3, "free")
    xl = 12.5
    u = P.evaluate(m, terms)
    U = P.helmholtz_rhs(m, terms, xl)
    U[0, :], U[-1, :], U[:, 0], U[:, -1] = u[0, :], u[-1, :], u[:, 0], u[:, -1]
    sol = oracle_mod.solvediag(bases[m], U, "d", xl)
    assert np.max(np.abs(sol - u)) <= 1e-11 * np.max(np.abs(u)) * max(1, m / 8)


@pytest.mark.parametrize("m", SIZES)
@pytest.mark.parametrize("xl", [2.0, 1.0e-3, 1016.0])
def test_neumann_manufactured(oracle_mod, bases, m, xl):
    rng = np.random.default_rng(31 + m)
    terms = P.separable_terms(m, rng, 3, "n")
    u = P.evaluate(m, terms)
    f = P.helmholtz_rhs(m, terms, xl)
    f_noisy_bdry = f.copy()
    f_noisy_bdry[0, :] = 1e3   # boundary of the input is ignored for 'n'
    sol = oracle_mod.solvediag(bases[m], f_noisy_bdry, "n", xl)
    tol = 1e-10 if xl < 1e-2 else 1e-11
    assert np.max(np.abs(sol - u)) <= tol * np.max(np.abs(u)) * max(1, m / 8)   # edges + corners too


@pytest.mark.parametrize("m", SIZES + [65])
def test_neumann_gauge(oracle_mod, bases, m):
    b = bases[m].b
    rng = np.random.default_rng(3 * m)
    terms = P.separable_terms(m, rng, 2, "n")
    u = P.evaluate(m, terms)
    f = P.helmholtz_rhs(m, terms, 0.0)
    sol = oracle_mod.solvediag(bases[m], f, "n", 0.0)
    a = b.iofzero
    assert abs(sol[a, a]) <= 1e-12 * np.max(np.abs(u))               # U(iofzero+1, iofzero+1) = wanted
    d = sol - u
    assert np.max(np.abs(d - d[a, a])) <= 1e-11 * np.max(np.abs(u)) * max(1, m / 8)


@pytest.mark.parametrize("m", [33, 65])
def test_dpval_known_answer(oracle_mod, bases, m):
    """-(D_yy + D_zz) U = dpval(aspect=1) gives max U = 1 (operators.F:2486-2498)."""
    dpval = 3.39344926982230
    U = _zero_boundary(np.full((m, m), -dpval))
    sol = oracle_mod.solvediag(bases[m], U, "d", 0.0)
    assert abs(sol.max() - 1.0) < 1e-12
    assert sol.argmax() == (m // 2) * m + m // 2     # at the duct centre


def test_helm3d_modes_and_k0(oracle_mod, bases):
    m, M, alp, alam = 17, 5, 0.7, 50.0
    ob = bases[m]
    rng = np.random.default_rng(5)
    f = np.zeros((M, m, m), dtype=np.complex128)
    us = []
    for k in range(M):
        tr = P.separable_terms(m, rng, 2, "d")
        ti = P.separable_terms(m, rng, 2, "d")
        xl = alam + (alp * k) ** 2
        ur, ui = P.evaluate(m, tr), P.evaluate(m, ti)
        f[k] = P.helmholtz_rhs(m, tr, xl) + 1j * P.helmholtz_rhs(m, ti, xl)
        us.append(ur + (1j * ui if k > 0 else 0))
    f[:, 0, :] = 7.0 + 3j   # boundary garbage: helm3d zeroes it in place
    sol, rhs_after = oracle_mod.helm3d(ob, f, "d", alam, alp=alp, ib=0)
    assert not rhs_after[:, 0, :].any()
    for k in range(M):
        assert np.max(np.abs(sol[k] - us[k])) <= 1e-11 * np.max(np.abs(us[k]))
    assert not sol[0].imag.any()                           # Im skipped at k=0
    # ib offset: the same field solved as modes 1..M gives different xlambda
    sol1, _ = oracle_mod.helm3d(ob, f, "d", alam, alp=alp, ib=1)
    assert np.max(np.abs(sol1[1] - sol[1])) > 1e-6 * np.max(np.abs(sol[1]))


def test_bad_bound_raises(oracle_mod, bases):
    with pytest.raises(ValueError):
        oracle_mod.solvediag(bases[9], np.zeros((9, 9)), "x", 1.0)
    with pytest.raises(ValueError):
        oracle_mod.helm3d(bases[9], np.zeros((2, 9, 9), dtype=np.complex128), "q", 1.0)


def test_blas_and_loop_gemm_agree(oracle_mod, bases):
    m = 33
    rng = np.random.default_rng(11)
    U = rng.standard_normal((m, m))
    a = oracle_mod.solvediag(bases[m], U, "n", 3.0)
    desc = oracle_mod.use_blas(True)
    try:
        b = oracle_mod.solvediag(bases[m], U, "n", 3.0)
    finally:
        oracle_mod.use_blas(False)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a)), desc
