"""Pins of the mixed-boundary (temperature) solve oracle (oracle/temp.py) on the CPU.

Manufactured solutions: u(y, z) a polynomial of degree <= m-1 in each variable on
[-sy, sy] x [-sz, sz], so collocation is exact.  The right-hand side is the exact
u_yy + u_zz - xlambda u (numpy chebder), the wall data are u (Dirichlet) or its exact
y / z derivative (Neumann) per wall, for every one of the 16 wall combinations; the solve
must return u on the interior and the edges (the corners keep the right-hand side's values,
as solvediag_temp.f90 leaves them).  Independent of the oracle's matrices: a mistake in the
elimination of a wall, in a wall-data term or in an edge formula fails the wall combinations
that use it.
"""
import itertools

import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from oracle import temp as T


def manufactured(m, sy, sz, xlam, seed, deg=None, mz=None):
    """u on the [mz, my] grid (my = m), degree <= deg in y and <= deg + mz - m in z."""
    mz = m if mz is None else mz
    deg = m - 3 if deg is None else deg
    degz = deg + mz - m
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((degz + 1, deg + 1)) * np.exp(-0.3 * np.add.outer(np.arange(degz + 1),
                                                                             np.arange(deg + 1)))
    xy = -np.cos(np.pi * np.arange(m) / (m - 1))
    xz = -np.cos(np.pi * np.arange(mz) / (mz - 1))

    def ev(Ac):                                    # Ac[q, p] -> values [z, y] at the nodes
        Ty = np.stack([C.chebval(xy, np.eye(Ac.shape[1])[p]) for p in range(Ac.shape[1])], 1)
        Tz = np.stack([C.chebval(xz, np.eye(Ac.shape[0])[q]) for q in range(Ac.shape[0])], 1)
        return Tz @ Ac @ Ty.T

    u = ev(A)
    uy = ev(C.chebder(A, 1, axis=1)) / sy
    uz = ev(C.chebder(A, 1, axis=0)) / sz
    uyy = ev(C.chebder(A, 2, axis=1)) / sy ** 2
    uzz = ev(C.chebder(A, 2, axis=0)) / sz ** 2
    f = uyy + uzz - xlam * u
    return u, uy, uz, f


def wall_data(bounds, u, uy, uz):
    by1, by2, bz1, bz2 = bounds
    bcy = np.stack([u[:, 0] if by1 == "d" else uy[:, 0], u[:, -1] if by2 == "d" else uy[:, -1]])   # over z
    bcz = np.stack([u[0, :] if bz1 == "d" else uz[0, :], u[-1, :] if bz2 == "d" else uz[-1, :]])   # over y
    return bcy, bcz


@pytest.mark.parametrize("bounds", ["".join(c) for c in itertools.product("dn", repeat=4)])
def test_manufactured_solution_every_wall_combination(bounds):
    m, sy, sz, xlam = 17, 1.0, 1.5, 2.5
    tb = T.TempBasis(m, bounds, sy, sz)
    u, uy, uz, f = manufactured(m, sy, sz, xlam, int(bounds.replace("d", "0").replace("n", "1"), 2))
    bcy, bcz = wall_data(bounds, u, uy, uz)
    got = T.solvediag_temp(tb, f, xlam, bcy, bcz)
    inner = np.ones((m, m), bool)
    inner[0, 0] = inner[0, -1] = inner[-1, 0] = inner[-1, -1] = False
    err = np.abs(got - u)[inner].max() / np.abs(u).max()
    assert err <= 1e-10, (bounds, err)
    for (a, b) in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
        assert got[a, b] == f[a, b]                                   # corners untouched


@pytest.mark.parametrize("my,mz", [(13, 21), (21, 13), (9, 33)])
@pytest.mark.parametrize("bounds", ["dndn", "nddn", "nnnn", "dddd"])
def test_manufactured_solution_rectangular_grid(my, mz, bounds):
    """my != mz: the y basis has my points, the z basis mz (helm3d_temp.f90:35-36); a transposed
    operand or a swapped size in any step fails these."""
    sy, sz, xlam = 1.0, 2.0, 1.5
    tb = T.TempBasis(my, bounds, sy, sz, mz=mz)
    assert tb.Py.shape == (my - 2, my - 2) and tb.Pz.shape == (mz - 2, mz - 2)
    u, uy, uz, f = manufactured(my, sy, sz, xlam, my * 100 + mz, mz=mz)
    assert f.shape == (mz, my)
    bcy, bcz = wall_data(bounds, u, uy, uz)
    assert bcy.shape == (2, mz) and bcz.shape == (2, my)
    got = T.solvediag_temp(tb, f, xlam, bcy, bcz)
    inner = np.ones((mz, my), bool)
    inner[0, 0] = inner[0, -1] = inner[-1, 0] = inner[-1, -1] = False
    err = np.abs(got - u)[inner].max() / np.abs(u).max()
    assert err <= 1e-10, (my, mz, bounds, err)
    with pytest.raises(ValueError):
        T.solvediag_temp(tb, f.T.copy(), xlam, bcy, bcz)


def test_helm3d_temp_modes_and_rhs():
    m, M, sy, sz, alam, alp = 13, 3, 1.0, 1.0, 4.0, 0.7
    bounds = "dndn"
    tb = T.TempBasis(m, bounds, sy, sz)
    rhs = np.zeros((M, m, m), complex)
    bcy = np.zeros((M, 2, m))
    bcz = np.zeros((M, 2, m))
    want = np.zeros((M, m, m), complex)
    for k in range(M):
        xlam = alam + (alp * k) ** 2
        u, uy, uz, f = manufactured(m, sy, sz, xlam, 7 + k)
        by, bz = wall_data(bounds, u, uy, uz)
        bcy[k], bcz[k] = by, bz
        rhs[k] = f + 1j * f * (k > 0)
        want[k] = u + 1j * u * (k > 0)
    r0 = rhs.copy()
    sol = T.helm3d_temp(tb, rhs, alam, bcy, bcz, alp=alp)
    assert np.array_equal(rhs, r0)                                    # the rhs is not modified
    assert not sol[0].imag.any()                                      # Im not solved at k = 0
    inner = np.ones((m, m), bool)
    inner[0, 0] = inner[0, -1] = inner[-1, 0] = inner[-1, -1] = False
    for k in range(M):
        assert np.abs(sol[k] - want[k])[inner].max() <= 1e-10 * np.abs(want[k]).max()


def test_bad_bounds_raise():
    with pytest.raises(ValueError):
        T.TempBasis(9, "dnx d")
