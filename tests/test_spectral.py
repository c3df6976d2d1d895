"""Pins of the pseudo-spectral oracle (oracle/spectral.py) on the CPU.

Each restated routine is checked against what the mathematics fixes: the
Chebyshev transform against direct evaluation of T_k at the nodes, the
derivative recurrence against numpy's Chebyshev differentiation, transform +
recurrence + inverse against the collocation matrix DC, the x-transforms
against direct DFT sums and the 3/2 rule against exact convolution of the
coefficient sequences, advection3d against products and derivatives evaluated
by other routes (direct sums, DC products).
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from oracle import spectral as S
from oracle.pseudomat import cgl_nodes, colloc_deriv


def _xi(m):
    return np.cos(np.pi * np.arange(m) / (m - 1))       # cos(pi j / N) = -x_j


@pytest.mark.parametrize("m", [9, 33, 129])
def test_cheb_forward_is_the_series_of_the_node_values(m):
    rng = np.random.default_rng(m)
    c = rng.standard_normal((3, m))
    u = C.chebval(_xi(m), c.T)                            # u(x_j) = sum c_k T_k(cos(pi j/N)), [3, m]
    assert np.allclose(S.cheb_forward(u, -1), c, rtol=0, atol=1e-13 * np.abs(c).max() * m)
    assert np.allclose(S.cheb_inverse(c, -1), u, rtol=0, atol=1e-13 * np.abs(u).max())


@pytest.mark.parametrize("m", [9, 33, 65])
def test_derivyz_recurrence_is_chebyshev_differentiation(m):
    rng = np.random.default_rng(2 * m)
    c = rng.standard_normal((2, 3, m)) + 1j * rng.standard_normal((2, 3, m))
    for ider in (1, 2):
        got = S.derivyz(c, "y", ider, scal=0.75)
        want = np.zeros_like(c)
        d = C.chebder(np.moveaxis(c, -1, 0), m=ider) * (-1.0) ** ider     # d/dx = -d/dxi
        want[..., : m - ider] = np.moveaxis(d, 0, -1)
        assert np.allclose(got, 0.75 * want, rtol=0, atol=1e-12 * np.abs(want).max())
    gz = S.derivyz(np.swapaxes(c, -1, -2), "z", 1, aspect=2.0)
    gy = S.derivyz(c, "y", 1)
    assert np.allclose(np.swapaxes(gz, -1, -2), gy / 2.0, rtol=0, atol=1e-13 * np.abs(gy).max())


@pytest.mark.parametrize("m", [17, 65, 129])
def test_transform_recurrence_inverse_equals_collocation_matrix(m):
    """transyz + derivyz + transyz (the reference's route) == DC applied in physical space."""
    rng = np.random.default_rng(3 * m)
    f = rng.standard_normal((2, m, m)) + 1j * rng.standard_normal((2, m, m))
    DC = colloc_deriv(m)
    dy = S.transyz(S.derivyz(S.transyz(f, "y", -1), "y", 1), "y", +1)
    dz = S.transyz(S.derivyz(S.transyz(f, "z", -1), "z", 1), "z", +1)
    want_y = f @ DC.T                                     # y is the last axis
    want_z = np.einsum("ij,mjk->mik", DC, f)
    scale = np.abs(DC).sum(1).max() * np.abs(f).max()
    assert np.max(np.abs(dy - want_y)) <= 1e-13 * scale
    assert np.max(np.abs(dz - want_z)) <= 1e-13 * scale


def test_fourx_inverse_is_the_c2r_sum_and_round_trips():
    M, mgalx, m = 16, 48, 5
    rng = np.random.default_rng(5)
    c = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    phys = S.fourx_inverse(c, mgalx)                      # [z, y, x]
    j = np.arange(mgalx)
    k = np.arange(1, M)
    w = np.exp(2j * np.pi * np.outer(k, j) / mgalx)       # [k, x]
    direct = c[0].real[..., None] + 2.0 * np.real(np.einsum("kzy,kx->zyx", c[1:], w))
    assert np.allclose(phys, direct, rtol=0, atol=1e-12 * np.abs(direct).max())
    back = S.fourx_forward(phys, M)
    c0 = c.copy()
    c0[0] = c0[0].real                                    # Im of mode 0 is not representable
    assert np.allclose(back, c0, rtol=0, atol=1e-13 * np.abs(c).max())


def test_three_halves_rule_makes_products_exact():
    """With mgalx >= 3 M, the kept modes of a product equal the exact convolution."""
    M, mgalx = 16, 48
    rng = np.random.default_rng(6)
    a = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    b = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    a[0], b[0] = a[0].real, b[0].real
    full_a = np.concatenate([np.conj(a[:0:-1]), a])       # modes -(M-1)..M-1
    full_b = np.concatenate([np.conj(b[:0:-1]), b])
    conv = np.convolve(full_a, full_b)                    # modes -(2M-2)..2M-2
    want = conv[2 * M - 2: 2 * M - 2 + M]
    pa = S.fourx_inverse(a[:, None, None], mgalx)
    pb = S.fourx_inverse(b[:, None, None], mgalx)
    got = S.fourx_forward(pa * pb, M)[:, 0, 0]
    assert np.allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_advection3d_against_direct_products_and_collocation():
    """The reference's route (FFT, DCT, recurrence) against direct x-sums and DC products."""
    m, M, mgalx, alp = 17, 8, 24, 1.3
    rng = np.random.default_rng(7)
    fld = []
    for _ in range(3):
        f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
        f[0] = f[0].real
        fld.append(f)
    adv = S.advection3d(*fld, mgalx=mgalx, alp=alp)
    # direct route
    j = np.arange(mgalx)
    E = np.exp(2j * np.pi * np.outer(np.arange(M), j) / mgalx)      # [k, x]
    wts = np.full(M, 2.0)
    wts[0] = 1.0

    def to_phys(c):
        cc = c.copy()
        cc[0] = cc[0].real
        return np.real(np.einsum("k,kzy,kx->zyx", wts, cc, E))

    def to_fou(p):
        return np.einsum("zyx,kx->kzy", p, np.conj(E)) / mgalx

    u, v, w = (to_phys(x) for x in fld)
    DC = colloc_deriv(m)
    k = np.arange(M)[:, None, None]

    def dx(f):
        return 1j * alp * k * f

    def dy(f):
        return f @ DC.T

    def dz(f):
        return np.einsum("ij,mjk->mik", DC, f)

    uu, vv, ww, uv, uw, vw = (to_fou(p) for p in (u * u, v * v, w * w, u * v, u * w, v * w))
    want = (dx(uu) + dy(uv) + dz(uw), dx(uv) + dy(vv) + dz(vw), dx(uw) + dy(vw) + dz(ww))
    for g, wnt in zip(adv, want):
        assert np.max(np.abs(g - wnt)) <= 1e-12 * np.abs(wnt).max()
    _ = cgl_nodes(m)
