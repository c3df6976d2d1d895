"""Oracle restatement of the pseudo-spectral pieces of one Runge-Kutta substep.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the duct_amd
package).  Each function follows the reference routine it names, in the
reference's own algorithm (FFTW r2c/c2r and REDFT00 transforms, the
coefficient recurrence), so that the GPU path, which takes a different route
(dense x-DFTs on matrix cores, collocation products in physical space), is
checked against the reference's arithmetic:

  cheb_forward / cheb_inverse   tcheb_fftw, src/tchebw.f:62-150 (REDFT00, nnyc = 0)
  derivyz                       src/chebyshev.f90:1-295 (first/second derivative recurrence)
  transyz                       src/chebyshev.f90:458-540
  cheb_interp                   src/chebyshev.f90:725-765 (read_4fields_fou_hdf5 onto another grid)
  fourx_inverse / fourx_forward src/fourier.F:1-115 (fill, c2r / r2c, pack_fftw with 1/mgalx)
  derivx                        src/fourier.F:117-188
  mult                          src/mult.F:11-44
  advection3d                   src/operators.F:9-276 (divergence form, one rank: chx2z/chz2x
                                are plain transposes, slip/CFL/statistics left out)

Array conventions: a complex field is indexed [mode, z, y] (C order), i.e. the
reference's u(my, mz, ib:ie) with y fastest; a physical field is [z, y, x].
The Chebyshev coefficients of a line x_j = -cos(pi j / N), j = 0..N, are those
of tcheb_fftw: u(x_j) = sum_k c_k T_k(cos(pi j / N)).
"""
from __future__ import annotations

import numpy as np
import scipy.fft


# ---- Chebyshev transforms (tchebw.f:62-150) -------------------------------------
def cheb_forward(u: np.ndarray, axis: int) -> np.ndarray:
    """phys -> cheb along axis (iopt < 0): REDFT00, then 1/N inside, 1/(2N) at the ends."""
    N = u.shape[axis] - 1
    y = scipy.fft.dct(u, type=1, axis=axis)          # FFTW REDFT00, unnormalised
    y = np.moveaxis(y, axis, -1)
    out = y / N
    out[..., 0] = y[..., 0] * (0.5 / N)
    out[..., -1] = y[..., -1] * (0.5 / N)
    return np.moveaxis(out, -1, axis)


def cheb_interp(plane: np.ndarray, my: int, mz: int) -> np.ndarray:
    """cheb_interp (chebyshev.f90:725-765) of one complex plane [z, y] of myr x mzr points onto my x mz.

    transyz_plane_fou(-1) (:767-800) takes the plane to Chebyshev coefficients in y, then z; the
    result at the new nodes y2 = -dglc(-1, 1, my) (:741-743, dglc :429-456) sums only the first
    min(20, myr) x min(20, mzr) coefficients, each times cos(k acos(y2)) cos(l acos(z2)). The
    small-coefficient test (:755) ends in `continue`, a no-op, so no term is skipped."""
    mzr, myr = plane.shape
    c = cheb_forward(cheb_forward(plane, axis=1), axis=0)          # [l, k]
    ky, kz = min(20, myr), min(20, mzr)
    pi = 4.0 * np.arctan(1.0)

    def nodes(n):
        x = -np.cos(np.arange(n) * (pi / (n - 1)))                 # dglc, then y2 = -x
        return -x

    ty = np.cos(np.outer(np.arange(ky), np.arccos(nodes(my))))    # [k, i]
    tz = np.cos(np.outer(np.arange(kz), np.arccos(nodes(mz))))    # [l, j]
    return tz.T @ c[:kz, :ky] @ ty                                  # [j, i]


def cheb_inverse(c: np.ndarray, axis: int) -> np.ndarray:
    """cheb -> phys along axis (iopt >= 0): double the end coefficients, REDFT00, times 1/2."""
    c = np.moveaxis(np.array(c, copy=True), axis, -1)
    c[..., 0] *= 2.0
    c[..., -1] *= 2.0
    y = scipy.fft.dct(c, type=1, axis=-1) * 0.5
    return np.moveaxis(y, -1, axis)


def transyz(field: np.ndarray, cdir: str, iopt: int) -> np.ndarray:
    """chebyshev.f90:458-540 on a complex [mode, z, y] field (real and imaginary parts alike)."""
    axis = -1 if cdir == "y" else -2
    f = cheb_forward if iopt < 0 else cheb_inverse
    return f(field.real, axis) + 1j * f(field.imag, axis)


# ---- derivative recurrence (chebyshev.f90:1-295) --------------------------------
def _deriv1(c: np.ndarray) -> np.ndarray:
    """One pass of the recurrence along the last axis: c_k = c_{k+2} + 2k u_k from the top,
    du_k = -c_{k+1}, du_0 = -c_1 / 2 (the minus sign: coefficients are in cos(pi j / N) = -x_j)."""
    n = c.shape[-1]
    du = np.zeros_like(c)
    aari = np.zeros(c.shape[:-1], dtype=c.dtype)
    bri = np.zeros_like(aari)
    tri = np.zeros_like(aari)
    for kk in range(n - 1, 0, -1):            # ii = 1..my1, kk = my - ii, ia - 1 = kk
        cri = bri + 2.0 * float(kk) * c[..., kk]
        du[..., kk] = -tri
        bri = aari
        tri = cri
        aari = cri
    du[..., 0] = -0.5 * tri
    return du


def derivyz(u: np.ndarray, cdir: str, ider: int, scal: float = 1.0, aspect: float = 1.0,
            du: np.ndarray | None = None, what: str = "n") -> np.ndarray:
    """d^ider/dy^ider or d/dz on Chebyshev coefficients [mode, z, y]; du = scal*d(u) ('n') or
    du + scal*d(u) (else); z derivatives carry 1/aspect per order."""
    if ider not in (1, 2):
        raise ValueError("this order of deriv not implemented")
    x = u if cdir == "y" else np.swapaxes(u, -1, -2)
    d = _deriv1(x)
    if ider == 2:
        d = _deriv1(d)
    if cdir == "z":
        d = np.swapaxes(d, -1, -2) / aspect ** ider
    d = scal * d
    if what == "n" or du is None:
        return d
    return du + d


# ---- Fourier transforms in x (fourier.F:1-115) ----------------------------------
def fourx_inverse(fou: np.ndarray, mgalx: int) -> np.ndarray:
    """fou [mode, z, y] (modes 0..mx1) -> phys [z, y, x] with mgalx points: fill (zeros above
    mx1), FFTW c2r (unnormalised; Im of mode 0 ignored)."""
    M = fou.shape[0]
    full = np.zeros((mgalx // 2 + 1,) + fou.shape[1:], dtype=np.complex128)
    full[:M] = fou
    phys = np.fft.irfft(full, n=mgalx, axis=0) * mgalx
    return np.moveaxis(phys, 0, -1)


def fourx_forward(phys: np.ndarray, M: int) -> np.ndarray:
    """phys [z, y, x] -> fou [mode, z, y]: FFTW r2c, keep modes 0..M-1 scaled by 1/mgalx (pack_fftw)."""
    mgalx = phys.shape[-1]
    y = np.fft.rfft(phys, axis=-1)[..., :M] / float(mgalx)
    return np.moveaxis(y, -1, 0)


def derivx(u: np.ndarray, alp: float, ib: int = 0, scal: float = 1.0) -> np.ndarray:
    """fourier.F:117-188, ider = 1: d/dx -> i * alp * k for modes ib.."""
    k = (ib + np.arange(u.shape[0]))[:, None, None]
    return scal * (1j * alp * k) * u


def mult(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """mult.F:11-44, what = 'n', plusm = '+'."""
    return a * b


# ---- advection3d (operators.F:9-276) --------------------------------------------
def advection3d(u, v, w, mgalx: int, alp: float = 1.0, aspect: float = 1.0, ib: int = 0):
    """Divergence form u.grad(u): returns (adv1, adv2, adv3) in f-p-p space [mode, z, y]."""
    M = u.shape[0]
    uf, vf, wf = (fourx_inverse(x, mgalx) for x in (u, v, w))

    def back(p):
        return fourx_forward(p, M)

    def dy(f):
        return transyz(derivyz(transyz(f, "y", -1), "y", 1), "y", +1)

    def dz(f):
        return transyz(derivyz(transyz(f, "z", -1), "z", 1, aspect=aspect), "z", +1)

    uu, vv, ww = back(mult(uf, uf)), back(mult(vf, vf)), back(mult(wf, wf))
    uv, uw, vw = back(mult(uf, vf)), back(mult(uf, wf)), back(mult(vf, wf))
    adv1 = derivx(uu, alp, ib)
    adv2 = dy(vv)
    adv3 = dz(ww)
    adv2 = adv2 + derivx(uv, alp, ib)
    adv1 = adv1 + dy(uv)
    adv3 = adv3 + derivx(uw, alp, ib)
    adv1 = adv1 + dz(uw)
    adv3 = adv3 + dy(vw)
    adv2 = adv2 + dz(vw)
    return adv1, adv2, adv3


# ---- the remaining operators of the Runge-Kutta substep (operators.F:278-605) ----
def _dy(f, ider=1, scal=1.0):
    return transyz(derivyz(transyz(f, "y", -1), "y", ider, scal=scal), "y", +1)


def _dz(f, ider=1, scal=1.0, aspect=1.0):
    return transyz(derivyz(transyz(f, "z", -1), "z", ider, scal=scal, aspect=aspect), "z", +1)


def derivx2(u: np.ndarray, alp: float, ib: int = 0, scal: float = 1.0) -> np.ndarray:
    """fourier.F:117-188, ider = 2: -(alp k)^2."""
    k = (ib + np.arange(u.shape[0]))[:, None, None]
    return -scal * (alp * k) ** 2 * u


def laplacian3d(u, scal: float, alp: float = 1.0, aspect: float = 1.0, ib: int = 0):
    """operators.F:278-347, what = 'n': scal (d_xx + d_yy + d_zz) u, the y and z parts through
    transyz / derivyz(ider = 2) / transyz."""
    flap = derivx2(u, alp, ib, scal)
    flap = flap + _dy(u, 2, scal)
    return flap + _dz(u, 2, scal, aspect)


def grad3d(u, scal: float, alp: float = 1.0, aspect: float = 1.0, ib: int = 0):
    """operators.F:349-425, what = 'n': scal grad(u) -> (dxu, dyu, dzu)."""
    return derivx(u, alp, ib, scal), _dy(u, 1, scal), _dz(u, 1, scal, aspect)


def div3d(u, v, w, scal: float, alp: float = 1.0, aspect: float = 1.0, ib: int = 0):
    """operators.F:427-501: scal (du/dx + dv/dy + dw/dz)."""
    div = derivx(u, alp, ib, scal)
    div = div + _dy(v, 1, scal)
    return div + _dz(w, 1, scal, aspect)


def update_uvwp(u, v, w, p, phi, akk: float, bkk: float, dt: float, fnu: float, alp: float = 1.0,
                aspect: float = 1.0, ib: int = 0):
    """operators.F:503-605 without FILTER: p += phi - (a+b) dt nu lap(phi);
    u_i += -dt (a+b) d_i phi.  Returns (u, v, w, p) as new arrays."""
    rhsphi = laplacian3d(phi, (akk + bkk) * dt * fnu, alp, aspect, ib)
    p = p + phi - rhsphi
    scal = -dt * (akk + bkk)
    u = u + derivx(phi, alp, ib, scal)
    v = v + _dy(phi, 1, scal)
    w = w + _dz(phi, 1, scal, aspect)
    return u, v, w, p
