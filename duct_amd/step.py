"""The pseudo-spectral step on device fields: advection term and one substep.

advection()  duct_advection_dev_: u.grad(u) in divergence form (advection3d,
             src/operators.F:9-276) from u, v, w in Fourier(x)-physical(y, z) space.
rk3_substep() duct_rk3_substep_dev_: one substep of the reference's low-storage RK3 for u, v, w
             and p (src/main1.f90:384-655): advection, viscous and pressure terms, the three
             Helmholtz solves, the Poisson solve and update_uvwp, all on the device.
rk3_step()   the three substeps of one time step (main1.f90:384-723).
c3_step()    one backward-Euler projection substep through the hot path:
               (D_y^2 + D_z^2 - k^2 - 1/(nu dt)) u* = adv/nu - u/(nu dt)   (helm3d 'd', u, v, w)
               phi = N(div u* / dt),  u_new = u* - dt grad(phi)           (helm3d 'n', projection)
             i.e. the structure of src/main1.f90:606-655 (predictor, div3d, Poisson,
             update_uvwp) with the explicit viscous and old-pressure terms left out.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .adjoint import projection_step
from .solver import _check_field, _use_current_stream


def advection(u, v, w, adv1, adv2, adv3, mgalx: int, alp: float = 1.0, aspect: float = 1.0) -> None:
    """adv_c = the c-component of u.grad(u) (duct_advection_dev_); all fields complex128 (M, m, m)."""
    m = u.shape[-1]
    for t in (u, v, w, adv1, adv2, adv3):
        _check_field(t, m)
    _use_current_stream()
    _lib.call("duct_advection_dev_", *[_lib.byref_ptr(t.data_ptr()) for t in (u, v, w, adv1, adv2, adv3)],
              _lib.byref_int(m), _lib.byref_int(u.shape[0]), _lib.byref_int(mgalx), _lib.byref_double(alp),
              _lib.byref_double(aspect))


def transyz(inp, out, cdir: str, iopt: int) -> None:
    """out = Chebyshev transform of inp along y or z (duct_transyz_dev_): iopt < 0 physical ->
    coefficients, iopt >= 0 coefficients -> physical."""
    m = inp.shape[-1]
    _check_field(inp, m)
    _check_field(out, m)
    _use_current_stream()
    _lib.call("duct_transyz_dev_", _lib.byref_ptr(inp.data_ptr()), _lib.byref_ptr(out.data_ptr()),
              _lib.byref_int(m), _lib.byref_int(inp.shape[0]), cdir.encode(), _lib.byref_int(iopt),
              ctypes.c_size_t(1))


def fourx(fou, phys, mgalx: int, iopt: int) -> None:
    """fourx on the z-cut layouts (duct_fourx_dev_): fou complex128 (nz, my, nmodes) [z, y, mode],
    phys float64 (nz, my, mgalx + 2) [z, y, x]; iopt >= 0 fou -> phys, < 0 phys -> fou."""
    nz, my, M = fou.shape
    if phys.shape != (nz, my, mgalx + 2) or not (fou.is_contiguous() and phys.is_contiguous()):
        raise ValueError("fou (nz, my, nmodes) and phys (nz, my, mgalx+2), contiguous")
    _use_current_stream()
    _lib.call("duct_fourx_dev_", _lib.byref_ptr(fou.data_ptr()), _lib.byref_ptr(phys.data_ptr()),
              _lib.byref_int(my), _lib.byref_int(nz), _lib.byref_int(M), _lib.byref_int(mgalx),
              _lib.byref_int(iopt))


def derivyz(u, du, cdir: str, ider: int = 1, scal: float = 1.0, what: str = "n", aspect: float = 1.0) -> None:
    """Chebyshev-coefficient derivative along y or z (duct_derivyz_dev_, derivyz)."""
    m = u.shape[-1]
    _check_field(u, m)
    _check_field(du, m)
    _use_current_stream()
    _lib.call("duct_derivyz_dev_", _lib.byref_ptr(u.data_ptr()), _lib.byref_ptr(du.data_ptr()), _lib.byref_int(m),
              _lib.byref_int(u.shape[0]), cdir.encode(), _lib.byref_int(ider), _lib.byref_double(scal),
              what.encode(), _lib.byref_double(aspect), ctypes.c_size_t(1), ctypes.c_size_t(1))


def c3_step(u, v, w, nu: float, dt: float, mgalx: int, alp: float = 1.0):
    """One substep; returns (u_new, v_new, w_new, phi, (adv1, adv2, adv3)).  Inputs are not modified."""
    import torch
    adv = [torch.empty_like(u) for _ in range(3)]
    advection(u, v, w, *adv, mgalx=mgalx, alp=alp)
    alam = 1.0 / (nu * dt)
    rhs = [a / nu - alam * x for a, x in zip(adv, (u, v, w))]
    un, vn, wn, phi = projection_step(*rhs, alam, alp=alp, ib=0, c=dt)
    return un, vn, wn, phi, tuple(adv)


# mod.f90:51-58 (the low-storage RK3 of Spalart, Moser & Rogers); the library holds the same table
RK_ALPHA = (29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0)
RK_BETA = (37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0)
RK_GAMMA = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
RK_ZETA = (0.0, -17.0 / 60.0, -5.0 / 12.0)

# operators.F:2494-2507: -(d_yy + d_zz) U = dpval(aspect) gives the laminar profile with max U = 1
_ASPECTVAL = tuple(1.0 + 0.5 * i for i in range(19))
_DPVAL = (3.39344926982230, 2.48086300035167, 2.19545075660646, 2.08474987413869, 2.03778220024811,
          2.01705223568555, 2.00773897291851, 2.00352109020906, 2.00160386398057, 2.00073094391977,
          2.00033319895779, 2.00015190434113, 2.00006925611912, 2.00003157591041, 2.00001439653781,
          2.00000656390479, 2.00000299272912, 2.00000136449820, 2.00000062212652)


def get_dpdx(aspect: float, fnu: float) -> float:
    """The driving pressure gradient of pressure-driven laminar duct flow, -fnu dpval(aspect)
    (get_dpdx, operators.F:2478-2522); raises for an aspect ratio that is not tabulated."""
    idx = [i for i, a in enumerate(_ASPECTVAL) if abs(a - aspect) <= 1e-5]
    if not idx:
        raise ValueError("this aspect ratio is not tabulated! %r" % aspect)
    return -fnu * _DPVAL[idx[-1]]


def rk3_substep(u, v, w, p, adv, krk: int, fnu: float, dt: float, mgalx: int, alp: float = 1.0,
                aspect: float = 1.0, dpdx: float | None = None) -> None:
    """Substep krk (1..3) in place (duct_rk3_substep_dev_): u, v, w, p advance from u^{k-1} to
    u^k; adv = (adv1, adv2, adv3) holds advection3d(u^{k-2}) on entry (read for krk > 1) and
    advection3d(u^{k-1}) on exit.  dpdx defaults to get_dpdx(aspect, fnu) (icase = 1); pass 0
    for no driving gradient.  All fields complex128 (nmodes, m, m), modes 0..nmodes-1."""
    m = u.shape[-1]
    fields = (u, v, w, p) + tuple(adv)
    if len(fields) != 7:
        raise ValueError("adv must hold three fields")
    for t in fields:
        _check_field(t, m)
        if t.shape != u.shape:
            raise ValueError("all fields must have the shape of u")
    if dpdx is None:
        dpdx = get_dpdx(aspect, fnu)
    _use_current_stream()
    _lib.call("duct_rk3_substep_dev_", *[_lib.byref_ptr(t.data_ptr()) for t in fields],
              _lib.byref_int(m), _lib.byref_int(u.shape[0]), _lib.byref_int(mgalx), _lib.byref_double(alp),
              _lib.byref_double(aspect), _lib.byref_double(fnu), _lib.byref_double(dt), _lib.byref_int(krk),
              _lib.byref_double(dpdx))


def rk3_step(u, v, w, p, adv, fnu: float, dt: float, mgalx: int, alp: float = 1.0, aspect: float = 1.0,
             dpdx: float | None = None) -> None:
    """One full time step: substeps krk = 1, 2, 3 (main1.f90:384-723), in place."""
    for krk in (1, 2, 3):
        rk3_substep(u, v, w, p, adv, krk, fnu, dt, mgalx, alp, aspect, dpdx)
