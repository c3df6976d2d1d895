"""The pseudo-spectral step on device fields: advection term and one substep.

advection()  duct_advection_dev_: u.grad(u) in divergence form (advection3d,
             src/operators.F:9-276) from u, v, w in Fourier(x)-physical(y, z) space.
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
