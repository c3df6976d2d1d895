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


def c3_step(u, v, w, nu: float, dt: float, mgalx: int, alp: float = 1.0):
    """One substep; returns (u_new, v_new, w_new, phi, (adv1, adv2, adv3)).  Inputs are not modified."""
    import torch
    adv = [torch.empty_like(u) for _ in range(3)]
    advection(u, v, w, *adv, mgalx=mgalx, alp=alp)
    alam = 1.0 / (nu * dt)
    rhs = [a / nu - alam * x for a, x in zip(adv, (u, v, w))]
    un, vn, wn, phi = projection_step(*rhs, alam, alp=alp, ib=0, c=dt)
    return un, vn, wn, phi, tuple(adv)
