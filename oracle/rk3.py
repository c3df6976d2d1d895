"""Oracle restatement of one Runge-Kutta substep of the velocity and pressure.

TEST INFRASTRUCTURE ONLY: imported by tests/ (never by the duct_amd package).
Follows main1.f90:384-655 item by item for one rank owning all modes (ib = 0),
icase = 1 without CMASSFLOW / FIXED_ZERO_MODE / forcing / FILTER, no temperature:

  (i)    h = -rkkc adv(u^{k-2})            (h = 0 at krk = 1)
  (ii)   adv = advection3d(u^{k-1})
  (iii)  h -= gkkc adv
  (iv)   h += nu akkc lap(u)
  (v)    h += u / dt
  (vi)   h -= (akkc + bkkc) grad(p)
  (vi-b) h1 (mode 0, real part) -= (akkc + bkkc) dpdx,  dpdx = get_dpdx(aspect, nu)
  (vii)  h *= -1 / (nu bkkc)
  (viii) u* = helm3d(h, 'd', 1 / (bkkc nu dt))   (u, v, w)
  (ix)   h1 = div3d(u*) / ((akkc + bkkc) dt)
  (x)    phi = helm3d(h1, 'n', 0)
  (xi)   update_uvwp(u*, p, phi)

The Helmholtz solves are the caller's `helm3d(rhs, bound, alambda) -> sol`
(the C restatement of solvediag in oracle/__init__.py); the operators are
oracle/spectral.py's FFT / DCT-I / recurrence route.
"""
from __future__ import annotations

import numpy as np

from . import spectral as S

# mod.f90:51-58 (Spalart, Moser & Rogers low-storage RK3)
AKKC = (29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0)
BKKC = (37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0)
GKKC = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
RKKC = (0.0, -17.0 / 60.0, -5.0 / 12.0)

# operators.F:2494-2507: -(d_yy + d_zz) U = dpval gives max U = 1, per aspect ratio
ASPECTVAL = tuple(1.0 + 0.5 * i for i in range(19))
DPVAL = (3.39344926982230, 2.48086300035167, 2.19545075660646, 2.08474987413869, 2.03778220024811,
         2.01705223568555, 2.00773897291851, 2.00352109020906, 2.00160386398057, 2.00073094391977,
         2.00033319895779, 2.00015190434113, 2.00006925611912, 2.00003157591041, 2.00001439653781,
         2.00000656390479, 2.00000299272912, 2.00000136449820, 2.00000062212652)


def get_dpdx(aspect: float, fnu: float) -> float:
    """operators.F:2478-2522: -fnu dpval(aspect); aspect must be tabulated (to 1e-5)."""
    idx = [i for i, a in enumerate(ASPECTVAL) if abs(a - aspect) <= 1e-5]
    if not idx:
        raise ValueError("this aspect ratio is not tabulated! %r" % aspect)
    return -fnu * DPVAL[idx[-1]]


def rk3_substep(u, v, w, p, adv, krk: int, fnu: float, dt: float, mgalx: int, helm3d, alp: float = 1.0,
                aspect: float = 1.0, dpdx: float | None = None):
    """One substep krk (1..3).  adv = (adv1, adv2, adv3) of the previous substep (ignored at
    krk = 1).  Returns (u, v, w, p, adv_new, phi) as new arrays."""
    k = krk - 1
    a, b, g, r = AKKC[k], BKKC[k], GKKC[k], RKKC[k]
    if dpdx is None:
        dpdx = get_dpdx(aspect, fnu)
    vel = (u, v, w)
    if krk > 1:
        h = [-r * x for x in adv]                                    # (i)
    else:
        h = [np.zeros_like(u) for _ in range(3)]
    advn = S.advection3d(u, v, w, mgalx=mgalx, alp=alp, aspect=aspect)   # (ii)
    h = [hc - g * an for hc, an in zip(h, advn)]                       # (iii)
    h = [hc + S.laplacian3d(x, fnu * a, alp, aspect) for hc, x in zip(h, vel)]   # (iv)
    h = [hc + x / dt for hc, x in zip(h, vel)]                         # (v)
    gp = S.grad3d(p, -(a + b), alp, aspect)                            # (vi)
    h = [hc + gc for hc, gc in zip(h, gp)]
    h[0] = h[0].copy()
    h[0][0] += -(a + b) * dpdx                                         # (vi-b): mode 0, real part
    h = [hc * (-1.0 / (fnu * b)) for hc in h]                          # (vii)
    helco = 1.0 / (b * fnu * dt)
    us = [helm3d(hc, "d", helco) for hc in h]                          # (viii)
    div = S.div3d(us[0], us[1], us[2], 1.0 / ((a + b) * dt), alp, aspect)   # (ix)
    phi = helm3d(div, "n", 0.0)                                        # (x)
    un, vn, wn, pn = S.update_uvwp(us[0], us[1], us[2], p, phi, a, b, dt, fnu, alp, aspect)   # (xi)
    return un, vn, wn, pn, tuple(advn), phi
