"""Differentiation of the solve stage and of one projection step.

Forward step (the Runge-Kutta substep tail of src/main1.f90:606-655):
  u*_c = H(f_c)                          helm3d 'd', alambda (u, v, w)       main1.f90:617-619
  d    = (1/c) (dx u* + dy v* + dz w*)   div3d, scal = 1/((alpha+beta) dt)   main1.f90:630, operators.F:427
  phi  = N(d)                            helm3d 'n', alambda = 0             main1.f90:643
  u    = u* - c grad(phi)                update_uvwp, scal = -(alpha+beta) dt operators.F:503-605
with c = (alpha+beta) dt, dx = i alp k (derivx, fourier.F:117), dy, dz the
collocation derivatives (d/dy, d/dz of the nodal interpolant, as transyz +
derivyz compute them, chebyshev.f90:1-540).

The backward is hand-written: the transposed solves (duct_helm3d_adjoint_batch_dev_,
the solve with P -> P^-T, P^-1 -> P^T plus the adjoints of the Neumann edges,
corners and k=0 gauge), and the transposed derivative products.  Every map is
linear, so nothing from the forward is stored: the backward needs only the
incoming gradient (the forward is recomputed if its output is needed).

Complex fields carry the real inner product <a, b> = Re sum conj(a) b, so the
adjoint of multiplication by i k is multiplication by -i k; torch's gradient
convention for a real loss of complex inputs (dL/dRe + i dL/dIm) matches it.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .solver import _bound_bytes, _check_field, _use_current_stream, helm3d_batch, helm3d_mixed_batch


def helm3d_adjoint_batch(solbar_list, rhsbar_list, bounds: str, alambdas, alp: float = 1.0, ib: int = 0):
    """rhsbar_f = S_f^T solbar_f (duct_helm3d_adjoint_batch_dev_)."""
    nf = len(solbar_list)
    if nf != len(rhsbar_list) or nf != len(bounds) or nf != len(alambdas):
        raise ValueError("lengths differ")
    m = solbar_list[0].shape[-1]
    for t in list(solbar_list) + list(rhsbar_list):
        _check_field(t, m)
    for b in bounds:
        _bound_bytes(b)
    _use_current_stream()
    S = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in solbar_list])
    R = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in rhsbar_list])
    A = (ctypes.c_double * nf)(*[float(a) for a in alambdas])
    _lib.call("duct_helm3d_adjoint_batch_dev_", _lib.byref_int(nf), S, R, bounds.encode(), A, _lib.byref_int(m),
              _lib.byref_int(solbar_list[0].shape[0]), _lib.byref_int(ib), _lib.byref_double(alp),
              ctypes.c_size_t(nf))


def colloc_deriv(inp, out, cdir: str, trans: bool = False, scal: float = 1.0, beta: float = 0.0):
    """out = beta*out + scal * (DC or DC^T along y or z) inp (duct_colloc_deriv_dev_)."""
    m = inp.shape[-1]
    _check_field(inp, m)
    _check_field(out, m)
    if cdir not in ("y", "z"):
        raise ValueError("cdir must be 'y' or 'z'")
    _use_current_stream()
    _lib.call("duct_colloc_deriv_dev_", _lib.byref_ptr(inp.data_ptr()), _lib.byref_ptr(out.data_ptr()),
              _lib.byref_int(m), _lib.byref_int(inp.shape[0]), cdir.encode(), _lib.byref_int(1 if trans else 0),
              _lib.byref_double(scal), _lib.byref_double(beta), ctypes.c_size_t(1))


def derivx(inp, out, ib: int = 0, alp: float = 1.0, scal: float = 1.0, beta: float = 0.0):
    """out = beta*out + scal * (i alp k) inp (duct_derivx_dev_)."""
    m = inp.shape[-1]
    _check_field(inp, m)
    _check_field(out, m)
    _use_current_stream()
    _lib.call("duct_derivx_dev_", _lib.byref_ptr(inp.data_ptr()), _lib.byref_ptr(out.data_ptr()),
              _lib.byref_int(m), _lib.byref_int(inp.shape[0]), _lib.byref_int(ib), _lib.byref_double(alp),
              _lib.byref_double(scal), _lib.byref_double(beta))


def projection_step(fu, fv, fw, alambda: float, alp: float = 1.0, ib: int = 0, c: float = 1.0,
                    mixed: tuple | None = None):
    """Forward step on device fields (inputs are not modified).  Returns u, v, w, phi.
    mixed = (scheme0, scheme1, niter) runs the three Helmholtz solves and the
    Poisson solve with the mixed-precision refined solver."""
    import torch
    f = [t.clone() for t in (fu, fv, fw)]              # helm3d zeroes the rhs boundary in place
    us = [torch.empty_like(t) for t in f]
    if mixed:
        helm3d_mixed_batch(f, us, "ddd", [alambda] * 3, alp=alp, ib=ib, scheme0=mixed[0], scheme1=mixed[1],
                           niter=mixed[2])
    else:
        helm3d_batch(f, us, "ddd", [alambda] * 3, alp=alp, ib=ib)
    d = torch.empty_like(us[0])
    derivx(us[0], d, ib=ib, alp=alp, scal=1.0 / c, beta=0.0)
    colloc_deriv(us[1], d, "y", scal=1.0 / c, beta=1.0)
    colloc_deriv(us[2], d, "z", scal=1.0 / c, beta=1.0)
    phi = torch.empty_like(d)
    if mixed:
        helm3d_mixed_batch([d], [phi], "n", [0.0], alp=alp, ib=ib, scheme0=mixed[0], scheme1=mixed[1],
                           niter=mixed[2])
    else:
        helm3d_batch([d], [phi], "n", [0.0], alp=alp, ib=ib)
    u, v, w = us
    derivx(phi, u, ib=ib, alp=alp, scal=-c, beta=1.0)
    colloc_deriv(phi, v, "y", scal=-c, beta=1.0)
    colloc_deriv(phi, w, "z", scal=-c, beta=1.0)
    return u, v, w, phi


def projection_step_adjoint(gu, gv, gw, alambda: float, alp: float = 1.0, ib: int = 0, c: float = 1.0):
    """(fu_bar, fv_bar, fw_bar) = step^T (gu, gv, gw)."""
    import torch
    ub = [t.clone() for t in (gu, gv, gw)]             # ubar* starts as ubar
    phib = torch.empty_like(ub[0])
    derivx(gu, phib, ib=ib, alp=alp, scal=c, beta=0.0)                 # (-c i k)^H = c i k
    colloc_deriv(gv, phib, "y", trans=True, scal=-c, beta=1.0)
    colloc_deriv(gw, phib, "z", trans=True, scal=-c, beta=1.0)
    db = torch.empty_like(phib)
    helm3d_adjoint_batch([phib], [db], "n", [0.0], alp=alp, ib=ib)
    derivx(db, ub[0], ib=ib, alp=alp, scal=-1.0 / c, beta=1.0)         # ((1/c) i k)^H
    colloc_deriv(db, ub[1], "y", trans=True, scal=1.0 / c, beta=1.0)
    colloc_deriv(db, ub[2], "z", trans=True, scal=1.0 / c, beta=1.0)
    fb = [torch.empty_like(t) for t in ub]
    helm3d_adjoint_batch(ub, fb, "ddd", [alambda] * 3, alp=alp, ib=ib)
    return tuple(fb)


def _autograd_cls():
    import torch

    class ProjectionStep(torch.autograd.Function):
        """(fu, fv, fw) -> (u, v, w): Helmholtz solves + pressure projection, with a
        hand-written adjoint backward.  No intermediates are saved."""

        @staticmethod
        def forward(ctx, fu, fv, fw, alambda, alp, ib, c, mixed):
            ctx.params = (alambda, alp, ib, c)
            u, v, w, _ = projection_step(fu, fv, fw, alambda, alp, ib, c, mixed)
            return u, v, w

        @staticmethod
        def backward(ctx, gu, gv, gw):
            alambda, alp, ib, c = ctx.params
            g = [t.contiguous() if t is not None else None for t in (gu, gv, gw)]
            ref = next(t for t in g if t is not None)
            g = [t if t is not None else torch.zeros_like(ref) for t in g]
            fb = projection_step_adjoint(g[0], g[1], g[2], alambda, alp, ib, c)
            return fb[0], fb[1], fb[2], None, None, None, None, None

    class HelmholtzSolve(torch.autograd.Function):
        """rhs -> sol of one field (helm3d), hand-written adjoint backward."""

        @staticmethod
        def forward(ctx, rhs, bound, alambda, alp, ib):
            ctx.params = (bound, alambda, alp, ib)
            r = rhs.clone()
            sol = torch.empty_like(r)
            helm3d_batch([r], [sol], bound, [alambda], alp=alp, ib=ib)
            return sol

        @staticmethod
        def backward(ctx, g):
            bound, alambda, alp, ib = ctx.params
            out = torch.empty_like(g)
            helm3d_adjoint_batch([g.contiguous()], [out], bound, [alambda], alp=alp, ib=ib)
            return out, None, None, None, None

    return ProjectionStep, HelmholtzSolve


_cls = {}


def projection_step_autograd(fu, fv, fw, alambda: float, alp: float = 1.0, ib: int = 0, c: float = 1.0,
                             mixed: tuple | None = None):
    """Differentiable projection step (torch.autograd)."""
    if not _cls:
        _cls["step"], _cls["solve"] = _autograd_cls()
    return _cls["step"].apply(fu, fv, fw, alambda, alp, ib, c, mixed)


def helmholtz_autograd(rhs, bound: str, alambda: float, alp: float = 1.0, ib: int = 0):
    """Differentiable helm3d of one field (torch.autograd)."""
    if not _cls:
        _cls["step"], _cls["solve"] = _autograd_cls()
    return _cls["solve"].apply(rhs, bound, alambda, alp, ib)
