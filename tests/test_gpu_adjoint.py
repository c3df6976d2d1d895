"""GPU tests of the hand-written adjoints and of gradients through the
projection step (129^2 x 64 modes: J = 0.5 ||u||^2 of the output velocity, gradient
with respect to the forcing f).

Pins: the FP64 adjoint identity <S f, g> = <f, S^T g> (real inner product
Re sum conj(a) b) for every adjoint kernel; central finite differences
(eps = 1e-6) of J through the GPU forward and through an independent FP64
forward built from the oracle (CPU) - the FP64 reference gradient.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    import duct_amd
    torch.cuda.set_device(0)
    duct_amd.init(0)
    yield duct_amd
    duct_amd.shutdown()


_cur = {"m": None}


def ensure(dev, m):
    if _cur["m"] != m:
        dev.setup(m)
        _cur["m"] = m
    return dev.solver._state["setup"]


def rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def inner(a, b):
    return float(np.real(np.vdot(a, b)))


def cuda(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("m,M,ib", [(9, 3, 0), (33, 6, 0), (33, 3, 2), (129, 4, 0), (257, 2, 0)])
@pytest.mark.parametrize("bound,alam", [("d", 1000.0), ("n", 0.0), ("n", 3.0)])
def test_solve_adjoint_identity(dev, m, M, ib, bound, alam):
    import torch
    from duct_amd import adjoint
    ensure(dev, m)
    f, g = rnd((M, m, m), m + M), rnd((M, m, m), 7 * m + M)
    sf = torch.empty_like(cuda(f))
    dev.helm3d_batch([cuda(f)], [sf], bound, [alam], ib=ib)
    stg = torch.empty_like(sf)
    adjoint.helm3d_adjoint_batch([cuda(g)], [stg], bound, [alam], ib=ib)
    torch.cuda.synchronize()
    Sf, STg = sf.cpu().numpy(), stg.cpu().numpy()
    lhs, rhs = inner(Sf, g), inner(f, STg)
    scale = np.linalg.norm(Sf) * np.linalg.norm(g)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)
    assert not STg[:, 0, :].any() and not STg[:, -1, :].any()      # rhs boundary is never read
    if ib == 0:
        assert not STg[0].imag.any()                                  # Im at k=0 is never read


@pytest.mark.parametrize("m", [17, 129, 257])
def test_colloc_deriv_and_derivx(dev, oracle_mod, m):
    """d/dy, d/dz against the oracle's transform-recurrence route (transyz + derivyz + transyz, the
    reference's algorithm, itself pinned against exact differentiation); the transposed product
    through the adjoint identity <D u, g> = <u, D^T g> with the forward product."""
    import torch
    from duct_amd import adjoint
    from oracle import spectral as S
    ensure(dev, m)
    M = 3
    u, g = rnd((M, m, m), m), rnd((M, m, m), m + 1)
    _, DC = dev.basis.collocation_matrix(m)
    scale = np.abs(DC).sum(1).max() * np.max(np.abs(u))
    fwd = {}
    for cdir in ("y", "z"):
        d = S.transyz(S.derivyz(S.transyz(u, cdir, -1), cdir, 1), cdir, +1)
        out = cuda(g.copy())
        adjoint.colloc_deriv(cuda(u), out, cdir, trans=False, scal=0.5, beta=2.0)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.max(np.abs(got - (0.5 * d + 2.0 * g))) <= 1e-14 * scale, cdir
        fwd[cdir] = (got - 2.0 * g) / 0.5
        out = cuda(np.zeros_like(g))
        adjoint.colloc_deriv(cuda(g), out, cdir, trans=True, scal=1.0, beta=0.0)
        torch.cuda.synchronize()
        lhs = np.vdot(fwd[cdir], g)
        rhs = np.vdot(u, out.cpu().numpy())
        assert abs(lhs - rhs) <= 1e-13 * scale * np.abs(g).sum(), cdir
    out = cuda(g.copy())
    adjoint.derivx(cuda(u), out, ib=2, alp=0.5, scal=3.0, beta=-1.0)
    torch.cuda.synchronize()
    k = np.arange(2, 2 + M)[:, None, None]
    assert np.allclose(out.cpu().numpy(), 3.0 * 1j * 0.5 * k * u - g, rtol=0, atol=1e-13 * np.max(np.abs(u)) * 10)


# ---- FP64 reference step built from the oracle (CPU) -----------------------------
def oracle_step(oracle_mod, ob, f3, alam, alp, ib, c):
    DC = ob.b.DC
    us = [oracle_mod.helm3d(ob, f, "d", alam, alp=alp, ib=ib)[0] for f in f3]
    k = (ib + np.arange(f3[0].shape[0]))[:, None, None] * alp
    d = (1j * k * us[0] + us[1] @ DC.T + np.einsum("ij,kjl->kil", DC, us[2])) / c
    phi = oracle_mod.helm3d(ob, d, "n", 0.0, alp=alp, ib=ib)[0]
    return (us[0] - c * 1j * k * phi, us[1] - c * (phi @ DC.T), us[2] - c * np.einsum("ij,kjl->kil", DC, phi))


def J(u3):
    return 0.5 * sum(float(np.sum(np.abs(u) ** 2)) for u in u3)


@pytest.mark.parametrize("m,M,ib", [(17, 4, 0), (33, 6, 1)])
def test_projection_gradient_fd_and_oracle(dev, oracle_mod, m, M, ib):
    import torch
    from duct_amd import adjoint
    ensure(dev, m)
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    alam, alp, c = 800.0, 0.8, 0.37
    f3 = [rnd((M, m, m), 100 * m + i) for i in range(3)]
    ft = [cuda(f).requires_grad_(True) for f in f3]
    u = adjoint.projection_step_autograd(*ft, alam, alp, ib, c)
    loss = 0.5 * sum((x.abs() ** 2).sum() for x in u)
    loss.backward()
    grads = [t.grad.cpu().numpy() for t in ft]
    # the GPU forward matches the oracle forward
    ref_u = oracle_step(oracle_mod, ob, f3, alam, alp, ib, c)
    for a, b in zip(u, ref_u):
        assert np.max(np.abs(a.detach().cpu().numpy() - b)) <= 1e-12 * np.max(np.abs(b))
    eps = 1e-6
    for s in range(3):
        dirs = [rnd((M, m, m), 999 + 10 * s + i) for i in range(3)]
        dd = sum(inner(g, d) for g, d in zip(grads, dirs))
        with torch.no_grad():
            up = adjoint.projection_step(*[cuda(f + eps * d) for f, d in zip(f3, dirs)], alam, alp, ib, c)[:3]
            um = adjoint.projection_step(*[cuda(f - eps * d) for f, d in zip(f3, dirs)], alam, alp, ib, c)[:3]
        fd_gpu = (J([x.cpu().numpy() for x in up]) - J([x.cpu().numpy() for x in um])) / (2 * eps)
        fd_ref = (J(oracle_step(oracle_mod, ob, [f + eps * d for f, d in zip(f3, dirs)], alam, alp, ib, c)) -
                  J(oracle_step(oracle_mod, ob, [f - eps * d for f, d in zip(f3, dirs)], alam, alp, ib, c))) / (2 * eps)
        print("directional derivative: adjoint %.12e  FD(gpu) %.12e  FD(oracle) %.12e" % (dd, fd_gpu, fd_ref),
              flush=True)
        assert abs(dd - fd_gpu) <= 1e-6 * abs(dd)
        assert abs(dd - fd_ref) <= 1e-6 * abs(dd)


def test_config4_gradient_fd_and_oracle(dev, oracle_mod):
    """129^2 x 64 modes (the adjoint workload: seeded forcing, alambda = 1000, alp = 1, c = 0.5), FP64:
    the hand-written gradient of J = 0.5 ||u||^2 through the projection step against central
    finite differences (eps = 1e-6) through the GPU forward and through the FP64 forward built from
    the oracle, along three directions, and the GPU forward against the oracle's."""
    import torch
    from duct_amd import adjoint, workloads
    cfg = workloads.ADJOINT_129
    m, M = cfg.m, cfg.nmodes
    ensure(dev, m)
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    alam, alp, ib, c = 1000.0, 1.0, 0, 0.5
    f3 = [workloads.rhs_field(cfg, i) for i in range(3)]
    ft = [cuda(f).requires_grad_(True) for f in f3]
    u = adjoint.projection_step_autograd(*ft, alam, alp, ib, c)
    loss = 0.5 * sum((x.abs() ** 2).sum() for x in u)
    loss.backward()
    grads = [t.grad.cpu().numpy() for t in ft]
    from oracle import spectral as S
    oracle_mod.use_blas(True)
    try:
        # d/dy and d/dz by the oracle's transform route (REDFT00 + coefficient recurrence): at m = 129
        # the textbook collocation matrix's rounding (~1e-10 here) would dominate the comparison
        def ostep(fs):
            us = [oracle_mod.helm3d(ob, f, "d", alam, alp=alp, ib=ib, nthreads=8)[0] for f in fs]
            k = (ib + np.arange(M))[:, None, None] * alp
            d = (1j * k * us[0] + S._dy(us[1]) + S._dz(us[2])) / c
            phi = oracle_mod.helm3d(ob, d, "n", 0.0, alp=alp, ib=ib, nthreads=8)[0]
            return (us[0] - c * 1j * k * phi, us[1] - c * S._dy(phi), us[2] - c * S._dz(phi))
        ref_u = ostep(f3)
        for a, b in zip(u, ref_u):
            e = np.max(np.abs(a.detach().cpu().numpy() - b)) / np.max(np.abs(b))
            print("129^2x64 forward vs oracle: %.2e" % e, flush=True)
            assert e <= 1e-11
        # directions with the forcing's spectrum (seeded like it, other seeds): J is quadratic, so the
        # central difference is exact up to rounding, ~eps_machine J / eps; smooth directions keep
        # the directional derivative O(J) and that rounding far below the 1e-6 bar (white-noise
        # directions are damped by the solve to ~1e-4 J, where it reaches ~1e-6)
        eps = 1e-6
        for s in range(3):
            dcfg = workloads.Config(40 + s, m, M, 3)       # the generator with other seeds
            dirs = [workloads.rhs_field(dcfg, i) for i in range(3)]
            dd = sum(inner(g, d) for g, d in zip(grads, dirs))
            with torch.no_grad():
                up = adjoint.projection_step(*[cuda(f + eps * d) for f, d in zip(f3, dirs)], alam, alp, ib, c)[:3]
                um = adjoint.projection_step(*[cuda(f - eps * d) for f, d in zip(f3, dirs)], alam, alp, ib, c)[:3]
            fd_gpu = (J([x.cpu().numpy() for x in up]) - J([x.cpu().numpy() for x in um])) / (2 * eps)
            fd_ref = (J(ostep([f + eps * d for f, d in zip(f3, dirs)])) -
                      J(ostep([f - eps * d for f, d in zip(f3, dirs)]))) / (2 * eps)
            print("129^2x64 directional derivative: adjoint %.12e  FD(gpu) %.12e  FD(oracle) %.12e" %
                  (dd, fd_gpu, fd_ref), flush=True)
            assert abs(dd - fd_gpu) <= 1e-6 * abs(dd)
            assert abs(dd - fd_ref) <= 1e-6 * abs(dd)
    finally:
        oracle_mod.use_blas(False)


def test_config4_step_adjoint_identity(dev):
    """129^2 x 64 modes, FP64: <step(f), g> = <f, step^T(g)> for the full step."""
    import torch
    from duct_amd import adjoint, workloads
    cfg = workloads.ADJOINT_129
    ensure(dev, cfg.m)
    f3 = [workloads.rhs_field(cfg, i) for i in range(3)]
    g3 = [rnd((cfg.nmodes, cfg.m, cfg.m), 40 + i) for i in range(3)]
    with torch.no_grad():
        u3 = adjoint.projection_step(*[cuda(f) for f in f3], 1000.0, 1.0, 0, 0.5)[:3]
        fb = adjoint.projection_step_adjoint(*[cuda(g) for g in g3], 1000.0, 1.0, 0, 0.5)
    lhs = sum(inner(u.cpu().numpy(), g) for u, g in zip(u3, g3))
    rhs = sum(inner(f, b.cpu().numpy()) for f, b in zip(f3, fb))
    scale = np.sqrt(sum(np.linalg.norm(u.cpu().numpy()) ** 2 for u in u3)) * \
        np.sqrt(sum(np.linalg.norm(g) ** 2 for g in g3))
    print("129^2x64 step adjoint identity: %.15e vs %.15e (scale %.3e)" % (lhs, rhs, scale), flush=True)
    assert abs(lhs - rhs) <= 1e-11 * scale


def test_config4_mixed_forward_gradient(dev):
    """257^2 x 128 modes: gradient of J through the mixed-precision forward
    (bf16x6, one refinement) against the same gradient through the FP64 forward."""
    import torch
    from duct_amd import adjoint, workloads
    cfg = workloads.ADJOINT_129
    m, M = 257, 128
    ensure(dev, m)
    f3 = [workloads.rhs_field(cfg, i, m=m, nmodes=M) for i in range(3)]
    grads = {}
    for name, mixed in (("fp64", None), ("mixed", ("bf16x6", "bf16x6", 1))):
        ft = [cuda(f).requires_grad_(True) for f in f3]
        u = adjoint.projection_step_autograd(*ft, 1000.0, 1.0, 0, 0.5, mixed)
        loss = 0.5 * sum((x.abs() ** 2).sum() for x in u)
        loss.backward()
        grads[name] = [t.grad.cpu().numpy() for t in ft]
        del ft, u, loss
    err = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(grads["mixed"], grads["fp64"]))
    print("257^2x128 mixed-forward gradient rel. diff vs FP64 forward: %.2e" % err, flush=True)
    assert err <= 1e-9
