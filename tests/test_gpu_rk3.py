"""GPU tests of the Runge-Kutta substep (duct_rk3_substep_dev_, main1.f90:384-655) against
the oracle's item-by-item restatement (oracle/rk3.py: FFT / DCT-I / recurrence operators,
the C solvediag).  Bar: 1e-10 relative, max norm per field (the pressure Poisson solve's
conditioning at these sizes is ~1e-11); the laminar steady state to 1e-12."""
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


def cuda(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def oracle_helm(oracle_mod, m, alp):
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))

    def helm3d(rhs, bound, alam):
        return oracle_mod.helm3d(ob, rhs, bound, alam, alp=alp, nthreads=8)[0]
    return helm3d


def initial_state(m, M, seed):
    """The step's velocity (laminar duct + divergence-free perturbation), a smooth pressure and a
    previous advection term."""
    from duct_amd import workloads
    from oracle import spectral as S
    cfg = workloads.STEP_129
    u, v, w = workloads.duct_velocity(cfg, m=m, nmodes=M)
    rng = np.random.default_rng(seed)
    x = -np.cos(np.pi * np.arange(m) / (m - 1))
    k = np.arange(M)[:, None, None]
    p = (rng.standard_normal((M, 1, 1)) + 1j * rng.standard_normal((M, 1, 1))) * np.exp(-0.3 * k) * \
        np.cos(1.3 * x)[None, :, None] * np.sin(0.7 * x + 0.2)[None, None, :] * 1e-2
    p[0] = p[0].real
    adv = S.advection3d(0.9 * u, 1.1 * v, w, mgalx=3 * M)
    return [u, v, w, p], list(adv)


def run_gpu(fields, adv, krks, fnu, dt, mgalx, alp, dpdx):
    import torch
    from duct_amd import step
    F = [cuda(f) for f in fields]
    A = [cuda(a) for a in adv]
    out = []
    for krk in krks:
        step.rk3_substep(*F, A, krk, fnu, dt, mgalx, alp=alp, dpdx=dpdx)
        torch.cuda.synchronize()
        out.append(([f.cpu().numpy() for f in F], [a.cpu().numpy() for a in A]))
    return out


@pytest.mark.parametrize("m,M,mgalx", [(33, 16, 48), (40, 32, 96), (200, 16, 48), (33, 24, 72)])
def test_rk3_step_matches_oracle(dev, oracle_mod, m, M, mgalx):
    """A full time step (krk = 1, 2, 3) from the same state, every substep compared."""
    from oracle import rk3 as R
    fnu, dt, alp = 0.01, 0.02, 1.0
    dev.setup(m)
    fields, adv = initial_state(m, M, m)
    dpdx = R.get_dpdx(1.0, fnu)
    got = run_gpu(fields, adv, (1, 2, 3), fnu, dt, mgalx, alp, dpdx)
    helm3d = oracle_helm(oracle_mod, m, alp)
    u, v, w, p = fields
    a = tuple(adv)
    for i, krk in enumerate((1, 2, 3)):
        u, v, w, p, a, _ = R.rk3_substep(u, v, w, p, a, krk, fnu, dt, mgalx, helm3d, alp, dpdx=dpdx)
        gf, ga = got[i]
        errs = [relerr(g, r) for g, r in zip(gf, (u, v, w, p))] + [relerr(g, r) for g, r in zip(ga, a)]
        print("m=%d M=%d krk=%d rel.err u v w p adv1..3: %s" % (m, M, krk, " ".join("%.1e" % e for e in errs)),
              flush=True)
        assert max(errs) <= 1e-10


def test_rk3_substep_config3(dev, oracle_mod):
    """129^2 CGL, 64 modes, mgalx 192: substep krk = 2 (reads the previous advection)."""
    from oracle import rk3 as R
    m, M, mgalx, fnu, dt, alp = 129, 64, 192, 0.01, 0.01, 1.0
    dev.setup(m)
    fields, adv = initial_state(m, M, 3)
    dpdx = R.get_dpdx(1.0, fnu)
    (gf, ga), = run_gpu(fields, adv, (2,), fnu, dt, mgalx, alp, dpdx)
    oracle_mod.use_blas(True)
    try:
        ref = R.rk3_substep(*fields, tuple(adv), 2, fnu, dt, mgalx, oracle_helm(oracle_mod, m, alp), alp, dpdx=dpdx)
    finally:
        oracle_mod.use_blas(False)
    errs = [relerr(g, r) for g, r in zip(gf + ga, list(ref[:4]) + list(ref[4]))]
    print("129^2x64 krk=2 rel.err u v w p adv1..3: %s" % " ".join("%.1e" % e for e in errs), flush=True)
    assert max(errs) <= 1e-10


def test_rk3_laminar_steady_state(dev, oracle_mod):
    """-(d_yy + d_zz) U = dpval(1) (operators.F:2494-2507) is a steady state of the scheme with
    dpdx = get_dpdx(1, nu): ten time steps leave it unchanged."""
    import torch
    from duct_amd import step
    from oracle import rk3 as R
    m, M, mgalx, fnu, dt = 65, 16, 48, 0.01, 0.05
    dev.setup(m)
    U = np.zeros((M, m, m), dtype=np.complex128)
    U[0] = -R.DPVAL[0]
    U = oracle_helm(oracle_mod, m, 1.0)(U, "d", 0.0)
    F = [cuda(U), cuda(np.zeros_like(U)), cuda(np.zeros_like(U)), cuda(np.zeros_like(U))]
    A = [cuda(np.zeros_like(U)) for _ in range(3)]
    for _ in range(10):
        step.rk3_step(*F, A, fnu, dt, mgalx)
    torch.cuda.synchronize()
    u, v, w, p = (f.cpu().numpy() for f in F)
    print("laminar steady state after 10 steps: |u-U| %.1e |v| %.1e |w| %.1e |p| %.1e" % (
        np.abs(u - U).max(), np.abs(v).max(), np.abs(w).max(), np.abs(p).max()), flush=True)
    assert np.abs(u - U).max() <= 1e-12
    assert max(np.abs(v).max(), np.abs(w).max(), np.abs(p).max()) <= 1e-12


def test_rk3_rejects_bad_arguments(dev):
    import torch
    from duct_amd import step
    from duct_amd._lib import DuctError
    m, M = 33, 16
    dev.setup(m)
    F = [torch.zeros((M, m, m), dtype=torch.complex128, device="cuda") for _ in range(7)]
    with pytest.raises(DuctError):
        step.rk3_substep(*F[:4], F[4:], 4, 0.01, 0.1, 48)              # krk outside 1..3
    with pytest.raises(DuctError):
        step.rk3_substep(*F[:4], [F[4], F[4], F[6]], 1, 0.01, 0.1, 48)  # aliasing
    with pytest.raises(DuctError):
        step.rk3_substep(*F[:4], F[4:], 1, 0.01, 0.1, 40)              # mgalx < 3 nmodes
    with pytest.raises(ValueError):
        step.rk3_substep(*F[:4], F[4:], 1, 0.01, 0.1, 48, aspect=1.2)  # untabulated aspect
