"""GPU tests of the advection term (duct_advection_dev_) and of the pseudo-spectral
step (192 x-points, 64 modes, 129^2 CGL points), against the oracle,
which follows the reference's own route (FFTW-convention x transforms, REDFT00
Chebyshev transforms and the derivyz recurrence, oracle/spectral.py) where the
GPU uses dense x-DFTs on the matrix cores (up to 64 modes and 192 points, batched
rocFFT transforms past that) and collocation products.

Bars: advection 1e-12 relative (max norm per component); the full step 1e-10
(it contains the k=0 pressure-Poisson solve, ~1e-11 conditioning at this size).
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


def cuda(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def random_velocity(m, M, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        f = (rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))) * \
            np.exp(-0.1 * np.arange(M))[:, None, None]
        f[0] = f[0].real
        out.append(f)
    return out


def run_adv(dev, fields, mgalx, alp, aspect=1.0):
    import torch
    from duct_amd import step
    u = [cuda(f) for f in fields]
    adv = [torch.zeros_like(u[0]) for _ in range(3)]
    step.advection(*u, *adv, mgalx=mgalx, alp=alp, aspect=aspect)
    torch.cuda.synchronize()
    return [a.cpu().numpy() for a in adv]


# m = 257 and 288 run the folded derivative products with three row tiles per wave (the largest
# instantiation; 288 is the largest m of the derivative kernels)
# mgalx = 192 with up to 64 modes runs the lane-FFT product kernel (xlane.hip): 129^2 x 64 (the
# pseudo-spectral workload), fewer modes than the transform length (48, 17), and line counts that
# leave the last workgroup ragged (17^2, 33^2)
@pytest.mark.parametrize("m,M,mgalx,alp", [(17, 16, 48, 1.3), (33, 32, 96, 1.0), (40, 16, 64, 0.7), (129, 64, 192, 1.0),
                                          (33, 48, 192, 1.1), (17, 64, 192, 0.9), (65, 17, 192, 1.0),
                                          (257, 16, 48, 1.0), (288, 16, 48, 0.8)])
def test_advection_matches_oracle(dev, m, M, mgalx, alp):
    from oracle import spectral
    ensure(dev, m)
    fields = random_velocity(m, M, m + M)
    got = run_adv(dev, fields, mgalx, alp)
    ref = spectral.advection3d(*fields, mgalx=mgalx, alp=alp)
    for c in range(3):
        e = relerr(got[c], ref[c])
        print("m=%d M=%d mgalx=%d adv%d rel.err %.2e" % (m, M, mgalx, c + 1, e), flush=True)
        assert e <= 1e-12


def test_advection_aspect_and_errors(dev):
    from duct_amd._lib import DuctError
    from oracle import spectral
    m, M, mgalx = 17, 16, 48
    ensure(dev, m)
    fields = random_velocity(m, M, 3)
    got = run_adv(dev, fields, mgalx, 1.0, aspect=2.0)
    ref = spectral.advection3d(*fields, mgalx=mgalx, alp=1.0, aspect=2.0)
    for c in range(3):
        assert relerr(got[c], ref[c]) <= 1e-12
    with pytest.raises(DuctError):
        run_adv(dev, fields, 47, 1.0)                               # odd mgalx
    with pytest.raises(DuctError):
        run_adv(dev, fields, 40, 1.0)                               # mgalx < 3 nmodes
    import torch
    from duct_amd import step
    u = [cuda(f) for f in fields]
    with pytest.raises(DuctError):
        step.advection(u[0], u[1], u[2], u[0], torch.zeros_like(u[0]), torch.zeros_like(u[0]), mgalx=48)


# Past the dense x-DFT kernels (M % 16 != 0, M > 64 or mgalx > 192): batched rocFFT real transforms
# along x with the fill, the products in physical space and the pack, then the same derivative terms.
# The last case is the full solve workload's grid: 257^2 x 256 modes, mgalx 768.
@pytest.mark.parametrize("m,M,mgalx,alp", [(17, 8, 24, 1.1), (33, 20, 64, 1.0), (65, 96, 288, 0.9),
                                          (257, 256, 768, 1.0)])
def test_advection_large_modes_matches_oracle(dev, m, M, mgalx, alp):
    from oracle import spectral
    ensure(dev, m)
    fields = random_velocity(m, M, 7 * m + M)
    print("advection m=%d M=%d mgalx=%d (rocFFT x-transforms)" % (m, M, mgalx), flush=True)
    got = run_adv(dev, fields, mgalx, alp)
    ref = spectral.advection3d(*fields, mgalx=mgalx, alp=alp)
    for c in range(3):
        e = relerr(got[c], ref[c])
        print("m=%d M=%d mgalx=%d adv%d rel.err %.2e" % (m, M, mgalx, c + 1, e), flush=True)
        assert e <= 1e-12


def test_config3_step(dev, oracle_mod):
    """129^2 x 64 modes, mgalx 192: laminar duct + 5% divergence-free perturbation; advection, three Helmholtz
    solves, the pressure Poisson solve and the projection, against the oracle."""
    import torch
    from duct_amd import step, workloads
    from oracle import spectral as S
    cfg = workloads.STEP_129
    m, M, mgalx, nu, dt = cfg.m, cfg.nmodes, 192, 0.01, 0.1
    ensure(dev, m)
    vel = workloads.duct_velocity(cfg)
    ut = [cuda(f) for f in vel]
    un, vn, wn, phi, adv = step.c3_step(*ut, nu=nu, dt=dt, mgalx=mgalx)
    torch.cuda.synchronize()
    got = [t.cpu().numpy() for t in (un, vn, wn)]
    # oracle step
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    alam = 1.0 / (nu * dt)
    radv = S.advection3d(*vel, mgalx=mgalx, alp=cfg.alp)
    for c in range(3):
        e = relerr(adv[c].cpu().numpy(), radv[c])
        print("129^2x64 adv%d rel.err %.2e" % (c + 1, e), flush=True)
        assert e <= 1e-12
    oracle_mod.use_blas(True)
    try:
        us = [oracle_mod.helm3d(ob, a / nu - alam * x, "d", alam, nthreads=8)[0] for a, x in zip(radv, vel)]

        def dy(f):
            return S.transyz(S.derivyz(S.transyz(f, "y", -1), "y", 1), "y", +1)

        def dz(f):
            return S.transyz(S.derivyz(S.transyz(f, "z", -1), "z", 1), "z", +1)

        d = (S.derivx(us[0], cfg.alp) + dy(us[1]) + dz(us[2])) / dt
        ph = oracle_mod.helm3d(ob, d, "n", 0.0, nthreads=8)[0]
    finally:
        oracle_mod.use_blas(False)
    ref = (us[0] - dt * S.derivx(ph, cfg.alp), us[1] - dt * dy(ph), us[2] - dt * dz(ph))
    for c in range(3):
        e = relerr(got[c], ref[c])
        print("129^2x64 step u%d rel.err %.2e" % (c + 1, e), flush=True)
        assert e <= 1e-10
    # the projection leaves the interior divergence the reference's own route leaves: its Poisson
    # operator D2B comes from the textbook DC (pseudomat.f90:90-110) while div3d and the gradient
    # take the transform route, and the two second derivatives differ by ~1e-5 relative at
    # m = 129, so div(u_new) / div(u*) is ~2e-8 there (1e-10 with the textbook DC everywhere)
    div = S.div3d(got[0], got[1], got[2], 1.0, cfg.alp)      # div3d (operators.F:427-501)
    div_ref = S.div3d(ref[0], ref[1], ref[2], 1.0, cfg.alp)
    dstar = d * dt
    rel = np.abs(div[:, 1:-1, 1:-1]).max() / np.abs(dstar[:, 1:-1, 1:-1]).max()
    rel_ref = np.abs(div_ref[:, 1:-1, 1:-1]).max() / np.abs(dstar[:, 1:-1, 1:-1]).max()
    print("129^2x64 interior divergence after projection / before: %.2e (oracle route %.2e)" % (rel, rel_ref),
          flush=True)
    assert rel <= 1.05 * rel_ref + 1e-12 and rel_ref <= 5e-8


@pytest.mark.parametrize("m", [33, 129])
def test_transyz_and_derivyz_match_oracle(dev, m):
    """duct_transyz_dev_ (dense transforms) and duct_derivyz_dev_ (the recurrence) against the
    oracle's REDFT00 transforms and recurrence; transform + derivative + inverse against DC."""
    import torch
    from duct_amd import step, basis
    from oracle import spectral as S
    ensure(dev, m)
    M = 3
    rng = np.random.default_rng(m)
    f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    for cdir in ("y", "z"):
        for iopt in (-1, 1):
            out = torch.zeros_like(cuda(f))
            step.transyz(cuda(f), out, cdir, iopt)
            torch.cuda.synchronize()
            ref = S.transyz(f, cdir, iopt)
            e = relerr(out.cpu().numpy(), ref)
            print("m=%d transyz %s %+d rel.err %.2e" % (m, cdir, iopt, e), flush=True)
            assert e <= 1e-13
        for ider in (1, 2):
            du = torch.zeros_like(cuda(f))
            step.derivyz(cuda(f), du, cdir, ider=ider, scal=0.5, aspect=1.5)
            g0 = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
            acc = cuda(g0)
            step.derivyz(cuda(f), acc, cdir, ider=ider, scal=0.5, what="+", aspect=1.5)
            torch.cuda.synchronize()
            ref = S.derivyz(f, cdir, ider, scal=0.5, aspect=1.5)
            e = relerr(du.cpu().numpy(), ref)
            e2 = relerr(acc.cpu().numpy(), ref + g0)
            print("m=%d derivyz %s ider=%d rel.err %.2e (accumulate %.2e)" % (m, cdir, ider, e, e2), flush=True)
            assert e <= 1e-13 and e2 <= 1e-13
    # the reference's route of a physical derivative equals the collocation product
    c = torch.zeros_like(cuda(f))
    d = torch.zeros_like(c)
    p = torch.zeros_like(c)
    step.transyz(cuda(f), c, "y", -1)
    step.derivyz(c, d, "y", 1)
    step.transyz(d, p, "y", 1)
    torch.cuda.synchronize()
    _, DC = basis.collocation_matrix(m)
    want = f @ DC.T
    assert np.max(np.abs(p.cpu().numpy() - want)) <= 1e-12 * np.abs(DC).sum(1).max() * np.abs(f).max()


@pytest.mark.parametrize("my,nz,M,mgalx", [(9, 5, 16, 48), (129, 7, 64, 192), (17, 9, 100, 300), (257, 33, 256, 768),
                                          (12, 3, 5, 10)])
def test_fourx_matches_oracle(dev, my, nz, M, mgalx):
    """duct_fourx_dev_ against numpy's FFT restatement of fourx (fill, c2r, r2c, pack_fftw); the
    dense kernel up to 64 modes and 192 points (multiples of 16), rocFFT past that."""
    import torch
    from duct_amd import step
    from oracle import spectral as S
    rng = np.random.default_rng(my + M)
    c = rng.standard_normal((M, nz, my)) + 1j * rng.standard_normal((M, nz, my))   # [mode, z, y]
    fou = cuda(np.ascontiguousarray(np.moveaxis(c, 0, -1)))                         # [z, y, mode]
    phys = torch.full((nz, my, mgalx + 2), 7.0, dtype=torch.float64, device="cuda")
    step.fourx(fou, phys, mgalx, +1)
    torch.cuda.synchronize()
    ref = S.fourx_inverse(c, mgalx)                                                 # [z, y, x]
    got = phys.cpu().numpy()
    assert relerr(got[..., :mgalx], ref) <= 1e-13 and not got[..., mgalx:].any()
    p2 = rng.standard_normal((nz, my, mgalx))
    ph = cuda(np.concatenate([p2, np.zeros((nz, my, 2))], axis=-1))
    back = torch.zeros_like(fou)
    step.fourx(back, ph, mgalx, -1)
    torch.cuda.synchronize()
    want = np.moveaxis(S.fourx_forward(p2, M), 0, -1)
    assert relerr(back.cpu().numpy(), want) <= 1e-13
