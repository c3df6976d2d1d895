"""GPU tests of the Shen-Legendre Galerkin solve (duct_setup_shen_ / duct_helm3d_shen_dev_, the
-DHELM_SHEN path: one-direction diagonalisation and batched tridiagonal solves, solveshen.f90)
against the oracle's independent two-direction diagonalisation of the same Galerkin system
(oracle/shen.py, pinned by manufactured solutions), and against exact manufactured solutions.
Bars: 1e-10 relative against the oracle (two routes through the same discrete system),
1e-11 for manufactured solutions in the Shen space."""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

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


def solve(m, rhs, bound, alam, alp, ib=0):
    import torch
    from duct_amd import shen
    shen.setup_shen(m, bound)
    r = cuda(rhs)
    sol = torch.full_like(r, 3.0)
    shen.helm3d_shen(r, sol, bound, alam, alp=alp, ib=ib)
    torch.cuda.synchronize()
    assert np.array_equal(r.cpu().numpy(), rhs)
    return sol.cpu().numpy()


@pytest.mark.parametrize("m", [17, 33, 65, 129])
@pytest.mark.parametrize("bound,alam", [("d", 1000.0), ("d", 2.0), ("n", 0.0), ("n", 5.0)])
def test_shen_matches_oracle(dev, m, bound, alam):
    from oracle import shen as OS
    M, alp = 3, 1.0
    rng = np.random.default_rng(m)
    x = -np.cos(np.pi * np.arange(m) / (m - 1))
    rhs = np.zeros((M, m, m), complex)
    for k in range(M):
        for part in (1.0, 1j):
            c = rng.standard_normal((8, 8)) * np.exp(-0.4 * np.add.outer(np.arange(8), np.arange(8)))
            rhs[k] += part * (C.chebvander(x, 7) @ c @ C.chebvander(x, 7).T)
    got = solve(m, rhs, bound, alam, alp)
    worst = 0.0
    for k in range(M):
        xl = alam + (alp * k) ** 2
        want = OS.shen_solve(m, rhs[k].real, xl, bound) + (1j * OS.shen_solve(m, rhs[k].imag, xl, bound) if k else 0)
        worst = max(worst, np.abs(got[k] - want).max() / np.abs(want).max())
    print("m=%d %s alambda=%g rel.err vs oracle %.2e" % (m, bound, alam, worst), flush=True)
    assert worst <= 1e-10
    assert not got[0].imag.any()


@pytest.mark.parametrize("m", [33, 129])
def test_shen_manufactured_dirichlet(dev, m):
    """u = (1 - y^2)(1 - z^2) p(y) q(z) lies in the Shen space; f = -lap(u) + xlambda u exactly."""
    x = -np.cos(np.pi * np.arange(m) / (m - 1))
    py = C.chebmul(C.chebsub([1.0], C.chebmul([0, 1], [0, 1])), [0.3, -0.5, 0.2, 0.1])
    pz = C.chebmul(C.chebsub([1.0], C.chebmul([0, 1], [0, 1])), [0.7, 0.1, -0.3])
    u = np.outer(C.chebval(x, pz), C.chebval(x, py))                      # [z, y]
    uyy = np.outer(C.chebval(x, pz), C.chebval(x, C.chebder(py, 2)))
    uzz = np.outer(C.chebval(x, C.chebder(pz, 2)), C.chebval(x, py))
    alam = 7.0
    rhs = np.stack([-(uyy + uzz) + (alam + k * k) * u for k in range(2)]).astype(complex)
    rhs[1] *= (1 + 1j)
    got = solve(m, rhs, "d", alam, 1.0)
    assert np.abs(got[0] - u).max() <= 1e-11 * np.abs(u).max()
    assert np.abs(got[1] - (1 + 1j) * u).max() <= 1e-11 * np.abs(u).max()


def test_shen_errors(dev):
    import torch
    from duct_amd import shen
    from duct_amd._lib import DuctError
    m = 17
    shen.setup_shen(m, "d")
    r = torch.zeros((2, m, m), dtype=torch.complex128, device="cuda")
    with pytest.raises(DuctError):
        shen.helm3d_shen(r, r, "d", 1.0)                                 # aliasing
    with pytest.raises(DuctError):
        shen.helm3d_shen(r, torch.zeros_like(r), "x", 1.0)               # bad bound
    with pytest.raises(ValueError):
        shen.setup_shen(m, "q")
