"""GPU tests of the mixed-boundary (temperature) solve (duct_setup_temp_ / duct_helm3d_temp_dev_,
helm3d_temp.f90 / solvediag_temp.f90) against the oracle's restatement (oracle/temp.py), for
wall combinations of every kind, rectangular domains (sy != sz), my != mz grids and several
sizes; plus the reference's semantics: the rhs is not modified, the corners keep the rhs's
values, Im is zero at k = 0.  Bar: 1e-11 relative (max norm per field)."""
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
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def case(m, M, seed, mz=None):
    mz = m if mz is None else mz
    rng = np.random.default_rng(seed)
    rhs = rng.standard_normal((M, mz, m)) + 1j * rng.standard_normal((M, mz, m))
    bcy = rng.standard_normal((M, 2, mz))
    bcz = rng.standard_normal((M, 2, m))
    return rhs, bcy, bcz


@pytest.mark.parametrize("m,bounds,sy,sz,alam", [
    (17, "dddd", 1.0, 1.0, 40.0), (17, "nnnn", 1.0, 1.0, 0.5), (33, "dndn", 1.0, 2.0, 10.0),
    (33, "ndnd", 1.5, 1.0, 3.0), (65, "ddnn", 1.0, 1.0, 100.0), (65, "nndd", 2.0, 1.0, 1.0),
    (129, "dnnd", 1.0, 1.0, 1000.0), (34, "dnnn", 1.0, 1.0, 2.0)])
def test_helm3d_temp_matches_oracle(dev, m, bounds, sy, sz, alam):
    import torch
    from duct_amd import temperature
    from oracle import temp as T
    M, alp = 5, 0.8
    rhs, bcy, bcz = case(m, M, m + len(bounds))
    temperature.setup_temp(m, bounds, sy, sz)
    r = cuda(rhs)
    sol = torch.full_like(r, 7.0)
    temperature.helm3d_temp(r, sol, cuda(bcy), cuda(bcz), alam, alp=alp)
    torch.cuda.synchronize()
    got = sol.cpu().numpy()
    ref = T.helm3d_temp(T.TempBasis(m, bounds, sy, sz), rhs, alam, bcy, bcz, alp=alp)
    e = relerr(got, ref)
    print("m=%d %s sy=%.1f sz=%.1f alambda=%g rel.err %.2e" % (m, bounds, sy, sz, alam, e), flush=True)
    assert e <= 1e-11
    assert np.array_equal(r.cpu().numpy(), rhs)                                    # rhs kept
    assert not got[0].imag.any()                                                   # Im zero at k = 0
    for (a, b) in ((0, 0), (0, -1), (-1, 0), (-1, -1)):                            # corners = rhs corners
        assert np.array_equal(got[1:, a, b], rhs[1:, a, b]) and got[0, a, b] == rhs[0, a, b].real


@pytest.mark.parametrize("my,mz,bounds,sy,sz,alam", [
    (33, 49, "dndn", 1.0, 2.0, 10.0), (49, 33, "ndnd", 1.0, 1.0, 3.0), (17, 65, "nnnn", 1.5, 1.0, 0.5),
    (65, 17, "dddd", 1.0, 1.0, 40.0), (129, 65, "dnnd", 1.0, 2.0, 1000.0), (34, 130, "nndd", 2.0, 1.0, 1.0)])
def test_helm3d_temp_rectangular_grid(dev, my, mz, bounds, sy, sz, alam):
    """my != mz: the y basis of size my and the z basis of size mz (helm3d_temp.f90:35-36), both
    padded to the larger interior in the projection passes; fields (my, mz, nmodes), bcy over z,
    bcz over y."""
    import torch
    from duct_amd import temperature
    from oracle import temp as T
    M, alp = 4, 0.8
    rhs, bcy, bcz = case(my, M, my * 1000 + mz, mz=mz)
    print("my=%d mz=%d %s" % (my, mz, bounds), flush=True)
    temperature.setup_temp(my, bounds, sy, sz, mz=mz)
    r = cuda(rhs)
    sol = torch.full_like(r, 7.0)
    temperature.helm3d_temp(r, sol, cuda(bcy), cuda(bcz), alam, alp=alp)
    torch.cuda.synchronize()
    got = sol.cpu().numpy()
    ref = T.helm3d_temp(T.TempBasis(my, bounds, sy, sz, mz=mz), rhs, alam, bcy, bcz, alp=alp)
    e = relerr(got, ref)
    print("my=%d mz=%d %s rel.err %.2e" % (my, mz, bounds, e), flush=True)
    assert e <= 1e-11
    assert np.array_equal(r.cpu().numpy(), rhs)
    assert not got[0].imag.any()
    for (a, b) in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
        assert np.array_equal(got[1:, a, b], rhs[1:, a, b]) and got[0, a, b] == rhs[0, a, b].real
    with pytest.raises(ValueError):                                                # bcy / bcz swapped
        temperature.helm3d_temp(r, sol, cuda(bcz), cuda(bcy), alam, alp=alp)


def test_helm3d_temp_modes_offset_and_errors(dev):
    import torch
    from duct_amd import temperature
    from duct_amd._lib import DuctError
    from oracle import temp as T
    m, M, ib = 33, 3, 2
    rhs, bcy, bcz = case(m, M, 5)
    temperature.setup_temp(m, "dndd", 1.0, 1.0)
    r = cuda(rhs)
    sol = torch.zeros_like(r)
    temperature.helm3d_temp(r, sol, cuda(bcy), cuda(bcz), 20.0, alp=1.0, ib=ib)
    torch.cuda.synchronize()
    ref = T.helm3d_temp(T.TempBasis(m, "dndd"), rhs, 20.0, bcy, bcz, alp=1.0, ib=ib)
    assert relerr(sol.cpu().numpy(), ref) <= 1e-11
    with pytest.raises(DuctError):
        temperature.helm3d_temp(r, r, cuda(bcy), cuda(bcz), 20.0)                  # rhs aliases sol
    with pytest.raises(ValueError):
        temperature.setup_temp(m, "dndx")
    with pytest.raises(DuctError):                                                 # set up for another m
        temperature.helm3d_temp(cuda(rhs[:, :17, :17]), torch.zeros_like(cuda(rhs[:, :17, :17])),
                                cuda(bcy[:, :, :17]), cuda(bcz[:, :, :17]), 20.0)
    with pytest.raises(DuctError):                                                 # my, mz swapped
        temperature.setup_temp(m, "dndd", 1.0, 1.0, mz=17)
        temperature.helm3d_temp(cuda(rhs[:, :33, :17]), torch.zeros_like(cuda(rhs[:, :33, :17])),
                                cuda(bcy[:, :, :33]), cuda(bcz[:, :, :17]), 20.0)
