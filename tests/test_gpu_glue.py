"""GPU tests of the transpose / pencil-transpose / invdiag kernels: exact
permutations and IEEE divisions, so the bar is bit-exactness against numpy."""
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


@pytest.mark.parametrize("Ny,Nx,nb", [(256, 256, 1), (255, 255, 3), (33, 7, 2), (1, 1, 1), (130, 64, 5)])
def test_transpose_gpu(dev, Ny, Nx, nb):
    import torch
    from duct_amd import glue
    rng = np.random.default_rng(Nx * 1000 + Ny)
    A = rng.standard_normal((nb, Ny, Nx))
    C = glue.transpose_gpu(torch.from_numpy(A).cuda())
    assert np.array_equal(C.cpu().numpy(), np.swapaxes(A, -1, -2))
    C2 = glue.transpose_gpu(torch.from_numpy(A).cuda(), alpha=-2.0)
    assert np.array_equal(C2.cpu().numpy(), -2.0 * np.swapaxes(A, -1, -2))


@pytest.mark.parametrize("nx,my,nz", [(65, 129, 17), (1, 33, 33), (43, 257, 3)])
def test_pencil_transposes(dev, nx, my, nz):
    """phys2(j,k,i) = phys(i,j,k) (parallel.F:476-501) and back (:504-529)."""
    import torch
    from duct_amd import glue
    rng = np.random.default_rng(nx + my + nz)
    phys = rng.standard_normal((nz, my, nx)) + 1j * rng.standard_normal((nz, my, nx))
    p2 = glue.cxyz2yzx(torch.from_numpy(phys).cuda())
    # Fortran phys(i,j,k) = numpy phys[k, j, i]; phys2(j,k,i) = numpy p2[i, k, j]
    ref = np.transpose(phys, (2, 0, 1))
    assert np.array_equal(p2.cpu().numpy(), ref)
    back = glue.cyzx2xyz(p2)
    assert np.array_equal(back.cpu().numpy(), phys)


def test_mat_invdiag_gpu(dev, oracle_mod):
    import torch
    from duct_amd import glue
    b = oracle_mod.pseudomat(33)
    rng = np.random.default_rng(4)
    F = rng.standard_normal(b.ED2.shape)
    for xl in (1000.0, 0.5, -3.0):
        C = glue.mat_invdiag_gpu(torch.from_numpy(F).cuda(), torch.from_numpy(np.ascontiguousarray(b.ED2)).cuda(), xl)
        assert np.array_equal(C.cpu().numpy(), F / (b.ED2 - xl))


def test_glue_rejects_bad_arguments(dev):
    import torch
    from duct_amd._lib import DuctError, call, byref_int, byref_double, byref_ptr
    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    with pytest.raises(DuctError):   # in place
        call("duct_transpose_dev_", byref_int(4), byref_int(4), byref_int(1), byref_double(1.0),
             byref_ptr(x.data_ptr()), byref_ptr(x.data_ptr()))
    with pytest.raises(DuctError):   # null device pointer
        call("duct_mat_invdiag_dev_", byref_int(4), byref_int(4), byref_double(1.0), byref_ptr(0),
             byref_ptr(x.data_ptr()), byref_ptr(x.data_ptr()))
