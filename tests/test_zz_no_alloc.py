"""GPU test of the allocation-free entry duct_helm3d_batch_noalloc_dev_ (scratch in a static device
array), called through the C-ABI: its solutions equal duct_helm3d_batch_dev_'s (unfolded FP64 passes)
to the bit, on two calls in a row, the second with a smaller input; a batch larger than the array is
rejected with an error."""
import ctypes

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
    duct_amd.set_fold(0)
    duct_amd.shutdown()


def field(M, m, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))


def call_noalloc(rhs, sol, bounds, alams, ib=0, alp=1.0):
    from duct_amd import _lib
    nf = len(rhs)
    R = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in rhs])
    S = (ctypes.c_size_t * nf)(*[t.data_ptr() for t in sol])
    A = (ctypes.c_double * nf)(*[float(a) for a in alams])
    m, M = rhs[0].shape[-1], rhs[0].shape[0]
    _lib.call("duct_helm3d_batch_noalloc_dev_", _lib.byref_int(nf), R, S, bounds.encode(), A, _lib.byref_int(m),
              _lib.byref_int(M), _lib.byref_int(ib), _lib.byref_double(alp), ctypes.c_size_t(nf))


@pytest.mark.parametrize("m", [129, 257])
def test_noalloc_equals_batch_two_calls(dev, m):
    import torch
    dev.setup(m)
    dev.set_fold(0)
    bounds, alams = "dddn", [1000.0, 1000.0, 250.0, 0.0]
    # first call: the largest batch that fits, then a smaller one (fewer modes and fields)
    M1 = 24 if m == 129 else 6
    cases = [(bounds, alams, M1, 0, 1.0), ("dn", [1000.0, 0.0], 3, 2, 0.7)]
    for i, (bd, al, M, ib, alp) in enumerate(cases):
        print("noalloc call %d: m=%d bounds=%s modes=%d ib=%d" % (i, m, bd, M, ib), flush=True)
        fs = [field(M, m, 100 * i + f) for f in range(len(bd))]
        rhs_a = [torch.from_numpy(x).cuda() for x in fs]
        rhs_b = [torch.from_numpy(x).cuda() for x in fs]
        sol_a = [torch.zeros_like(r) for r in rhs_a]
        sol_b = [torch.full_like(r, float("nan")) for r in rhs_b]   # every output point is written
        dev.helm3d_batch(rhs_a, sol_a, bd, al, alp=alp, ib=ib)
        torch.cuda.synchronize()
        call_noalloc(rhs_b, sol_b, bd, al, ib=ib, alp=alp)
        torch.cuda.synchronize()
        for f in range(len(bd)):
            a, b = sol_a[f].cpu().numpy(), sol_b[f].cpu().numpy()
            assert np.array_equal(a, b), (i, f, np.nanmax(np.abs(a - b)))
            assert np.array_equal(rhs_a[f].cpu().numpy(), rhs_b[f].cpu().numpy())   # same in-place zeroing


def test_noalloc_rejects_a_batch_larger_than_its_scratch(dev):
    import torch
    from duct_amd._lib import DuctError
    m = 257
    dev.setup(m)
    x = torch.zeros((40, m, m), dtype=torch.complex128, device="cuda")
    y = torch.zeros_like(x)
    with pytest.raises(DuctError, match="allocation-free entry holds"):
        call_noalloc([x], [y], "d", [1000.0])
