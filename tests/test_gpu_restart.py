"""A restart from a snapshot file. The fields of a reference-layout uvwp HDF5 file
(tests/golden/uvwp_m25_poly.h5, written by the HDF5 library) are read by the library's reader
(duct_h5_read_field_: 25^2 interpolated onto 33^2 as cheb_interp does; the file's 2 modes followed by
14 zero modes, as read_4fields_fou_hdf5 leaves modes past the file's mx1) and uploaded. They then
advance one full RK3 time step (duct_rk3_substep_dev_, krk = 1, 2, 3, with the file's alp and
aspect = 1.5), compared with the oracle's step from the same arrays. Bar: 1e-10 relative, max norm
per field, as for the other RK3 tests."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def dev():
    import torch
    import duct_amd
    torch.cuda.set_device(0)
    duct_amd.init(0)
    yield duct_amd
    duct_amd.shutdown()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_restart_from_file_rk3_step(dev, oracle_mod):
    import torch
    from duct_amd import h5io, step
    from oracle import rk3 as R
    path = os.path.join(GOLD, "uvwp_m25_poly.h5")
    m, M, mgalx, fnu, dt = 33, 16, 48, 0.01, 0.002
    h = h5io.read_header(path)
    alp, aspect = h["alp"], h["aspect"]
    fields = h5io.read_fields(path, my=m, nmodes=M)
    assert sorted(fields) == sorted("uvwp")
    u, v, w, p = (fields[c] for c in "uvwp")
    assert u.shape == (M, m, m) and np.all(u[2:] == 0) and np.all(u[0].imag == 0)
    print("restart m=%d M=%d from %s (file grid %d^2, mx1=%d)" % (m, M, os.path.basename(path), h["my"], h["mx1"]),
          flush=True)
    adv = [np.zeros_like(u) for _ in range(3)]
    dpdx = R.get_dpdx(aspect, fnu)
    dev.setup(m)
    F = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in (u, v, w, p)]
    A = [torch.from_numpy(a.copy()).cuda() for a in adv]
    got = []
    for krk in (1, 2, 3):
        step.rk3_substep(*F, A, krk, fnu, dt, mgalx, alp=alp, aspect=aspect, dpdx=dpdx)
        torch.cuda.synchronize()
        got.append(([f.cpu().numpy() for f in F], [a.cpu().numpy() for a in A]))
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))

    def helm3d(rhs, bound, alam):
        return oracle_mod.helm3d(ob, rhs, bound, alam, alp=alp, nthreads=8)[0]

    a = tuple(adv)
    for i, krk in enumerate((1, 2, 3)):
        u, v, w, p, a, _ = R.rk3_substep(u, v, w, p, a, krk, fnu, dt, mgalx, helm3d, alp, aspect=aspect, dpdx=dpdx)
        gf, ga = got[i]
        errs = [relerr(g, r) for g, r in zip(gf, (u, v, w, p))] + [relerr(g, r) for g, r in zip(ga, a)]
        print("restart krk=%d rel.err u v w p adv1..3: %s" % (krk, " ".join("%.1e" % e for e in errs)), flush=True)
        assert max(errs) <= 1e-10
