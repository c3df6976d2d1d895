"""GPU tests of the mixed-precision refined solve against the FP64 oracle.

Bar: relative error max|u_mixed - u_oracle| / max|u_oracle| <= 1e-10 per field
(the tolerance the task states for the refined solve), for every precision
scheme with the refinement count it needs; the FP64 residual kernel is checked
against a numpy FP64 residual of the same iterate.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


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


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def run_mixed(dev, fields, bounds, alams, **kw):
    import torch
    rhs = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in fields]
    sol = [torch.zeros_like(r) for r in rhs]
    st = dev.helm3d_mixed_batch(rhs, sol, bounds, alams, stats=True, **kw)
    torch.cuda.synchronize()
    return [s.cpu().numpy() for s in sol], st


# Refinement contracts the error by roughly eps(scheme1) * cond(L); cond(L) grows
# like m^4 and is largest for the alpha = 0 Poisson field, so the number of
# corrections each scheme needs grows with m (measured: at m = 129 one bf16x3
# correction leaves ~1e-9 on the Poisson field, two leave ~1e-11).
SCHEMES = {
    33: [("bf16x3", "bf16x6", 1), ("bf16x3", "bf16x3", 2), ("f32", "f32", 1), ("bf16x6", "bf16x3", 1),
         ("bf16x6", "bf16x6", 1)],
    129: [("bf16x3", "bf16x6", 2), ("bf16x3", "bf16x3", 2), ("f32", "f32", 1), ("bf16x6", "bf16x6", 1),
          ("fold16", "fold16", 1), ("fold16q", "fold16q", 1)],
    # even interiors: n = 128 (both halves 64) and n = 198 (halves of 99, a ragged last group of
    # four points in the fold and unfold)
    130: [("fold16q", "fold16q", 1)],
    200: [("fold16", "fold16", 1), ("fold16q", "fold16q", 1)],
}
# (the folded schemes give the Poisson modes with xlambda < 64 to the unfolded FP64 passes: the batches
# at m >= 129 hold more modes than that, so that the Poisson field has refined modes too)
CASES = [(m, M, s0, s1, it) for m, M in ((33, 16), (129, 12), (130, 10), (200, 10)) for s0, s1, it in SCHEMES[m]]


@pytest.mark.parametrize("m,M,s0,s1,niter", CASES)
def test_mixed_matches_fp64_oracle(dev, oracle_mod, m, M, s0, s1, niter):
    from duct_amd import workloads
    ensure(dev, m)
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    cfg = workloads.SOLVE_257
    fields = [workloads.rhs_field(cfg, f, m=m, nmodes=M) for f in range(4)]
    alams = [workloads.alambda_of(cfg, f) for f in range(4)]
    got, st = run_mixed(dev, fields, workloads.BOUNDS, alams, scheme0=s0, scheme1=s1, niter=niter)
    oracle_mod.use_blas(True)
    try:
        for f in range(4):
            ref, _ = oracle_mod.helm3d(ob, fields[f], workloads.BOUNDS[f], alams[f], nthreads=8)
            e = relerr(got[f], ref)
            print("m=%d %s->%s x%d field %s: rel.err %.2e, residual history %s" % (
                m, s0, s1, niter, workloads.FIELDS[f], e, " ".join("%.1e" % (r / fm) for r, fm in st[:, f])),
                flush=True)
            assert e <= TOL
            assert not got[f][0].imag.any()
    finally:
        oracle_mod.use_blas(False)
    # refinement reduces the residual (for the Poisson field's k=0 plane the
    # residual is measured after removing the part of f in the null space of L)
    assert np.all(st[-1, :, 0] < 1e-4 * st[0, :, 0])


def test_residual_kernel_matches_numpy(dev):
    """niter=0: stats[0] is max|f - L u0| of the returned u0, computed by the FP64 MFMA kernel."""
    import torch
    from duct_amd import workloads
    m, M = 65, 5
    s = ensure(dev, m)
    cfg = workloads.SOLVE_257
    fields = [workloads.rhs_field(cfg, f, m=m, nmodes=M) for f in (0, 3)]
    alams = [1000.0, 0.0]
    got, st = run_mixed(dev, fields, "dn", alams, scheme0="bf16x3", scheme1="bf16x3", niter=0)
    for i, op in enumerate((s.dirichlet.op, s.neumann.op)):
        u = got[i][:, 1:-1, 1:-1]
        f = fields[i][:, 1:-1, 1:-1]
        k = np.arange(M)
        xl = alams[i] + k.astype(float) ** 2
        r = f - (u @ op.T + np.einsum("ij,kjl->kil", op, u) - xl[:, None, None] * u)
        r[0].imag = 0.0
        if i == 1:   # the singular k=0 Poisson plane is reported after removing its null-space part
            b = s.neumann
            q, p = b.Pinv[b.iofzero - 1, :], b.P[:, b.iofzero - 1]
            r0 = r[0].real
            r[0] = r0 - (q @ r0 @ q) * np.outer(p, p)
        def cmax(a):   # the kernel reports the largest real or imaginary component
            return max(np.max(np.abs(a.real)), np.max(np.abs(a.imag)))
        want = cmax(r)
        scale = np.abs(op).sum(1).max() * np.max(np.abs(u)) * 1e-13
        assert abs(st[0, i, 0] - want) <= scale + 1e-9 * want, (i, st[0, i, 0], want)
        f0 = f.copy()
        f0[0] = f0[0].real
        assert st[0, i, 1] == cmax(f0)


def test_mixed_rejects_bad_scheme(dev):
    import torch
    from duct_amd._lib import DuctError, call, byref_int, byref_double
    import ctypes
    ensure(dev, 33)
    x = torch.zeros((2, 33, 33), dtype=torch.complex128, device="cuda")
    P = (ctypes.c_size_t * 1)(x.data_ptr())
    A = (ctypes.c_double * 1)(1.0)
    with pytest.raises(DuctError):
        call("duct_helm3d_mixed_batch_dev_", byref_int(1), P, P, b"d", A, byref_int(33), byref_int(2),
             byref_int(0), byref_double(1.0), byref_int(9), byref_int(2), byref_int(1), None, ctypes.c_size_t(1))


def _check_all_modes(label, got, refs):
    """Every mode of every field against the oracle: max-norm error over the field's maximum."""
    from duct_amd import workloads
    for f in range(4):
        e = relerr(got[f], refs[f])
        d = np.max(np.abs(got[f] - refs[f]).reshape(got[f].shape[0], -1), axis=1)
        print("%s 257^2x256 field %s: %.2e of the field over all modes (worst mode %d)" % (
            label, workloads.FIELDS[f], e, int(np.argmax(d))), flush=True)
        assert e <= TOL


def test_mixed_full_size_config2(dev, config2_refs):
    """257^2 x 256 modes x 4 fields, bf16x6 solve + one bf16x6 refinement (the bench's unfolded
    refined path): every mode <= 1e-10 of its field."""
    from duct_amd import workloads
    ensure(dev, workloads.SOLVE_257.m)
    fields, alams, refs = config2_refs
    got, st = run_mixed(dev, fields, workloads.BOUNDS, alams, scheme0="bf16x6", scheme1="bf16x6", niter=1)
    _check_all_modes("bf16x6 x1", got, refs)


@pytest.mark.parametrize("scheme,niter", [("fold16", 1), ("fold16q", 1), ("f32", 1), ("f32", 2)])
def test_fold16_full_size_config2(dev, config2_refs, scheme, niter):
    """257^2 x 256 modes x 4 fields: the folded fused solve with one refinement (scheme 5: FP64-MFMA
    residual; scheme 6: the residual from exact int8 digit products), and the FP32-MFMA eigenbasis
    products (scheme 1) with one and two FP64 refinements: every mode <= 1e-10 of its field (the
    folded schemes solve the Poisson modes with xlambda < 64 with the FP64 passes)."""
    from duct_amd import workloads
    ensure(dev, workloads.SOLVE_257.m)
    fields, alams, refs = config2_refs
    got, st = run_mixed(dev, fields, workloads.BOUNDS, alams, scheme0=scheme, scheme1=scheme, niter=niter)
    print("%s x%d residual history:" % (scheme, niter), [" ".join("%.1e" % (r / fm) for r, fm in st[:, f]) for f in range(4)],
          flush=True)
    _check_all_modes("%s x%d" % (scheme, niter), got, refs)
    assert not got[0][0].imag.any()
    # the correction reduces the residual by >= 1e4 on every field
    assert np.all(st[-1, :, 0] < 1e-4 * st[0, :, 0])


@pytest.mark.parametrize("scheme", ["fold16", "fold16q"])
def test_fold16_singular_plane_past_mode_0(dev, oracle_mod, scheme):
    """Neumann with alambda = -100 at modes 8 ... 15: the singular plane (xlambda = 0 at k = 10) and
    the other modes with xlambda < 64 take the FP64 passes, modes 13-15 are refined; <= 1e-10."""
    import torch
    m, M, ib, alam = 129, 8, 8, -100.0
    ensure(dev, m)
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    rng = np.random.default_rng(103)
    f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    ref, _ = oracle_mod.helm3d(ob, f, "n", alam, ib=ib)
    r = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    u = torch.zeros_like(r)
    dev.helm3d_mixed_batch([r], [u], "n", [alam], ib=ib, scheme0=scheme, scheme1=scheme, niter=1)
    torch.cuda.synchronize()
    got = u.cpu().numpy()
    e = relerr(got, ref)
    print("%s, singular plane at k=10: rel.err %.2e (refined modes %.2e)" % (scheme, e, relerr(got[5:], ref[5:])),
          flush=True)
    assert e <= TOL


def test_fold16q_equals_fold16(dev):
    """Scheme 6 (int8-digit residual) against scheme 5 (FP64-MFMA residual) on every mode of a
    129^2 x 12-mode batch: the two refined solutions agree to 2e-11 of the field maximum (each is
    within ~1e-11 of the converged solution of the folded operator); the first solve of scheme 6 is
    rounded to 2^-24 of its quarter's maximum and its residual is 1e2-1e3 times scheme 5's, which the
    correction removes."""
    from duct_amd import workloads
    m, M = 129, 12
    ensure(dev, m)
    cfg = workloads.SOLVE_257
    fields = [workloads.rhs_field(cfg, f, m=m, nmodes=M) for f in range(4)]
    alams = [workloads.alambda_of(cfg, f) for f in range(4)]
    u5, _ = run_mixed(dev, fields, workloads.BOUNDS, alams, scheme0="fold16", scheme1="fold16", niter=1)
    u6, _ = run_mixed(dev, fields, workloads.BOUNDS, alams, scheme0="fold16q", scheme1="fold16q", niter=1)
    for f in range(4):
        e = relerr(u6[f], u5[f])
        print("field %s: fold16q vs fold16 rel. diff %.2e" % (workloads.FIELDS[f], e), flush=True)
        assert e <= 2e-11


def test_fold16_rejects_unsupported_sizes(dev):
    import torch
    from duct_amd._lib import DuctError
    ensure(dev, 33)
    x = torch.zeros((2, 33, 33), dtype=torch.complex128, device="cuda")
    y = torch.zeros_like(x)
    with pytest.raises(DuctError):
        dev.helm3d_mixed_batch([x], [y], "d", [1000.0], scheme0="fold16", scheme1="fold16", niter=1)
    ensure(dev, 129)
    x = torch.zeros((2, 129, 129), dtype=torch.complex128, device="cuda")
    y = torch.zeros_like(x)
    with pytest.raises(DuctError):
        dev.helm3d_mixed_batch([x], [y], "d", [1000.0], scheme0="fold16", scheme1="bf16x6", niter=1)
