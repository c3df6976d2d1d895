"""GPU tests of the even/odd folded Dirichlet solve (duct_set_fold_).

Bar: relative error <= 1e-12 against the FP64 oracle, which runs the reference's
unfolded products.  The folded basis is the parity projection of the
reference's PD (a nearly degenerate even/odd pair returned rotated included); on random data the two differ by ~1e-13 at m <= 257 (measured
on the CPU with the oracle and a parity-projected basis), so 1e-12 is the
plain FP64 bar, not a loosened one.  Sizes cover NH = 64, 128, 256 and odd and
even n.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def dev():
    import torch
    import duct_amd
    torch.cuda.set_device(0)
    duct_amd.init(0)
    yield duct_amd
    duct_amd.set_fold(False)
    duct_amd.shutdown()


_cur = {"m": None}


def ensure(dev, oracle_mod, m):
    if _cur["m"] != m:
        dev.setup(m)
        _cur["m"] = m
    return oracle_mod.OracleBasis(oracle_mod.pseudomat(m))


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def solve(dev, f, bound, alam, fold, ib=0):
    import torch
    dev.set_fold(fold)
    rhs = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    sol = torch.zeros_like(rhs)
    dev.helm3d(rhs, sol, bound, alam, ib=ib)
    torch.cuda.synchronize()
    dev.set_fold(False)
    return sol.cpu().numpy()


@pytest.mark.parametrize("m,M,ib", [(67, 5, 0), (129, 4, 0), (130, 3, 1), (200, 3, 0), (257, 3, 0), (257, 2, 7),
                                    (300, 2, 0), (514, 2, 0)])
def test_fold_matches_oracle(dev, oracle_mod, m, M, ib):
    ob = ensure(dev, oracle_mod, m)
    assert dev.fold_size("d") == {67: 64, 129: 64, 130: 64, 200: 128, 257: 128, 300: 256, 514: 256}[m]
    rng = np.random.default_rng(m + ib)
    f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    ref, _ = oracle_mod.helm3d(ob, f, "d", 1000.0, ib=ib)
    got = solve(dev, f, "d", 1000.0, True, ib=ib)
    plain = solve(dev, f, "d", 1000.0, False, ib=ib)
    e, e0 = relerr(got, ref), relerr(plain, ref)
    print("m=%d folded %.2e  unfolded %.2e" % (m, e, e0), flush=True)
    assert e <= TOL and e0 <= TOL
    if ib == 0:
        assert not got[0].imag.any()
    assert not got[:, 0, :].any() and not got[:, :, -1].any()     # Dirichlet boundary stays zero


def test_fold_small_m_and_neumann_unchanged(dev, oracle_mod):
    """m = 33 does not fold (NH would be 16); Neumann groups never fold."""
    ob = ensure(dev, oracle_mod, 33)
    assert dev.fold_size("d") == 0
    rng = np.random.default_rng(5)
    f = rng.standard_normal((3, 33, 33)) + 1j * rng.standard_normal((3, 33, 33))
    a = solve(dev, f, "d", 1000.0, True)
    b = solve(dev, f, "d", 1000.0, False)
    assert np.array_equal(a, b)
    ob = ensure(dev, oracle_mod, 129)
    f = rng.standard_normal((3, 129, 129)) + 1j * rng.standard_normal((3, 129, 129))
    a = solve(dev, f, "n", 0.0, True)
    b = solve(dev, f, "n", 0.0, False)
    assert np.array_equal(a, b)


def test_fold_config2_batch(dev, oracle_mod):
    """257^2 x 256 modes, u, v, w, p batch with folding on: Dirichlet fields
    folded, p unfolded; sampled modes against the oracle."""
    import torch
    from duct_amd import workloads
    cfg = workloads.SOLVE_257
    ob = ensure(dev, oracle_mod, cfg.m)
    fields = [workloads.rhs_field(cfg, f) for f in range(4)]
    rhs = [torch.from_numpy(x).cuda() for x in fields]
    sol = [torch.zeros_like(r) for r in rhs]
    alams = [workloads.alambda_of(cfg, f) for f in range(4)]
    dev.set_fold(True)
    try:
        dev.helm3d_batch(rhs, sol, workloads.BOUNDS, alams)
        torch.cuda.synchronize()
    finally:
        dev.set_fold(False)
    oracle_mod.use_blas(True)
    try:
        for f in range(4):
            for k0 in (0, 1, 200, 255):
                ref, _ = oracle_mod.helm3d(ob, fields[f][k0:k0 + 1], workloads.BOUNDS[f], alams[f], ib=k0)
                e = relerr(sol[f][k0:k0 + 1].cpu().numpy(), ref)
                print("257^2x256 folded field %s mode %d rel.err %.2e" % (workloads.FIELDS[f], k0, e), flush=True)
                assert e <= TOL
    finally:
        oracle_mod.use_blas(False)


@pytest.mark.parametrize("m,M,ib,alam", [(67, 4, 6, 0.0), (129, 10, 0, 0.0), (129, 3, 7, 2.5), (200, 3, 6, 0.0),
                                        (257, 10, 0, 0.0), (257, 3, 30, 0.0), (300, 2, 7, 0.0), (514, 2, 7, 0.0)])
def test_fold_neumann(dev, oracle_mod, m, M, ib, alam):
    """Folded Neumann solve (mode 2): the leading modes with xlambda < 64 (the singular k=0
    plane and the ill-conditioned low Poisson modes) take the unfolded FP64 passes, the rest the
    folded ones; the batches here straddle that switch.  Field-level bar 1e-12 up to m = 257
    (the folded modes from xlambda = 64 up measured <= 2e-13 of the field at m = 257); the
    condition number of the planes grows as m^4, and so does the bound past m = 257."""
    tol = TOL * max(1.0, (m / 257.0) ** 4)
    ob = ensure(dev, oracle_mod, m)
    assert dev.fold_size("n") > 0
    rng = np.random.default_rng(7 * m + ib)
    f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    ref, _ = oracle_mod.helm3d(ob, f, "n", alam, ib=ib)
    got = solve(dev, f, "n", alam, 2, ib=ib)
    if ib == 0 and alam == 0.0:
        io = dev.solver._state["setup"].neumann.iofzero   # gauge node (iofzero+1, 1-based grid)
        assert abs(got[0, io, io]) <= 1e-12 * np.abs(got[0]).max()
        assert not got[0].imag.any()
    nlow = sum(1 for k in range(ib, ib + M) if alam + k * k < 64.0)
    e = relerr(got, ref)
    ef = relerr(got[nlow:], ref[nlow:]) if nlow < M else 0.0
    print("m=%d ib=%d alambda=%g Neumann, %d unfolded + %d folded modes: rel.err %.2e (folded modes %.2e)" % (
        m, ib, alam, nlow, M - nlow, e, ef), flush=True)
    assert e <= tol


def test_fold_neumann_singular_plane_past_mode_0(dev, oracle_mod):
    """alambda = -100, modes 8 ... 15: xlambda = k^2 - 100 is 0 at k = 10 (a singular plane with
    its gauge, solvediag.f90:151, past mode 0) and below 64 up to k = 12, so modes 8-12 take the
    unfolded passes and 13-15 the folded ones; both parts against the oracle at the FP64 bar."""
    m, M, ib, alam = 129, 8, 8, -100.0
    ob = ensure(dev, oracle_mod, m)
    rng = np.random.default_rng(101)
    f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    ref, _ = oracle_mod.helm3d(ob, f, "n", alam, ib=ib)
    got = solve(dev, f, "n", alam, 2, ib=ib)
    io = dev.solver._state["setup"].neumann.iofzero
    k0 = 10 - ib
    assert abs(got[k0, io, io]) <= 1e-12 * np.abs(got[k0]).max()   # the gauge node holds `wanted` (0)
    e, e_low, e_fold = relerr(got, ref), relerr(got[:5], ref[:5]), relerr(got[5:], ref[5:])
    print("singular plane at k=10: rel.err %.2e (unfolded modes %.2e, folded modes %.2e)" % (e, e_low, e_fold),
          flush=True)
    assert e <= TOL


def test_fold_all_config2_batch(dev, oracle_mod, config2_refs):
    """257^2 x 256 x 4 with fold mode 2 (every field folded but the Poisson modes with xlambda < 64,
    which take the unfolded passes): every mode of every field against the oracle at the FP64 bar
    (max-norm error over the field's maximum); the folded Poisson modes from k = 8 up are also
    reported against their own maxima."""
    import torch
    from duct_amd import workloads
    cfg = workloads.SOLVE_257
    ensure(dev, oracle_mod, cfg.m)
    fields, alams, refs = config2_refs
    rhs = [torch.from_numpy(x).cuda() for x in fields]
    sol = [torch.zeros_like(r) for r in rhs]
    dev.set_fold(2)
    try:
        dev.helm3d_batch(rhs, sol, workloads.BOUNDS, alams)
        torch.cuda.synchronize()
    finally:
        dev.set_fold(0)
    for f in range(4):
        got = sol[f].cpu().numpy()
        e = relerr(got, refs[f])
        d = np.max(np.abs(got - refs[f]).reshape(got.shape[0], -1), axis=1)
        own = d / np.maximum(np.max(np.abs(refs[f]).reshape(got.shape[0], -1), axis=1), 1e-300)
        print("257^2x256 fold-all field %s: %.2e of the field over all modes; worst mode %d; worst mode "
              "against its own maximum %.2e (k = %d)" % (workloads.FIELDS[f], e, int(np.argmax(d)),
                                                          float(own.max()), int(np.argmax(own))), flush=True)
        assert e <= TOL


def _rotated_pair_basis(P, Pinv, lam, theta):
    """P R and R^T P^-1 with R the rotation by theta of the even/odd eigenvector pair whose
    eigenvalues are closest (relative): what an eigensolver may return for a nearly degenerate
    pair.  Returns the new (P, P^-1), the pair and its relative gap."""
    J = P[::-1]
    se, so = ((P + J) ** 2).sum(0), ((P - J) ** 2).sum(0)
    fe = se / (se + so)
    ev, od = np.where(fe > 0.5)[0], np.where(fe <= 0.5)[0]
    L1, L2 = lam[ev][:, None], lam[od][None, :]
    gap = np.abs(L1 - L2) / np.maximum(np.abs(L1), np.abs(L2))
    i, j = np.unravel_index(np.argmin(gap), gap.shape)
    a, b = ev[i], od[j]
    c, s = np.cos(theta), np.sin(theta)
    P2, Pi2 = P.copy(), Pinv.copy()
    P2[:, a], P2[:, b] = c * P[:, a] + s * P[:, b], -s * P[:, a] + c * P[:, b]
    Pi2[a], Pi2[b] = c * Pinv[a] + s * Pinv[b], -s * Pinv[a] + c * Pinv[b]
    return P2, Pi2, (a, b), float(gap[i, j])


@pytest.mark.parametrize("bound,theta", [("d", np.pi / 4), ("d", 0.5), ("n", np.pi / 4)])
def test_fold_with_rotated_near_degenerate_pair(dev, oracle_mod, bound, theta):
    """A basis whose nearly degenerate even/odd eigenvector pair comes back rotated (mixed
    parities, as another eigensolver may return it; the pair's relative gap is ~2e-5 at
    m = 257) still folds: each mixed column is projected onto its parity, the folded inverse is
    recomputed and the pair's eigenvalues come from the folded operator.  The folded solve then
    agrees with the oracle as closely as with the unrotated basis (FP64 folded passes; for 'd'
    also the int8-residual mixed solve)."""
    m, M = 257, 3
    ob = ensure(dev, oracle_mod, m)
    s = dev.solver._state["setup"]
    b = s.dirichlet if bound == "d" else s.neumann
    rng = np.random.default_rng(31 + (bound == "n"))
    f = rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))
    if bound == "d":
        f[:, 0, :] = f[:, -1, :] = f[:, :, 0] = f[:, :, -1] = 0
    alam = 1000.0 if bound == "d" else 100.0   # xlambda >= 64: every Neumann mode folds
    mode = 1 if bound == "d" else 2
    ref, _ = oracle_mod.helm3d(ob, f, bound, alam)
    base = relerr(solve(dev, f, bound, alam, mode), ref)
    P2, Pi2, pair, gap = _rotated_pair_basis(b.P, b.Pinv, b.lam, theta)
    print("rotated pair %s (relative gap %.1e), theta %.3f" % (pair, gap, theta), flush=True)
    dev.solver.upload_basis(bound, m, P2, Pi2, b.lam, s.DC, b.AM1, b.iofzero, b.wanted)
    try:
        assert dev.fold_size(bound) > 0
        got = solve(dev, f, bound, alam, mode)
        e = relerr(got, ref)
        print("%s folded: unrotated basis %.2e, rotated %.2e" % (bound, base, e), flush=True)
        assert e <= max(2.0 * base, 1e-13)
        if bound == "d":
            import torch
            r = torch.from_numpy(np.ascontiguousarray(f)).cuda()
            u = torch.zeros_like(r)
            dev.helm3d_mixed_batch([r], [u], "d", [alam], scheme0="fold16q", scheme1="fold16q", niter=1)
            torch.cuda.synchronize()
            em = relerr(u.cpu().numpy(), ref)
            print("d fold16q with the rotated basis: %.2e" % em, flush=True)
            assert em <= 1e-10
    finally:
        dev.solver.upload_basis(bound, m, b.P, b.Pinv, b.lam, s.DC, b.AM1, b.iofzero, b.wanted)


@pytest.mark.parametrize("m,M,ib", [(131, 3, 0), (200, 3, 0), (257, 3, 0), (257, 2, 5), (258, 2, 0)])
def test_fused_quarters_bit_identical(dev, oracle_mod, m, M, ib):
    """Folded passes 2 and 3 on LDS-resident quarters (duct_set_fused_quarters_, the default at
    folded size 128) against the two planar launches they replace: the same MFMA sequence and
    divide, so the u, v, w, p batch must agree to the bit (p's modes with xlambda < 64 take the
    unfolded passes either way); padded quarters (m = 131: 65 of 128 rows), ragged
    even n (m = 200, 258) and a mode offset are covered.  One mode is also checked against the
    oracle."""
    import torch
    from duct_amd import workloads
    ensure(dev, oracle_mod, m)
    assert dev.fold_size("d") == 128 and dev.fold_size("n") == 128
    rng = np.random.default_rng(11 * m + ib)
    fields = [rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m)) for _ in range(4)]
    alams = [1000.0, 1000.0, 1000.0, 0.0]

    def run(fused):
        rhs = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in fields]
        sol = [torch.zeros_like(r) for r in rhs]
        dev.set_fold(2)
        dev.set_fused_quarters(fused)
        try:
            dev.helm3d_batch(rhs, sol, workloads.BOUNDS, alams, ib=ib)
            torch.cuda.synchronize()
        finally:
            dev.set_fold(0)
            dev.set_fused_quarters(True)
        return [s.cpu().numpy() for s in sol]

    a, b = run(True), run(False)
    for f in range(4):
        assert np.array_equal(a[f], b[f]), (f, float(np.max(np.abs(a[f] - b[f]))))
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    ref, _ = oracle_mod.helm3d(ob, fields[0][1:2], "d", 1000.0, ib=ib + 1)
    e = relerr(a[0][1:2], ref)
    print("m=%d ib=%d fused quarters: bit-identical, Dirichlet mode %d rel.err %.2e" % (m, ib, ib + 1, e), flush=True)
    assert e <= TOL
