import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libduct_hip.so on cuda:0)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()  # builds liboracle_duct.so with gcc if it is missing
    return oracle


@pytest.fixture(scope="session")
def config2_refs(oracle_mod):
    """The oracle's solution of the whole 257^2 x 256-mode x 4-field workload (the seeded right-hand
    sides of the benchmarked solve stage), computed once per session with BLAS on up to 16 host
    cores: (fields, alambdas, solutions)."""
    from duct_amd import workloads
    cfg = workloads.SOLVE_257
    fields = [workloads.rhs_field(cfg, f) for f in range(cfg.nfields)]
    alams = [workloads.alambda_of(cfg, f) for f in range(cfg.nfields)]
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(cfg.m))
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    oracle_mod.use_blas(True)
    try:
        refs = [oracle_mod.helm3d(ob, fields[f], workloads.BOUNDS[f], alams[f], alp=cfg.alp,
                                  nthreads=max(1, min(16, cores)))[0] for f in range(cfg.nfields)]
    finally:
        oracle_mod.use_blas(False)
    return fields, alams, refs
