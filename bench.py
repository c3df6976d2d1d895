"""Benchmark of the cross-plane Helmholtz/Poisson solve stage on one MI355X.

Workload (N=1): 257 x 257 CGL points x 256 complex Fourier modes x 4 fields
(u, v, w: Dirichlet Helmholtz, alambda = 1000; p: Neumann Poisson, alambda = 0,
k=0 gauge), seeded synthetic right-hand sides, inputs resident in HBM.
One step = one batched solve of all four fields (4 projection passes per
boundary group + gauge + boundary kernels).

Metric: mode-solves/s, a mode-solve being one (field, k_x) complex 2-D solve.
Roofline: the projection kernel (k_proj, FP64 MFMA) against the FP64 MFMA peak,
with the algorithmic 2 n^3 FLOP per real plane and pass (n = m-2, padding not
counted, the imaginary plane at k=0 not counted), timed with hipEvents the
library records around every launch on its stream.
cpu_baseline: the oracle (C restatement of solvediag/int2d3d/helm3d with a
BLAS dgemm per product) on the host cores over the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_MFMA_PEAK = 78.6e12   # MI355X dense FP64 matrix, FLOP/s
HBM_PEAK = 8.0e12          # bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(cfg, fields, alams, sols_gpu, threads):
    """Oracle helm3d on the host cores over the whole workload; also the checker
    of the GPU result (max relative error over all fields)."""
    import oracle
    blas = oracle.use_blas(True)
    ob = oracle.OracleBasis(oracle.pseudomat(cfg.m))
    from duct_amd import workloads
    t0 = time.perf_counter()
    worst = 0.0
    refs = []
    for f in range(cfg.nfields):
        ref, _ = oracle.helm3d(ob, fields[f], workloads.BOUNDS[f], alams[f], alp=cfg.alp, ib=0, nthreads=threads)
        refs.append(ref)
    t_all = time.perf_counter() - t0
    for f in range(cfg.nfields):
        ref = refs[f]
        worst = max(worst, float(np.max(np.abs(sols_gpu[f] - ref)) / np.max(np.abs(ref))))
    # one core: a few modes of u, scaled
    k1 = 4
    t0 = time.perf_counter()
    oracle.helm3d(ob, fields[0][:k1], "d", alams[0], alp=cfg.alp, ib=1, nthreads=1)
    t_one = time.perf_counter() - t0
    oracle.use_blas(False)
    nsolves = cfg.nfields * cfg.nmodes
    return {
        "value": nsolves / t_all,
        "unit": "mode-solves/s",
        "cores": threads,
        "kind": "port",
        "sample": "full workload (%d fields x %d modes, %d^2) with the oracle (C restatement of "
                  "helm3d/int2d3d/solvediag, %s), OpenMP over modes; %.2f s" % (
                      cfg.nfields, cfg.nmodes, cfg.m, blas, t_all),
        "value_1core": k1 / t_one,
        "gpu_max_rel_err_vs_oracle": worst,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the solve shards over Fourier modes with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    import duct_amd
    from duct_amd import workloads

    cfg = workloads.CONFIGS[args.config]
    m, M, F = cfg.m, cfg.nmodes, cfg.nfields
    n = m - 2
    torch.cuda.set_device(0)
    duct_amd.init(0)
    duct_amd.setup(m)
    fields = [workloads.rhs_field(cfg, f) for f in range(F)]
    alams = [workloads.alambda_of(cfg, f) for f in range(F)]
    bounds = workloads.BOUNDS[:F]
    rhs = [torch.from_numpy(x).cuda() for x in fields]
    sol = [torch.zeros_like(r) for r in rhs]

    def step():
        duct_amd.helm3d_batch(rhs, sol, bounds, alams, alp=cfg.alp, ib=0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    with duct_amd.KernelTimer() as kt:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    dt = (t1 - t0) / args.steps
    nsolves = F * M
    value = nsolves / dt

    # roofline of the projection kernel (all four passes share k_proj)
    planes = F * (2 * M - 1)                    # Im at k=0 not solved
    flop_step = 8.0 * n ** 3 * planes           # 4 products x 2 n^3 per real plane
    proj = ["proj_c2p", "proj_p2p_div", "proj_p2p", "proj_p2c"]
    proj_ms = sum(kt.ms[k] for k in proj)
    proj_launches = sum(kt.count[k] for k in proj)
    avg_launch_s = proj_ms / 1e3 / proj_launches
    flop_per_launch = flop_step * args.steps / proj_launches
    achieved = flop_per_launch / avg_launch_s
    kernels = {k: {"launches": kt.count[k], "avg_us": (kt.ms[k] / kt.count[k] * 1e3) if kt.count[k] else 0.0}
               for k in kt.ms if kt.count[k]}
    step_kernel_ms = sum(kt.ms.values()) / args.steps

    sols_host = [s.cpu().numpy() for s in sol]
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(cfg, fields, alams, sols_host, threads)

    out = {
        "metric": "Helmholtz mode-solves/s at 257^2 x256 modes: mixed vs FP64 @1e-10 err; %MFMA peak",
        "value": value,
        "unit": "mode-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Chebyshev-series right-hand sides, inputs resident in HBM)",
        "config": {"workload": "cross-plane solve stage, plain FP64 path: %d^2 CGL x %d modes x %d fields "
                               "(u,v,w Dirichlet alambda=1000; p Neumann alambda=0)" % (m, M, F),
                   "m": m, "nmodes": M, "nfields": F, "precision_path": "fp64"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": FP64_MFMA_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK, "traffic": None,
                     "kernel": "k_proj (FP64 MFMA projection pass, 4 launches per boundary group per step)",
                     "flop_per_launch": flop_per_launch, "avg_launch_us": avg_launch_s * 1e6},
        "kernels": kernels,
        "kernel_ms_per_step": step_kernel_ms,
        "stage_tflops": flop_step / dt / 1e12,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    duct_amd.shutdown()


if __name__ == "__main__":
    main()
