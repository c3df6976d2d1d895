"""Benchmark of the cross-plane Helmholtz/Poisson solve stage on one MI355X.

Workload (N=1): 257 x 257 CGL points x 256 complex Fourier modes x 4 fields
(u, v, w: Dirichlet Helmholtz, alambda = 1000; p: Neumann Poisson, alambda = 0,
k=0 gauge), seeded synthetic right-hand sides, inputs resident in HBM.
One step = one batched solve of all four fields.

Paths timed on the same inputs:
  fp64         the plain FP64 form: four FP64-MFMA eigenbasis passes per boundary
               group, the reference's unfolded products (agrees with the oracle to ~1e-15)
  fp64_fold_d  the Dirichlet fields in the even/odd folded eigenbasis (half the FLOPs;
               ~1e-13 of the oracle), the Neumann field unfolded
  fp64_fold    every field folded except the Poisson modes with xlambda < 64 (the
               singular k=0 plane among them), which take the unfolded passes
               (~2e-13 of the oracle)
  mixed        the refined form of --mixed (default bf16x6,bf16x6,1): low-precision
               passes (FP32 or split-bf16 MFMA), FP64 residual, correction(s)
               accumulated into the FP64 solution
  mixed_f32    the same with FP32-MFMA passes (variant A), --f32-iter refinements
  mixed_fold   the refined form in the folded basis (scheme fold16): fused per-quarter
               fp16 hi/lo solves, FP64 residual of the folded operator, one correction;
               the Poisson modes with xlambda < 64 by the unfolded FP64 passes
  mixed_fold_q the same with the residual's products formed exactly from int8 digits on the
               int8 matrix cores (scheme fold16q), summed to FP64 accuracy
The reported `value` is the fastest path whose result is within 1e-10 (relative,
max norm, every field) of the oracle, the metric's tolerance; value_path,
value_tier and value_max_rel_err_vs_oracle say which path and how close.  The
plain FP64 paths are also held to the FP64 bar 1e-12 (fp64_tier: the fastest of
them that meets it); mixed_tier is the fastest refined path within 1e-10 and
mixed_vs_fp64 its ratio to the plain FP64 solve.  Every path's throughput, error
and kernel times are in the line.  Without the oracle (--no-cpu-baseline) nothing
is verified and value is the plain FP64 path.

Metric: mode-solves/s, a mode-solve being one (field, k_x) complex 2-D solve.
Roofline: the chosen path's dominant kernel class, algorithmic FLOP per launch
(2 n^3 per real plane and product for an unfolded projection pass,
2 (h_e^2 + h_o^2) n for a folded one with h_e = ceil(n/2), h_o = floor(n/2),
4 n^3 for the dense FP64 residual; n = m-2, padding and the unsolved imaginary
plane at k=0 not counted) over its mean launch time from hipEvents recorded on
the library's stream during the timed steps.
cpu_baseline: the oracle (C restatement of solvediag/int2d3d/helm3d, one BLAS
dgemm per product) over the whole workload on the host cores this process may
use (its CPU affinity, capped by OMP_NUM_THREADS when that is set: the box's
share for one GPU); cores, host_cores and affinity_cores are all reported.  It
also serves as the checker of every GPU result.
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

PEAK = {"f64": 78.6e12, "f32": 157.3e12, "bf16": 2.5e15, "f16": 2.5e15,   # MI355X dense matrix peaks,
        "i8": 5.0e15}                                                   # FLOP/s (int8: OP/s)
# HBM bytes per launch of a kernel class at a workload (m, nmodes, nfields), from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs scaled by the same counters over a plain 1 GiB copy
# (profiles/k_projf_k_projq_f64_fold_257x256x4_pmc.md, profiles/k_lpq_fold16q_257x256x4_traffic.md:
# fetch counted 0.5 per byte, write 1.0).
TRAFFIC = {("fold", 257, 256, 4): 1124.0e6, ("fold_mid", 257, 256, 4): 1110.2e6, ("lp_p2p", 257, 256, 4): 609.6e6}
TOL = 1e-10        # the metric's tolerance (refined and folded paths)
TOL_FP64 = 1e-12   # the FP64 bar (plain FP64 paths)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2,
                    help="workload (duct_amd.workloads.CONFIGS): 0 = 33^2 x 16 x 1 field, 1 = 129^2 x 64 x 4, "
                         "2 = 257^2 x 256 x 4 (default)")
    ap.add_argument("--mixed", default="bf16x6,bf16x6,1",
                    help="scheme0,scheme1,niter of the refined solve")
    ap.add_argument("--path", choices=["auto", "fp64", "fp64_fold_d", "fp64_fold", "mixed", "mixed_f32",
                                       "mixed_fold", "mixed_fold_q"],
                    default="auto")
    ap.add_argument("--f32-iter", type=int, default=1,
                    help="refinements of the FP32-MFMA refined solve (path mixed_f32; one leaves 5.1e-11 of "
                         "the oracle at 257^2 x 256 x 4, two 2.7e-11)")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the oracle: the line is then unverified and value is the plain FP64 path")
    ap.add_argument("--events-only", action="store_true",
                    help="time only the pass with per-launch events (profiled runs: the profiler's\n                         durations then cover the launches the kernel classes come from)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="solve the Neumann group after the Dirichlet group on one stream (profiled runs:\n                         every launch then runs alone, as in the pass the kernel classes come from)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a float64 buffer of MIB MiB (steps times): the known-size copy the HBM "
                         "traffic counters are scaled by")
    ap.add_argument("--c3", action="store_true",
                    help="only time the pseudo-spectral step at 129^2 x 64 modes, mgalx 192: advection3d, "
                         "the backward-Euler projection substep, one full RK3 time step (three substeps "
                         "of duct_rk3_substep_dev_), and one field of the mixed-boundary and Shen-Legendre "
                         "solves at the same size")
    return ap.parse_args()


def copy_probe(mib, steps):
    """A plain device copy of known size; under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per
    dispatch give the counters' scale (bytes counted per byte moved)."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 8, "copies": steps}), flush=True)


def _timed_twice(fn, steps, warmup):
    """Warmup, then `steps` calls with nothing recorded between the launches (the time), then the
    same `steps` again with the library's hipEvents around every launch (the kernel classes)."""
    import torch
    import duct_amd
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    with duct_amd.KernelTimer() as kt:
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt_ev = (time.perf_counter() - t0) / steps
    return dt, dt_ev, kt


def c3_bench(steps, warmup):
    """The pseudo-spectral step at 129^2 CGL, 64 modes, mgalx = 192: time of advection3d and of one projection
    substep on device; the kernel classes come from the library's own hipEvents."""
    import torch
    import duct_amd
    from duct_amd import step, workloads
    cfg = workloads.STEP_129
    torch.cuda.set_device(0)
    duct_amd.init(0)
    duct_amd.setup(cfg.m)
    u, v, w = (torch.from_numpy(x).cuda() for x in workloads.duct_velocity(cfg))
    adv = [torch.empty_like(u) for _ in range(3)]
    # the RK3 state evolves in place: a copy of the velocity, zero pressure and advection
    ru, rv, rw = u.clone(), v.clone(), w.clone()
    rp = torch.zeros_like(u)
    radv = [torch.zeros_like(u) for _ in range(3)]
    out = {}
    for name, fn in (("advection3d", lambda: step.advection(u, v, w, *adv, mgalx=192, alp=cfg.alp)),
                     ("step", lambda: step.c3_step(u, v, w, nu=0.01, dt=0.1, mgalx=192, alp=cfg.alp)),
                     ("rk3_step", lambda: step.rk3_step(ru, rv, rw, rp, radv, fnu=0.01, dt=0.01, mgalx=192,
                                                        alp=cfg.alp))):
        dt, dt_ev, kt = _timed_twice(fn, steps, warmup)
        out[name] = {"ms": dt * 1e3, "ms_with_events": dt_ev * 1e3,
                     "kernels": {k: {"launches": kt.count[k] / steps, "avg_us": kt.ms[k] / kt.count[k] * 1e3}
                                 for k in kt.ms if kt.count[k]}}
    # the widened solves at the same 129^2 x 64-mode size: mixed-boundary (temperature) and Shen-Legendre
    from duct_amd import shen, temperature, workloads as W
    rhs = torch.from_numpy(W.rhs_field(workloads.SOLVE_129, 0)).cuda()
    sol = torch.empty_like(rhs)
    mm, MM = rhs.shape[-1], rhs.shape[0]
    bc = torch.zeros((MM, 2, mm), dtype=torch.float64, device="cuda")
    temperature.setup_temp(mm, "dndn", 1.0, 1.0)
    shen.setup_shen(mm, "d")
    for name, fn in (("helm3d_temp dndn", lambda: temperature.helm3d_temp(rhs, sol, bc, bc, 1000.0)),
                     ("helm3d_shen d", lambda: shen.helm3d_shen(rhs, sol, "d", 1000.0))):
        dt, dt_ev, kt = _timed_twice(fn, steps, warmup)
        out[name] = {"ms": dt * 1e3, "ms_with_events": dt_ev * 1e3, "mode_solves_per_s": MM / dt,
                     "kernels": {k: {"launches": kt.count[k] / steps, "avg_us": kt.ms[k] / kt.count[k] * 1e3}
                                 for k in kt.ms if kt.count[k]}}
    lines, X, M = cfg.m * cfg.m, 192, cfg.nmodes
    xp = out["advection3d"]["kernels"].get("xprod")
    if xp:
        # algorithmic bytes of the product kernel: 3 fields of M complex modes in and 6 products out
        # per (y, z) line, against the HBM peak (the x-transforms run as FFTs on the vector units)
        gbs = 9 * lines * M * 16 / (xp["avg_us"] * 1e-6) / 1e9
        out["advection3d"]["xprod_hbm_gbs"] = gbs
        out["advection3d"]["xprod_hbm_frac"] = gbs * 1e9 / 8.0e12
    print(json.dumps({"workload": "pseudo-spectral step, %d^2 CGL x %d modes, mgalx 192" % (cfg.m, M),
                      "steps": steps, **out}), flush=True)
    duct_amd.shutdown()


def cpu_baseline(cfg, fields, alams, threads, reps=3):
    """The oracle on the host cores over the whole workload, `reps` times (the first pass is also
    the checker): about 20 core-seconds of CPU work at 257^2 x 256 modes on 16 cores."""
    import oracle
    from duct_amd import workloads
    blas = oracle.use_blas(True)
    ob = oracle.OracleBasis(oracle.pseudomat(cfg.m))
    t0 = time.perf_counter()
    for rep in range(reps):
        out_r = [oracle.helm3d(ob, fields[f], workloads.BOUNDS[f], alams[f], alp=cfg.alp, ib=0, nthreads=threads)[0]
                 for f in range(cfg.nfields)]
        if rep == 0:
            refs = out_r
    t_all = time.perf_counter() - t0
    k1 = 4
    t0 = time.perf_counter()
    oracle.helm3d(ob, fields[0][:k1], "d", alams[0], alp=cfg.alp, ib=1, nthreads=1)
    t_one = time.perf_counter() - t0
    oracle.use_blas(False)
    out = {
        "value": reps * cfg.nfields * cfg.nmodes / t_all,
        "unit": "mode-solves/s",
        "cores": threads,
        "host_cores": os.cpu_count(),
        "affinity_cores": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
        "kind": "port",
        "sample": "the full workload (%d fields x %d modes, %d^2) %d times with the oracle (C restatement "
                  "of helm3d/int2d3d/solvediag, %s), OpenMP over modes on %d threads; %.2f s wall, "
                  "%.1f core-seconds" % (cfg.nfields, cfg.nmodes, cfg.m, reps, blas, threads, t_all,
                                         t_all * threads),
        "value_1core": k1 / t_one,
    }
    return out, refs


def max_rel_err(sols, refs):
    return max(float(np.max(np.abs(s - r)) / np.max(np.abs(r))) for s, r in zip(sols, refs))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the solve shards over Fourier modes with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe, args.steps)
        return
    if args.c3:
        c3_bench(args.steps, args.warmup)
        return
    import torch

    import duct_amd
    from duct_amd import workloads

    cfg = workloads.CONFIGS[args.config]
    m, M, F = cfg.m, cfg.nmodes, cfg.nfields
    n = m - 2
    s0, s1, niter = args.mixed.split(",")
    niter = int(niter)
    torch.cuda.set_device(0)
    duct_amd.init(0)
    duct_amd.set_overlap(not args.no_overlap)
    duct_amd.setup(m)
    fields = [workloads.rhs_field(cfg, f) for f in range(F)]
    alams = [workloads.alambda_of(cfg, f) for f in range(F)]
    bounds = workloads.BOUNDS[:F]
    rhs = [torch.from_numpy(x).cuda() for x in fields]
    # path -> (tier, bar): the FP64 paths are held to the FP64 bar, the refined solves to the
    # metric's 1e-10
    paths = {"fp64": ("fp64", TOL_FP64), "fp64_fold_d": ("fp64", TOL_FP64), "fp64_fold": ("fp64", TOL_FP64),
             "mixed": ("mixed", TOL), "mixed_f32": ("mixed", TOL), "mixed_fold": ("mixed", TOL),
             "mixed_fold_q": ("mixed", TOL)}
    sol = {p: [torch.zeros_like(r) for r in rhs] for p in paths}

    def fp64_step(path, fold):
        def run():
            duct_amd.set_fold(fold)
            duct_amd.helm3d_batch(rhs, sol[path], bounds, alams, alp=cfg.alp, ib=0)
        return run

    def mixed_step(path, a, b, it):
        return lambda: duct_amd.helm3d_mixed_batch(rhs, sol[path], bounds, alams, alp=cfg.alp, ib=0, scheme0=a,
                                                   scheme1=b, niter=it)

    schemes = {"mixed": (s0, s1, niter), "mixed_f32": ("f32", "f32", args.f32_iter),
               "mixed_fold": ("fold16", "fold16", 1), "mixed_fold_q": ("fold16q", "fold16q", 1)}
    steps = {"fp64": fp64_step("fp64", 0), "fp64_fold_d": fp64_step("fp64_fold_d", 1),
             "fp64_fold": fp64_step("fp64_fold", 2)}
    for p, (a, b, it) in schemes.items():
        steps[p] = mixed_step(p, a, b, it)
    he, ho = (n + 1) // 2, n // 2
    planes = F * (2 * M - 1)                      # Im at k=0 not solved
    nprod = {"f32": 1, "bf16x1": 1, "bf16x3": 3, "bf16x6": 6}
    g = 2.0 * n ** 3 * planes                    # one unfolded product of every plane

    # real planes solved by unfolded and by folded passes on each path: the folded paths give the
    # Poisson modes with xlambda < 64 (kLowXlam of the library) to the unfolded FP64 passes
    def nlow(f):
        return sum(1 for k in range(M) if alams[f] + (cfg.alp * k) ** 2 < 64.0) if bounds[f] == "n" else 0

    def planes_of(k0, k1):
        return sum(1 if k == 0 else 2 for k in range(k0, k1))

    def plane_split(path):
        unf = fol = 0
        for f in range(F):
            if path in ("fp64", "mixed", "mixed_f32") or (path == "fp64_fold_d" and bounds[f] == "n"):
                unf += planes_of(0, M)
            else:
                unf += planes_of(0, nlow(f))
                fol += planes_of(nlow(f), M)
        return unf, fol

    def work_per_step(kind, path):
        """Executed operations of one step of a kernel class and the peak they are held to."""
        unf, fol = plane_split(path)
        per = 2.0 * (he * he + ho * ho) * n       # one folded product of one real plane
        if kind.startswith("proj_"):
            return 2.0 * n ** 3 * unf, "f64"
        if kind.startswith("fold_"):
            return (2.0 if kind == "fold_mid" else 1.0) * per * fol, "f64"   # k_projq: passes 2 and 3
        if path in ("mixed_fold", "mixed_fold_q"):
            if kind == "lp_p2p":                   # k_lpq: 4 passes x 3 fp16 products, two solves
                return 2 * 4 * 3 * per * fol, "f16"
            if kind == "resid":
                if path == "mixed_fold":           # two folded FP64 products
                    return 2.0 * per * fol, "f64"
                return 2 * 18 * per * fol, "i8"    # two products x 18 int8 digit products
            return 0.0, None                       # fold / unfold: bandwidth-bound
        if path not in schemes:
            return 0.0, None                       # gauge, boundary: not matrix-core work
        a, b, it = schemes[path]
        if kind.startswith("lp_"):
            return g * (nprod[a] + it * nprod[b]), ("f32" if a == "f32" else "bf16")
        if kind == "resid":
            return 2.0 * g * it, "f64"             # 4 n^3 per plane, one residual per correction
        return 0.0, None

    res = {}
    # --path X times only X (a profiled run then holds only X's kernels)
    for path in (paths if args.path == "auto" else (args.path,)):
        for _ in range(args.warmup):
            steps[path]()
        torch.cuda.synchronize()
        # the step time: K steps with nothing recorded between the launches
        if not args.events_only:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                steps[path]()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
        # the kernel classes: the same K steps again, with the library's hipEvents around every
        # launch (on its stream); the events add ~0.1 ms of gaps per step, so they stay out of dt
        with duct_amd.KernelTimer() as kt:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                steps[path]()
            torch.cuda.synchronize()
            dt_ev = (time.perf_counter() - t0) / args.steps
        if args.events_only:
            dt = dt_ev
        classes = {}
        for k in kt.ms:
            if not kt.count[k]:
                continue
            classes[k] = {"launches_per_step": kt.count[k] / args.steps, "avg_us": kt.ms[k] / kt.count[k] * 1e3,
                          "ms_per_step": kt.ms[k] / args.steps}
        res[path] = {"ms_per_step": dt * 1e3, "mode_solves_per_s": F * M / dt, "kernels": classes,
                     "ms_per_step_with_events": dt_ev * 1e3, "tier": paths[path][0], "bar": paths[path][1]}
        if path in schemes:
            res[path]["schemes"] = {"scheme0": schemes[path][0], "scheme1": schemes[path][1],
                                    "refinements": schemes[path][2]}
    duct_amd.set_fold(0)

    cpu, refs = None, None
    if not args.no_cpu_baseline:
        share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = args.cpu_threads or (min(share, env) if env > 0 else share)
        cpu, refs = cpu_baseline(cfg, fields, alams, threads)
        for path in res:
            res[path]["max_rel_err_vs_oracle"] = max_rel_err([t.cpu().numpy() for t in sol[path]], refs)
            res[path]["meets_bar"] = res[path]["max_rel_err_vs_oracle"] <= res[path]["bar"]
            res[path]["meets_1e-10"] = res[path]["max_rel_err_vs_oracle"] <= TOL

    def fastest(cands):
        cands = [p for p in cands if p in res]
        return max(cands, key=lambda p: res[p]["mode_solves_per_s"]) if cands else None

    if args.path != "auto":
        chosen = args.path
    elif refs is None:
        chosen = "fp64"                            # nothing verified: the reference's own arithmetic
    else:
        chosen = fastest([p for p in res if res[p]["meets_1e-10"]]) or "fp64"

    # roofline of the chosen path's dominant kernel class (executed operations per launch)
    ks = res[chosen]["kernels"]
    groups = {}
    for k, v in ks.items():
        key = "proj" if k.startswith("proj_") else "lp" if k.startswith("lp_") and chosen not in (
            "mixed_fold", "mixed_fold_q") else "fold" if k.startswith("fold_") else k
        if k == "fold_mid":
            key = "fold_mid"                       # k_projq, its own kernel
        groups.setdefault(key, []).append(k)
    kernel_names = {"proj": "k_proj (FP64 MFMA projection pass)",
                    "fold": "k_projf (FP64 MFMA projection pass, even/odd folded basis)",
                    "fold_mid": "k_projq (FP64 MFMA folded passes 2 and 3 on LDS-resident quarters)",
                    "lp": "k_proj_lp (low-precision MFMA projection pass)",
                    "lp_p2p": "k_lpq (fused per-quarter fp16 hi/lo solve)",
                    "resid": {"mixed_fold": "k_resid_fold (FP64 MFMA folded residual)",
                              "mixed_fold_q": "k_resid_i8 (int8 digit products, FP64-accurate residual)"}.get(
                                  chosen, "k_resid (FP64 MFMA residual)")}

    def roofline_of(gk):
        ms = sum(ks[k]["ms_per_step"] for k in groups[gk])
        launches = sum(ks[k]["launches_per_step"] for k in groups[gk])
        work = sum(work_per_step(k, chosen)[0] for k in groups[gk])
        unit = work_per_step(groups[gk][0], chosen)[1] or "f64"
        avg_s = ms / 1e3 / launches if launches else float("nan")
        per_launch = work / launches if launches else 0.0
        ach = per_launch / avg_s if launches else 0.0
        return {"bound": "mfma", "achieved": ach / 1e12, "peak": PEAK[unit] / 1e12,
                "unit": "TOP/s" if unit == "i8" else "TFLOP/s", "frac": ach / PEAK[unit],
                "traffic": TRAFFIC.get((gk, m, M, F)), "kernel": kernel_names.get(gk, gk), "peak_kind": unit,
                "ops": "executed", "ops_per_launch": per_launch, "avg_launch_us": avg_s * 1e6,
                "ms_per_step": ms, "launches_per_step": launches}

    mfma_groups = [gk for gk in groups if work_per_step(groups[gk][0], chosen)[0] > 0]
    dom = max(groups, key=lambda gk: sum(ks[k]["ms_per_step"] for k in groups[gk]))
    roof = roofline_of(dom)
    # every matrix-core kernel class of the step, longest first
    roof_all = sorted((roofline_of(gk) for gk in mfma_groups), key=lambda r: -r["ms_per_step"])

    fp64_best = fastest([p for p in res if res[p]["tier"] == "fp64" and res[p].get("meets_bar", refs is None)])
    mixed_best = fastest([p for p in res if res[p]["tier"] == "mixed" and res[p].get("meets_bar", False)])
    value = res[chosen]["mode_solves_per_s"]
    dtype = {"fp64": "f64", "fp64_fold_d": "f64", "fp64_fold": "f64",
             "mixed_fold": "fp16 hi/lo folded, f64 residual",
             "mixed_fold_q": "fp16 hi/lo folded, f64-accurate int8-digit residual",
             "mixed_f32": "f32 refined, f64 residual"}.get(chosen, "%s+%s refined, f64 residual" % (s0, s1))
    out = {
        "metric": "Helmholtz mode-solves/s at 257^2 x256 modes: mixed vs FP64 @1e-10 err; %MFMA peak",
        "value": value,
        "unit": "mode-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[chosen]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded Chebyshev-series right-hand sides, inputs resident in HBM)",
        "config": {"workload": "cross-plane solve stage: %d^2 CGL x %d modes x %d fields (u,v,w Dirichlet "
                               "alambda=1000; p Neumann alambda=0)" % (m, M, F),
                   "m": m, "nmodes": M, "nfields": F, "path": chosen, "overlap": not args.no_overlap},
        # what `value` is: the fastest path within the metric's tolerance of the FP64 oracle
        "value_path": chosen,
        "value_tier": res[chosen]["tier"],
        "value_tolerance": res[chosen]["bar"],   # the bar of the chosen path's tier
        "selection_tolerance": TOL,              # value: the fastest path within this of the oracle
        "value_max_rel_err_vs_oracle": res[chosen].get("max_rel_err_vs_oracle"),
        "verified_against_oracle": refs is not None,
        # the FP64 tier at the FP64 bar, and the refined solves against the plain FP64 solve
        "fp64_tier": ({"path": fp64_best, "bar": TOL_FP64, "mode_solves_per_s": res[fp64_best]["mode_solves_per_s"],
                       "max_rel_err_vs_oracle": res[fp64_best].get("max_rel_err_vs_oracle")} if fp64_best else None),
        "mixed_tier": ({"path": mixed_best, "bar": TOL, "mode_solves_per_s": res[mixed_best]["mode_solves_per_s"],
                        "max_rel_err_vs_oracle": res[mixed_best].get("max_rel_err_vs_oracle")} if mixed_best else None),
        "mixed_vs_fp64": (res[mixed_best]["mode_solves_per_s"] / res["fp64"]["mode_solves_per_s"]
                          if mixed_best and "fp64" in res else None),
        # and against the fastest FP64 path at the FP64 bar (the folded solve)
        "mixed_vs_fp64_tier": (res[mixed_best]["mode_solves_per_s"] / res[fp64_best]["mode_solves_per_s"]
                               if mixed_best and fp64_best else None),
        # the stage's algorithmic work (8 n^3 per real plane, four unfolded products) over the step time;
        # the folded and refined paths execute other amounts, which the roofline counts
        "algorithmic_tflops": 4 * g / (res[chosen]["ms_per_step"] * 1e-3) / 1e12,
        # the same over the FP64 MFMA peak: above 1 for the folded paths, which execute half the FLOPs
        "algorithmic_frac_of_f64_peak": 4 * g / (res[chosen]["ms_per_step"] * 1e-3) / PEAK["f64"],
        "roofline": roof,
        "roofline_kernels": roof_all,
        "paths": res,
        "fp64_mfma_frac": None,
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        kf = res["fp64"]["kernels"]
        proj = [k for k in kf if k.startswith("proj_")]
        t = sum(kf[k]["ms_per_step"] for k in proj) / 1e3
        out["fp64_mfma_frac"] = (4 * g / t) / PEAK["f64"] if t else None
    print(json.dumps(out), flush=True)
    duct_amd.shutdown()

if __name__ == "__main__":
    main()
