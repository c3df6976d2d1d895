"""duct_amd - MI355X (gfx950) implementation of duct's per-Fourier-mode
cross-plane Helmholtz/Poisson solve stage (solvediag matrix diagonalisation).

Layout
  csrc/        HIP kernels (FP64 MFMA projection passes, boundary kernels) and
               the extern "C" boundary compiled into libduct_hip.so
  _lib.py      ctypes binding of that C-ABI (include/duct_hip.h)
  basis.py     one-time FP64 host eigensolve (the pseudomat step)
  solver.py    helm3d / int2d3d / solvediag mirrors over the library
  workloads.py seeded synthetic right-hand sides of the benchmark configs
  h5io.py      reading the reference's uvwp_NNNN.h5 field files (host code in the library)
"""
from . import _lib  # noqa: F401
from .solver import (  # noqa: F401
    KernelTimer, fold_size, helm3d, helm3d_batch, helm3d_host, helm3d_mixed_batch, init, int2d3d, set_fold,
    set_fused_quarters, set_overlap, set_stream, setup, shutdown, solvediag, synchronize, upload_basis,
)
