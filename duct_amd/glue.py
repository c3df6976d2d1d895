"""Python mirrors of the reference's GPU glue entry points (torch device tensors).

  transpose_gpu(A)            transpose_gpu_  (src/mytranspose_cuda.cu:132-162)
  mat_invdiag_gpu(F, ED2, xl) mat_invdiag_gpu_ (src/mytranspose_cuda.cu:181-196)
  cxyz2yzx(phys) / cyzx2xyz   src/parallel.F:476-529 (Fourier <-> physical pencils)

Tensors are C-ordered torch views of Fortran column-major arrays: a Fortran
A(Nx, Ny) is a torch tensor of shape (Ny, Nx); phys(0:mx1, my, nz) complex is
a complex128 tensor of shape (nz, my, nx).
"""
from __future__ import annotations

from . import _lib
from .solver import _use_current_stream


def _dev(t, dtype):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise TypeError("expected a contiguous %s CUDA tensor" % dtype)


def transpose_gpu(A, alpha: float = 1.0):
    """A: (..., Ny, Nx) float64 (Fortran A(Nx,Ny) per batch) -> (..., Nx, Ny) holding alpha*A^T."""
    import torch
    _dev(A, torch.float64)
    Ny, Nx = A.shape[-2], A.shape[-1]
    nb = A.numel() // max(1, Nx * Ny)
    C = torch.empty(A.shape[:-2] + (Nx, Ny), dtype=A.dtype, device=A.device)
    _use_current_stream()
    _lib.call("duct_transpose_dev_", _lib.byref_int(Nx), _lib.byref_int(Ny), _lib.byref_int(nb),
              _lib.byref_double(alpha), _lib.byref_ptr(A.data_ptr()), _lib.byref_ptr(C.data_ptr()))
    return C


def mat_invdiag_gpu(F, ED2, xlam: float):
    """C = F / (ED2 - xlam) elementwise (float64 tensors of equal shape)."""
    import torch
    _dev(F, torch.float64)
    _dev(ED2, torch.float64)
    if F.shape != ED2.shape:
        raise ValueError("F and ED2 shapes differ")
    C = torch.empty_like(F)
    Nx = F.shape[-1] if F.dim() else 1
    Ny = F.numel() // max(1, Nx)
    _use_current_stream()
    _lib.call("duct_mat_invdiag_dev_", _lib.byref_int(Nx), _lib.byref_int(Ny), _lib.byref_double(xlam),
              _lib.byref_ptr(F.data_ptr()), _lib.byref_ptr(ED2.data_ptr()), _lib.byref_ptr(C.data_ptr()))
    return C


def cxyz2yzx(phys):
    """phys (nz, my, nx) complex128 [= phys(0:mx1, my, kb:ke)] -> phys2 (nx, nz, my) [= phys2(my, kb:ke, 0:mx1)]."""
    import torch
    _dev(phys, torch.complex128)
    nz, my, nx = phys.shape
    out = torch.empty((nx, nz, my), dtype=phys.dtype, device=phys.device)
    _use_current_stream()
    _lib.call("duct_cxyz2yzx_dev_", _lib.byref_int(nx), _lib.byref_int(my), _lib.byref_int(nz),
              _lib.byref_ptr(phys.data_ptr()), _lib.byref_ptr(out.data_ptr()))
    return out


def cyzx2xyz(phys2):
    """Inverse of cxyz2yzx: phys2 (nx, nz, my) -> phys (nz, my, nx)."""
    import torch
    _dev(phys2, torch.complex128)
    nx, nz, my = phys2.shape
    out = torch.empty((nz, my, nx), dtype=phys2.dtype, device=phys2.device)
    _use_current_stream()
    _lib.call("duct_cyzx2xyz_dev_", _lib.byref_int(nx), _lib.byref_int(my), _lib.byref_int(nz),
              _lib.byref_ptr(phys2.data_ptr()), _lib.byref_ptr(out.data_ptr()))
    return out
