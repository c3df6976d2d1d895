"""Reading the Fourier-space field files (uvwp_NNNN.h5, uvwpt_NNNN.h5) through the library.

The reference writes one HDF5 file per snapshot (save_4fields_fou_hdf5, readwrite_hdf5.f90:266-381):
datasets "u", "v", "w", "p" (and "t" with temperature) of Fortran extents (2, my, mz, mx1+1), and a
header of one-element datasets (writeheader_hdf5, :853-918). read_4fields_fou_hdf5 (:2-197) reads them
back, interpolating onto the run's grid when my and mz both differ (cheb_interp, chebyshev.f90:725-765).

The library parses the files itself (no libhdf5 is needed); these wrappers return numpy arrays in the
layout of the rest of the package: complex128 (nmodes, mz, my) = [mode, z, y].
"""
import ctypes
import os

import numpy as np

from . import _lib

_REAL_KEYS = ("alp", "aspect", "time", "fnu", "cfl", "dtfixed", "cmassflow")


def _b(s: str) -> bytes:
    return os.fsencode(s) if not isinstance(s, bytes) else s


def read_header(path: str) -> dict:
    """The header (readheader_hdf5, readwrite_hdf5.f90:798-850): mgalx, mx1, my, mz, nparams and
    alp, aspect, time, fnu, cfl, dtfixed, cmassflow (NaN when the file lacks an optional one)."""
    p = _b(path)
    ints = (ctypes.c_int * 5)()
    reals = (ctypes.c_double * 7)()
    _lib.call("duct_h5_read_header_", p, ints, reals, len(p))
    h = dict(zip(("mgalx", "mx1", "my", "mz", "nparams"), list(ints)))
    h.update(zip(_REAL_KEYS, list(reals)))
    return h


def read_field(path: str, name: str, my: int, mz: int = None, ib: int = 0, nmodes: int = None) -> np.ndarray:
    """Field `name` as complex128 (nmodes, mz, my) for the modes ib .. ib+nmodes-1 (nmodes defaults to
    the file's mx1+1-ib). Modes past the file's mx1 are zero; a file of another grid is interpolated
    as the reference does."""
    mz = my if mz is None else mz
    if nmodes is None:
        nmodes = read_header(path)["mx1"] + 1 - ib
    if nmodes < 1:
        raise ValueError("nmodes=%d < 1" % nmodes)
    out = np.zeros((nmodes, mz, my), dtype=np.complex128)
    p, n = _b(path), _b(name)
    _lib.call("duct_h5_read_field_", p, n, out.ctypes.data_as(ctypes.c_void_p), _lib.byref_int(my),
              _lib.byref_int(mz), _lib.byref_int(ib), _lib.byref_int(nmodes), len(p), len(n))
    return out


def read_fields(path: str, my: int = None, mz: int = None, ib: int = 0, nmodes: int = None,
                names: str = None) -> dict:
    """The fields of a snapshot, as read_4fields_fou_hdf5 / read_5fields_fou_hdf5 read them: "uvwp",
    or "uvwpt" for a temperature file (nparams 13), onto the grid (my, mz) (default: the file's)."""
    h = read_header(path)
    my = h["my"] if my is None else my
    mz = my if mz is None else mz
    if names is None:
        names = "uvwpt" if h["nparams"] == 13 else "uvwp"
    if nmodes is None:
        nmodes = h["mx1"] + 1 - ib
    return {c: read_field(path, c, my, mz, ib, nmodes) for c in names}


def read_doubles(path: str, name: str, shape) -> np.ndarray:
    """A numeric dataset (header scalar, bcy, bcz) as float64 of the given C-order shape, which must
    hold exactly the dataset's elements."""
    out = np.zeros(shape, dtype=np.float64)
    p, n = _b(path), _b(name)
    cnt = ctypes.c_longlong(out.size)
    _lib.call("duct_h5_read_doubles_", p, n, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cnt),
              len(p), len(n))
    return out


def read_chars(path: str, name: str, capacity: int = 256) -> str:
    """A character dataset (the wall types boundy1 ... boundz2 of a temperature file)."""
    buf = ctypes.create_string_buffer(capacity)
    length = ctypes.c_int(0)
    p, n = _b(path), _b(name)
    _lib.call("duct_h5_read_chars_", p, n, buf, _lib.byref_int(capacity), ctypes.byref(length), len(p), len(n))
    return buf.raw[:length.value].decode("latin-1")


def read_walls(path: str, nmodes: int = None, ib: int = 0):
    """The temperature file's wall types and data: (bounds, bcy, bcz), bounds = boundy1 + boundy2 +
    boundz1 + boundz2 (e.g. "dndn"), bcy float64 (nmodes, 2, mz), bcz (nmodes, 2, my) for the modes
    ib .. ib+nmodes-1: the layout duct_amd.temperature.helm3d_temp takes."""
    h = read_header(path)
    M = h["mx1"] + 1
    bounds = "".join(read_chars(path, k).strip() or " " for k in ("boundy1", "boundy2", "boundz1", "boundz2"))
    bcy = read_doubles(path, "bcy", (M, 2, h["mz"]))
    bcz = read_doubles(path, "bcz", (M, 2, h["my"]))
    nmodes = M - ib if nmodes is None else nmodes
    return bounds, bcy[ib:ib + nmodes], bcz[ib:ib + nmodes]
