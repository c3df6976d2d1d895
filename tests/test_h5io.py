"""The field-file reader (duct_h5_read_*): HDF5 files of the reference's layout, written by the HDF5
library itself (tests/golden/make_uvwp_h5.c), read back by the library's own parser.

- Direct reads are bit-exact: the fixtures hold dyadic values, recomputed here.
- Interpolation onto another grid (cheb_interp, chebyshev.f90:725-765) is checked against the oracle
  (oracle.spectral.cheb_interp) and, to pin both, against the closed form. The fixture is a sum of
  Chebyshev polynomials: terms of degree <= 19 survive exactly, and the T_20, T_21 and T_22 terms
  fall outside the reference's 20 x 20 coefficient window.
- The other cases follow read_4fields_fou_hdf5 (readwrite_hdf5.f90:2-197): the ib/mode window, modes
  past the file's mx1 (zero), the header with its mx1 fallback, and the temperature file's extra
  fields and walls. Malformed files end in an error, not a crash.
Host code only: no device is needed.
"""
import ctypes
import os
import shutil

import numpy as np
import pytest

from duct_amd import _lib, h5io
from oracle import spectral

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F17 = os.path.join(GOLD, "uvwp_m17.h5")
F25 = os.path.join(GOLD, "uvwp_m25_poly.h5")
FT9 = os.path.join(GOLD, "uvwpt_m9.h5")


def exact_field(f, nm, m):
    """The generator's exact_value for field f: [mode, z, y] complex."""
    k = np.arange(nm)[:, None, None]
    z = np.arange(m)[None, :, None]
    y = np.arange(m)[None, None, :]
    v = (f + 1) * 1.0 + k * 0.5 + z * 0.03125 + y * 0.0009765625
    return v - 1j * np.where(k == 0, 0.0, v + 0.25)        # mode 0 real


def cheb(n, x):
    return np.cos(n * np.arccos(np.clip(x, -1.0, 1.0)))


def poly(f, k, part, Y, Z, truncated):
    """The generator's poly_value; truncated drops the terms of degree >= 20 in y or z."""
    s = (f + 1) * (k + 1) * ((-0.5 if k else 0.0) if part else 1.0)      # mode 0 real
    v = (1.0 + 0.5 * cheb(1, Y) - 0.25 * cheb(2, Z) + 0.125 * cheb(3, Y) * cheb(1, Z)
         + 0.0625 * cheb(4, Y) * cheb(4, Z) + 0.03125 * cheb(19, Y) * cheb(2, Z))
    if not truncated:
        v = v + 0.1 * cheb(20, Z) + 0.3 * cheb(22, Y) + 0.2 * cheb(1, Y) * cheb(21, Z)
    return s * v


def poly_field(f, nm, m, truncated):
    x = -np.cos(np.pi * np.arange(m) / (m - 1))
    Y, Z = np.meshgrid(x, x)                      # [z, y]
    return np.stack([poly(f, k, 0, Y, Z, truncated) + 1j * poly(f, k, 1, Y, Z, truncated)
                     for k in range(nm)])


def test_header_mx1_fallback():
    """A header without mx1 gives 2*(mgalx/3)/2 - 1 (readwrite_hdf5.f90:824-826)."""
    h = h5io.read_header(F25)
    assert (h["mgalx"], h["mx1"], h["my"], h["mz"], h["time"]) == (6, 1, 25, 25, 3.0)
    assert not _lib.load().duct_h5_read_doubles_(F25.encode(), b"mgalx", np.zeros(1).ctypes.data_as(
        ctypes.c_void_p), ctypes.byref(ctypes.c_longlong(1)), len(F25), 5)
    with pytest.raises(_lib.DuctError, match="no dataset 'mx1'"):
        h5io.read_doubles(F25, "mx1", (1,))


def test_header_default_format():
    h = h5io.read_header(F17)
    assert (h["mgalx"], h["mx1"], h["my"], h["mz"], h["nparams"]) == (9, 2, 17, 17, 7)
    assert (h["alp"], h["aspect"], h["time"], h["fnu"]) == (1.0, 1.5, 12.5, 0.00125)
    assert (h["cfl"], h["dtfixed"], h["cmassflow"]) == (0.5, 0.01, 1.0)


@pytest.mark.parametrize("f,name", list(enumerate("uvwp")))
def test_fields_bit_exact(f, name):
    got = h5io.read_field(F17, name, 17)
    assert got.shape == (3, 17, 17)
    np.testing.assert_array_equal(got, exact_field(f, 3, 17))


@pytest.mark.parametrize("ib,nmodes", [(0, 1), (1, 2), (2, 1), (1, 4), (3, 2), (0, 5)])
def test_mode_window(ib, nmodes):
    """Modes ib .. ib+nmodes-1; those past the file's mx1 = 2 are zero (readwrite_hdf5.f90:120-128)."""
    got = h5io.read_field(F17, "v", 17, ib=ib, nmodes=nmodes)
    want = np.zeros((nmodes, 17, 17), complex)
    full = exact_field(1, 3, 17)
    n = max(0, min(nmodes, 3 - ib))
    want[:n] = full[ib:ib + n]
    np.testing.assert_array_equal(got, want)


def test_temperature_file_newest_format():
    """Superblock 3, version-2 object headers and compact links behind a 512-byte user block."""
    h = h5io.read_header(FT9)
    assert (h["mgalx"], h["mx1"], h["my"], h["mz"], h["nparams"]) == (6, 1, 9, 9, 13)
    fields = h5io.read_fields(FT9)
    assert sorted(fields) == sorted("uvwpt")
    for f, c in enumerate("uvwpt"):
        np.testing.assert_array_equal(fields[c], exact_field(f, 2, 9))
    bounds, bcy, bcz = h5io.read_walls(FT9)
    assert bounds == "dndn"
    k = np.arange(2)[:, None, None]
    w = np.arange(2)[None, :, None]
    i = np.arange(9)[None, None, :]
    np.testing.assert_array_equal(bcy, 100.0 + k * 8.0 + w * 4.0 + i * 0.125)
    np.testing.assert_array_equal(bcz, -200.0 - k * 8.0 - w * 4.0 - i * 0.0625)
    assert h5io.read_doubles(FT9, "gravy", (1,))[0] == -1.0
    assert h5io.read_doubles(FT9, "numerop", (1,))[0] == 1.0


def test_oracle_interp_closed_form():
    """The oracle's cheb_interp reproduces the polynomial without its degree >= 20 terms."""
    raw = poly_field(2, 2, 25, truncated=False)
    for k in range(2):
        got = spectral.cheb_interp(raw[k], 17, 17)
        np.testing.assert_allclose(got, poly_field(2, 2, 17, truncated=True)[k], rtol=0, atol=1e-13)


@pytest.mark.parametrize("m", [17, 33])
def test_interp_matches_oracle(m):
    raw = h5io.read_field(F25, "w", 25)
    # the generator's libm and numpy evaluate the polynomial with different roundings
    np.testing.assert_allclose(raw, poly_field(2, 2, 25, truncated=False), rtol=0, atol=1e-13)
    got = h5io.read_field(F25, "w", m)
    assert got.shape == (2, m, m)
    for k in range(2):
        np.testing.assert_allclose(got[k], spectral.cheb_interp(raw[k], m, m), rtol=0, atol=1e-13)
    np.testing.assert_allclose(got, poly_field(2, 2, m, truncated=True), rtol=0, atol=1e-12)


def test_interp_rectangular_and_window():
    got = h5io.read_field(F25, "u", 13, mz=21, ib=1, nmodes=3)
    raw = h5io.read_field(F25, "u", 25)
    np.testing.assert_allclose(got[0], spectral.cheb_interp(raw[1], 13, 21), rtol=0, atol=1e-13)
    np.testing.assert_array_equal(got[1:], 0)


def test_one_grid_size_differs_is_unsupported():
    with pytest.raises(_lib.DuctError) as e:
        h5io.read_field(F25, "u", 25, mz=17)
    assert e.value.status == 5


def test_fortran_strings_trailing_blanks():
    """Fortran passes blank-padded characters with their hidden lengths."""
    path = F17.encode() + b" " * 7
    name = b"p   "
    out = np.zeros((3, 17, 17), complex)
    i = ctypes.c_int
    st = _lib.load().duct_h5_read_field_(path, name, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(17)),
                                         ctypes.byref(i(17)), ctypes.byref(i(0)), ctypes.byref(i(3)),
                                         len(path), len(name))
    assert st == 0
    np.testing.assert_array_equal(out, exact_field(3, 3, 17))


def test_errors(tmp_path):
    with pytest.raises(_lib.DuctError, match="no dataset 't'"):
        h5io.read_field(F17, "t", 17)
    with pytest.raises(_lib.DuctError, match="cannot open"):
        h5io.read_header(str(tmp_path / "missing.h5"))
    junk = tmp_path / "junk.h5"
    junk.write_bytes(b"not an hdf5 file" * 100)
    with pytest.raises(_lib.DuctError, match="superblock"):
        h5io.read_header(str(junk))
    with pytest.raises(_lib.DuctError, match="elements"):
        h5io.read_doubles(F17, "alp", (2,))
    with pytest.raises(_lib.DuctError, match="not a Fourier-space field"):
        h5io.read_field(F17, "alp", 17)
    with pytest.raises(_lib.DuctError):
        h5io.read_field(F17, "u", 17, ib=-1, nmodes=1)


@pytest.mark.parametrize("keep", [100, 1000, 4000, 20000, 60000])
def test_truncated_file_is_an_error(tmp_path, keep):
    """A cut-off file ends in DUCT_ERR_ARG (read past the end), never a crash or garbage."""
    cut = tmp_path / "cut.h5"
    with open(F17, "rb") as src:
        cut.write_bytes(src.read(keep))
    with pytest.raises(_lib.DuctError):
        h5io.read_fields(str(cut))


def test_copy_reads_the_same(tmp_path):
    """Addresses are file-relative: a copy under another name reads identically."""
    dst = tmp_path / "uvwp_0007.h5"
    shutil.copy(F17, dst)
    np.testing.assert_array_equal(h5io.read_field(str(dst), "u", 17), exact_field(0, 3, 17))


def test_short_layout_message_is_an_error(tmp_path):
    """A data layout message too short for the fields its version and class imply (here: a compact
    layout of 2 bytes, whose size field would sit at bytes 2-3) is reported as truncated; the parser
    reads no byte past the message. The object header stays walkable: the bytes the message gives
    up become a NIL message."""
    data = bytearray(open(F17, "rb").read())
    at = data.find(bytes([0x08, 0x00, 0x18, 0x00, 0, 0, 0, 0, 0x03, 0x01]))   # v1 header, layout v3 contiguous
    assert at > 0
    data[at + 2:at + 4] = (2).to_bytes(2, "little")        # message size 2
    data[at + 8:at + 10] = bytes([0x03, 0x00])             # version 3, class 0 (compact)
    nil = at + 8 + 2                                       # 22 bytes left of the original 32
    data[nil:nil + 8] = bytes([0x00, 0x00]) + (14).to_bytes(2, "little") + bytes(4)
    data[nil + 8:nil + 22] = bytes(14)
    bad = tmp_path / "short_layout.h5"
    bad.write_bytes(bytes(data))
    with pytest.raises(_lib.DuctError, match="truncated layout message"):
        h5io.read_fields(str(bad))
