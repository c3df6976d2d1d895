"""ctypes binding of libduct_hip.so (the C-ABI declared in include/duct_hip.h).

The library must be built in-tree (duct_amd/libduct_hip.so, see
__graft_entry__.build()).  There is no CPU fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libduct_hip.so")

# Every symbol include/duct_hip.h declares, with (restype, argtypes).
_vp, _i, _d, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
_pi, _pd, _pc = ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.c_char_p
_psz = ctypes.POINTER(_sz)

SIGNATURES = {
    "duct_init_": (_i, [_pi]),
    "duct_shutdown_": (_i, []),
    "duct_set_stream_": (_i, [_psz]),
    "duct_synchronize_": (_i, []),
    "duct_last_error_": (_i, [ctypes.c_char_p, _pi]),
    "duct_setup_": (_i, [_pc, _pi, _vp, _vp, _vp, _vp, _vp, _pi, _pd, _sz]),
    "duct_helm3d_dev_": (_i, [_psz, _psz, _pi, _pi, _pi, _pd, _pd, _pc, _sz]),
    "duct_helm3d_batch_dev_": (_i, [_pi, _psz, _psz, _pc, _pd, _pi, _pi, _pi, _pd, _sz]),
    "duct_helm3d_batch_noalloc_dev_": (_i, [_pi, _psz, _psz, _pc, _pd, _pi, _pi, _pi, _pd, _sz]),
    "duct_helm3d_": (_i, [_vp, _vp, _pi, _pi, _pi, _pd, _pd, _pc, _sz]),
    "duct_int2d3d_": (_i, [_vp, _vp, _pc, _pd, _pi, _sz]),
    "duct_solvediag_": (_i, [_vp, _pc, _pd, _sz]),
    "duct_helm3d_mixed_batch_dev_": (_i, [_pi, _psz, _psz, _pc, _pd, _pi, _pi, _pi, _pd, _pi, _pi, _pi,
                                          _vp, _sz]),
    "duct_helm3d_adjoint_batch_dev_": (_i, [_pi, _psz, _psz, _pc, _pd, _pi, _pi, _pi, _pd, _sz]),
    "duct_colloc_deriv_dev_": (_i, [_psz, _psz, _pi, _pi, _pc, _pi, _pd, _pd, _sz]),
    "duct_derivx_dev_": (_i, [_psz, _psz, _pi, _pi, _pi, _pd, _pd, _pd]),
    "duct_advection_dev_": (_i, [_psz, _psz, _psz, _psz, _psz, _psz, _pi, _pi, _pi, _pd, _pd]),
    "duct_transyz_dev_": (_i, [_psz, _psz, _pi, _pi, _pc, _pi, _sz]),
    "duct_fourx_dev_": (_i, [_psz, _psz, _pi, _pi, _pi, _pi, _pi]),
    "duct_rk3_substep_dev_": (_i, [_psz] * 7 + [_pi, _pi, _pi, _pd, _pd, _pd, _pd, _pi, _pd]),
    "duct_setup_temp_": (_i, [_pc, _pi, _pi, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz]),
    "duct_helm3d_temp_dev_": (_i, [_psz, _psz, _psz, _psz, _pi, _pi, _pi, _pi, _pd, _pd]),
    "duct_setup_shen_": (_i, [_pc, _pi, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz]),
    "duct_helm3d_shen_dev_": (_i, [_psz, _psz, _pi, _pi, _pi, _pd, _pd, _pc, _sz]),
    "duct_derivyz_dev_": (_i, [_psz, _psz, _pi, _pi, _pc, _pi, _pd, _pc, _pd, _sz, _sz]),
    "duct_transpose_dev_": (_i, [_pi, _pi, _pi, _pd, _psz, _psz]),
    "duct_cxyz2yzx_dev_": (_i, [_pi, _pi, _pi, _psz, _psz]),
    "duct_cyzx2xyz_dev_": (_i, [_pi, _pi, _pi, _psz, _psz]),
    "duct_mat_invdiag_dev_": (_i, [_pi, _pi, _pd, _psz, _psz, _psz]),
    "duct_set_fold_": (_i, [_pi]),
    "duct_set_overlap_": (_i, [_pi]),
    "duct_set_fused_quarters_": (_i, [_pi]),
    "duct_fold_size_": (_i, [_pc, _pi, _sz]),
    "duct_prof_enable_": (_i, [_pi]),
    "duct_h5_read_header_": (_i, [_pc, _pi, _pd, _sz]),
    "duct_h5_read_field_": (_i, [_pc, _pc, _vp, _pi, _pi, _pi, _pi, _sz, _sz]),
    "duct_h5_read_doubles_": (_i, [_pc, _pc, _vp, ctypes.POINTER(ctypes.c_longlong), _sz, _sz]),
    "duct_h5_read_chars_": (_i, [_pc, _pc, _vp, _pi, _pi, _sz, _sz]),
    "duct_prof_read_": (_i, [_pi, _pd, _pi]),
}

STATUS = {0: "OK", 1: "DUCT_ERR_ARG", 2: "DUCT_ERR_NOT_SETUP", 3: "DUCT_ERR_HIP",
          4: "DUCT_ERR_ALLOC", 5: "DUCT_ERR_UNSUPPORTED"}

PROF_CLASSES = ["proj_c2p", "proj_p2p_div", "gauge", "proj_p2p", "proj_p2c", "boundary", "bc_correct",
                "transpose", "invdiag", "lp_c2p", "lp_p2p_div", "lp_p2p", "lp_p2c", "resid", "deriv", "adjoint",
                "fold_c2p", "fold_p2p_div", "fold_p2p", "fold_p2c", "xprod", "tridiag", "fold_mid"]


class DuctError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__("%s -> %s: %s" % (fn, STATUS.get(status, status), msg))
        self.status = status


_lib = None


def open_library(path: str) -> ctypes.CDLL:
    """Open a build of the library and attach the signatures.  Raises if the
    file is missing or lacks a declared symbol: there is no fallback."""
    if not os.path.exists(path):
        raise RuntimeError("libduct_hip.so not built at %s; run __graft_entry__.build()" % path)
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def load() -> ctypes.CDLL:
    """The in-tree library, opened once."""
    global _lib
    if _lib is None:
        _lib = open_library(LIB_PATH)
    return _lib


def last_error() -> str:
    L = load()
    buf = ctypes.create_string_buffer(1024)
    n = _i(len(buf))
    L.duct_last_error_(buf, ctypes.byref(n))
    return buf.value.decode(errors="replace")


def check(fn: str, status: int) -> None:
    if status != 0:
        raise DuctError(fn, status, last_error())


def call(name: str, *args) -> None:
    """Call an entry point and raise DuctError on a nonzero status."""
    check(name, getattr(load(), name)(*args))


def byref_int(v: int):
    return ctypes.byref(_i(int(v)))


def byref_double(v: float):
    return ctypes.byref(_d(float(v)))


def byref_ptr(v: int):
    return ctypes.byref(_sz(int(v)))
