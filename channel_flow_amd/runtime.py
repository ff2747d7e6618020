"""ctypes binding of libchf.so (the C ABI declared in include/chf.h).

The library is loaded from channel_flow_amd/lib/libchf.so; there is no other
implementation of the timestep: if the library or a HIP device is missing the
calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libchf.so")

EXPORTS = ("chf_create", "chf_destroy", "chf_last_error", "chf_set_state", "chf_get_state",
           "chf_init_fields", "chf_nonlinear", "chf_solve", "chf_step", "chf_sync",
           "chf_get_fields", "chf_dims", "chf_device_bytes", "chf_set_timing", "chf_kernel_count",
           "chf_kernel_name", "chf_timing", "chf_ensemble_create", "chf_ensemble_destroy",
           "chf_ensemble_last_error", "chf_ensemble_size", "chf_ensemble_member", "chf_ensemble_step",
           "chf_physical", "chf_physical_noalloc", "chf_statistics", "chf_tile_choice",
           "chf_set_tile_choice")


class ChfParams(ctypes.Structure):
    _fields_ = [("nnx", ctypes.c_int), ("nny", ctypes.c_int), ("nnz", ctypes.c_int),
                ("aphi", ctypes.c_double), ("beta", ctypes.c_double), ("dt", ctypes.c_double),
                ("re", ctypes.c_double), ("re_p", ctypes.c_double), ("precision", ctypes.c_int)]


STATE_FIELDS = ("clamb1n", "clamb1nm1", "clamb3n", "clamb3nm1", "au0", "w0",
                "vnp1", "omegaynp1", "rhsvn", "rhsomegayn", "unp1", "wnp1")
MEAN_FIELDS = STATE_FIELDS[:6]


class ChfState(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in STATE_FIELDS]


_lib = None


def load():
    """Load libchf.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m channel_flow_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    lib.chf_create.argtypes = [ctypes.POINTER(ChfParams), ctypes.POINTER(vp)]
    lib.chf_destroy.argtypes = [vp]
    lib.chf_last_error.argtypes = [vp]
    lib.chf_last_error.restype = ctypes.c_char_p
    lib.chf_set_state.argtypes = [vp, ctypes.POINTER(ChfState)]
    lib.chf_get_state.argtypes = [vp, ctypes.POINTER(ChfState)]
    lib.chf_init_fields.argtypes = [vp, dp, dp, dp, dp]
    lib.chf_nonlinear.argtypes = [vp]
    lib.chf_solve.argtypes = [vp]
    lib.chf_step.argtypes = [vp, ctypes.c_int]
    lib.chf_sync.argtypes = [vp]
    lib.chf_get_fields.argtypes = [vp, dp, dp, dp, dp]
    lib.chf_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int)]
    lib.chf_device_bytes.argtypes = [vp]
    lib.chf_device_bytes.restype = ctypes.c_longlong
    lib.chf_set_timing.argtypes = [vp, ctypes.c_int]
    lib.chf_kernel_count.argtypes = []
    lib.chf_kernel_name.argtypes = [ctypes.c_int]
    lib.chf_kernel_name.restype = ctypes.c_char_p
    lib.chf_timing.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_int)]
    lib.chf_tile_choice.argtypes = [vp, ctypes.c_int, dp]
    lib.chf_set_tile_choice.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.chf_physical.argtypes = [vp, ctypes.c_int, dp]
    lib.chf_physical_noalloc.argtypes = [vp, ctypes.c_int, vp]
    lib.chf_statistics.argtypes = [vp, dp]
    lib.chf_ensemble_create.argtypes = [ctypes.POINTER(ChfParams), ctypes.c_int, ctypes.POINTER(vp)]
    lib.chf_ensemble_destroy.argtypes = [vp]
    lib.chf_ensemble_last_error.argtypes = [vp]
    lib.chf_ensemble_last_error.restype = ctypes.c_char_p
    lib.chf_ensemble_size.argtypes = [vp]
    lib.chf_ensemble_member.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    lib.chf_ensemble_step.argtypes = [vp, ctypes.c_int]
    _lib = lib
    return lib


def device_count() -> int:
    """Number of visible HIP devices (0 when none or no runtime)."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def dptr(a: np.ndarray):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


NONFINITE = -2        # CHF_NONFINITE: the state holds a NaN or an infinity


class ChfError(RuntimeError):
    """A failed libchf call; `code` is its status (see include/chf.h)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class ChfNonFinite(ChfError):
    """The state became (or was loaded) non-finite: a blow-up."""


def error(msg, code):
    return (ChfNonFinite if code == NONFINITE else ChfError)(msg, code)
