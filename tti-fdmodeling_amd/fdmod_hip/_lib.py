"""ctypes binding of libfdmod_hip (the C-ABI declared in include/fdmod_hip.h).

The shared library is built in-tree by tti-fdmodeling_amd/Makefile into
tti-fdmodeling_amd/_build/libfdmod_hip.so.  There is no CPU fallback: if the
library is missing, loading fails with an error.
"""
import ctypes as ct
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "_build", "libfdmod_hip.so")

FORMS = {"coord": 0, "tensor": 1, "bond": 2}
ARRAYS = ["vp", "vs", "rho", "eps", "delta", "theta",
          "invrho", "c11", "c13", "c15", "c33", "c35", "c55",
          "sin2", "cos2", "sin2t", "sin", "cos", "sinv", "cosv",
          "D", "Beta", "coef1", "coef2",
          "vx", "vz", "sxx", "szz", "sxz", "txx", "tzz", "txz"]
AID = {n: i for i, n in enumerate(ARRAYS)}
ERRORS = {0: "FDMOD_OK", -1: "FDMOD_EINVAL", -2: "FDMOD_EUNSTABLE", -3: "FDMOD_ERANGE",
          -4: "FDMOD_EDEVICE", -5: "FDMOD_ENOMEM", -6: "FDMOD_ESTATE"}

# Every symbol include/fdmod_hip.h declares, with its ctypes signature.
_fp, _ip, _dp = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int), ct.POINTER(ct.c_double)


class Params(ct.Structure):
    _fields_ = [("form", ct.c_int), ("nz", ct.c_int), ("nx", ct.c_int), ("nt", ct.c_int),
                ("dx", ct.c_float), ("dz", ct.c_float), ("dt", ct.c_float),
                ("ealThick", ct.c_int), ("eal_alpha", ct.c_float), ("freesurface", ct.c_int),
                ("fpeak", ct.c_float), ("scalar", ct.c_float), ("mode", ct.c_int)]


class GradOpts(ct.Structure):
    _fields_ = [("ncheckpoints", ct.c_int), ("storage", ct.c_int), ("history", ct.c_int)]


STORAGE = {"f32": 0, "f16": 1, "f16all": 2}
HISTORY = {"checkpoints": 0, "full": 1, "reconstruct": 2, "reconstruct_interior": 3}  # fdmod_grad_opts.history

SNAPSHOT_FN = ct.CFUNCTYPE(ct.c_int, ct.c_void_p, ct.c_int, ct.c_char_p, _fp, ct.c_int, ct.c_int)

SIGNATURES = {
    "fdmod_create": (ct.c_int, [ct.POINTER(Params), ct.POINTER(ct.c_void_p)]),
    "fdmod_destroy": (None, [ct.c_void_p]),
    "fdmod_last_error": (ct.c_char_p, [ct.c_void_p]),
    "fdmod_strerror": (ct.c_char_p, [ct.c_int]),
    "fdmod_dims": (ct.c_int, [ct.c_void_p, _ip]),
    "fdmod_check_stability": (ct.c_int, [ct.POINTER(Params), _fp, _dp]),
    "fdmod_set_model": (ct.c_int, [ct.c_void_p] + [_fp] * 6 + [_dp]),
    "fdmod_set_source": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip, _ip, _fp]),
    "fdmod_set_receivers": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _ip, _ip]),
    "fdmod_set_snapshots": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, SNAPSHOT_FN, ct.c_void_p]),
    "fdmod_set_energy": (ct.c_int, [ct.c_void_p, ct.c_int]),
    "fdmod_run": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _dp]),
    "fdmod_run_energy_noalloc": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _fp, _fp, _fp, _dp]),
    "fdmod_reset": (ct.c_int, [ct.c_void_p]),
    "fdmod_get_traces": (ct.c_int, [ct.c_void_p, ct.c_int, _fp]),
    "fdmod_get_energy": (ct.c_int, [ct.c_void_p, _fp, _fp, _fp]),
    "fdmod_get_array": (ct.c_int, [ct.c_void_p, ct.c_int, _fp]),
    "fdmod_set_array": (ct.c_int, [ct.c_void_p, ct.c_int, _fp]),
    "fdmod_time_sweeps": (ct.c_int, [ct.c_void_p, ct.c_int, _dp, _dp]),
    "fdmod_batch_create": (ct.c_int, [ct.POINTER(Params), ct.c_int, _ip, ct.POINTER(ct.c_void_p)]),
    "fdmod_batch_destroy": (None, [ct.c_void_p]),
    "fdmod_batch_last_error": (ct.c_char_p, [ct.c_void_p]),
    "fdmod_batch_set_model": (ct.c_int, [ct.c_void_p, ct.c_int] + [_fp] * 6 + [_dp]),
    "fdmod_batch_set_source": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _ip, _ip, _fp]),
    "fdmod_batch_set_receivers": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _ip, _ip]),
    "fdmod_batch_shot": (ct.c_void_p, [ct.c_void_p, ct.c_int]),
    "fdmod_batch_tune": (ct.c_int, [ct.c_void_p, ct.c_int, _ip, _ip]),
    "fdmod_batch_set_tiles": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int]),
    "fdmod_batch_run": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _dp]),
    "fdmod_batch_reset": (ct.c_int, [ct.c_void_p]),
    "fdmod_batch_get_traces": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _fp]),
    "fdmod_gradient": (ct.c_int, [ct.c_void_p, _fp, _fp, _fp, ct.c_int, _dp, _dp]),
    "fdmod_update_model": (ct.c_int, [ct.c_void_p] + [_fp] * 6),
    "fdmod_gradient_opts": (ct.c_int, [ct.c_void_p, _fp, _fp, _fp, ct.POINTER(GradOpts), _dp, _dp]),
    "fdmod_batch_gradient": (ct.c_int, [ct.c_void_p, ct.POINTER(_fp), ct.POINTER(_fp), ct.POINTER(_fp),
                                        ct.POINTER(GradOpts), _ip, ct.c_int, _dp, _dp]),
    "fdmod_adjoint_source": (ct.c_int, [ct.c_void_p, _fp, _fp, _fp, _dp]),
    "fdmod_adjoint_dot": (ct.c_int, [ct.c_void_p, _fp, _fp, _fp, _dp, _dp]),
}

_LIB = None


def load(path=None):
    """Load libfdmod_hip and bind every declared entry point (raises if absent)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libfdmod_hip is not built ({p}); run __graft_entry__.build() or make -C tti-fdmodeling_amd")
    lib = ct.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _LIB = lib
    return lib


class FdmodError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return a.ctypes.data_as(ct.POINTER(t))
