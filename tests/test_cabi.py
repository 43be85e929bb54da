"""The C-ABI library loads and exports every entry point include/fdmod_hip.h
declares; the host-only entry points work without a GPU."""
import ctypes as ct
import os
import re

import numpy as np

from fdmod_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "fdmod_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fdmod_[a-z_0-9]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n


def test_no_oracle_symbols_in_product_library():
    lib = ct.CDLL(_lib.LIB_PATH)
    for n in ("oracle32_run", "oracle64_run", "oracle32_create"):
        assert not hasattr(lib, n)


def test_strerror_and_host_stability_check():
    lib = _lib.load()
    assert lib.fdmod_strerror(-2).decode().startswith("finite-difference")
    p = _lib.Params(2, 6, 7, 10, 10.0, 10.0, 1e-3, 5, -1.0, 0, 15.0, 1.0, 1)
    vp = np.full(42, 4800.0, np.float32)
    dtm = ct.c_double()
    rc = lib.fdmod_check_stability(ct.byref(p), vp.ctypes.data_as(ct.POINTER(ct.c_float)), ct.byref(dtm))
    assert rc == 0 and abs(dtm.value - 1.30604e-3) < 2e-8
    p.dt = 1.4e-3
    rc = lib.fdmod_check_stability(ct.byref(p), vp.ctypes.data_as(ct.POINTER(ct.c_float)), ct.byref(dtm))
    assert rc == -2


def test_create_rejects_bad_parameters_before_device():
    lib = _lib.load()
    h = ct.c_void_p()
    for bad in (dict(form=5), dict(nz=0), dict(dt=-1.0), dict(mode=3), dict(freesurface=2)):
        kw = dict(form=2, nz=6, nx=7, nt=10, dx=10.0, dz=10.0, dt=1e-3, ealThick=5, eal_alpha=-1.0, freesurface=0,
                  fpeak=15.0, scalar=1.0, mode=1)
        kw.update(bad)
        p = _lib.Params(*kw.values())
        assert lib.fdmod_create(ct.byref(p), ct.byref(h)) == -1, bad
