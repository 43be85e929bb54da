"""ctypes wrapper for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads liboracle_f32.so (the FP32 restatement of the
reference's OpenMP program) or liboracle_f64.so (the FP64 oracle), built by
oracle/Makefile from oracle/fdmod_oracle.c.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_BUILD_DIR selects another build of the same sources (the sanitizer build)
BUILD = os.environ.get("ORACLE_BUILD_DIR") or os.path.join(HERE, "_build")

FORMS = {"coord": 0, "tensor": 1, "bond": 2}
ARRAYS = ["vp", "vs", "rho", "eps", "delta", "theta",
          "invrho", "c11", "c13", "c15", "c33", "c35", "c55",
          "sin2", "cos2", "sin2t", "sin", "cos", "sinv", "cosv",
          "D", "Beta", "coef1", "coef2",
          "vx", "vz", "sxx", "szz", "sxz", "txx", "tzz", "txz"]
AID = {n: i for i, n in enumerate(ARRAYS)}

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib(prec):
    if prec in _libs:
        return _libs[prec]
    name = {"f32": "liboracle_f32.so", "f64": "liboracle_f64.so"}[prec]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        build()
    lib = ct.CDLL(path)
    p = "oracle32_" if prec == "f32" else "oracle64_"
    fp, ip, dp = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int), ct.POINTER(ct.c_double)
    sig = {
        "create": (ct.c_void_p, [ct.c_int] * 4 + [ct.c_float] * 3 + [ct.c_int, ct.c_float, ct.c_int,
                                                                       ct.c_float, ct.c_float, ct.c_int]),
        "destroy": (None, [ct.c_void_p]),
        "last_error": (ct.c_char_p, [ct.c_void_p]),
        "dims": (ct.c_int, [ct.c_void_p, ip]),
        "set_model": (ct.c_int, [ct.c_void_p] + [fp] * 6 + [dp]),
        "set_source": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ip, ip, fp]),
        "set_receivers": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ip, ip]),
        "set_energy": (ct.c_int, [ct.c_void_p, ct.c_int]),
        "run": (ct.c_double, [ct.c_void_p, ct.c_int, ct.c_int]),
        "v_sweep": (None, [ct.c_void_p]),
        "s_sweep": (None, [ct.c_void_p]),
        "free_surface": (None, [ct.c_void_p]),
        "sources": (None, [ct.c_void_p, ct.c_int, ct.c_int]),
        "get_array": (ct.c_int, [ct.c_void_p, ct.c_int, dp]),
        "set_array": (ct.c_int, [ct.c_void_p, ct.c_int, dp]),
        "get_traces": (ct.c_int, [ct.c_void_p, ct.c_int, dp]),
        "get_energy": (ct.c_int, [ct.c_void_p, dp, dp, dp]),
        "set_threads": (ct.c_int, [ct.c_int]),
        "gradient": (ct.c_int, [ct.c_void_p, fp, fp, fp, dp, dp, dp]),
        "adjoint_source": (ct.c_int, [ct.c_void_p, dp, dp, dp, dp]),
        "set_model_frozen": (ct.c_int, [ct.c_void_p] + [fp] * 6),
    }
    fns = {}
    for k, (res, args) in sig.items():
        f = getattr(lib, p + k)
        f.restype = res
        f.argtypes = args
        fns[k] = f
    _libs[prec] = fns
    return fns


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return a.ctypes.data_as(ct.POINTER(t))


class Oracle:
    """One modelling run on the CPU oracle, mirroring the libfdmod_hip context API.

    Arrays are in the reference layout: model files and padded fields are
    [ix][iz] with iz fastest; returned padded arrays have shape (Nx, Nz).
    """

    def __init__(self, form, nz, nx, nt, dx, dz, dt, ealThick=30, eal_alpha=-1.0, freesurface=0,
                 fpeak=10.0, scalar=1.0, mode=1, prec="f32"):
        self.L = _lib(prec)
        self.prec = prec
        self.form = form
        self.nz, self.nx, self.nt = nz, nx, nt
        self.c = self.L["create"](FORMS[form], nz, nx, nt, dx, dz, dt, ealThick, eal_alpha, freesurface,
                                  fpeak, scalar, mode)
        d = (ct.c_int * 7)()
        self.L["dims"](self.c, d)
        self.Nz, self.Nx, self.e, self.izStart, self.izEnd, self.ixStart, self.ixEnd = list(d)
        self.nr = 0
        self.ntsample = 0

    def close(self):
        if getattr(self, "c", None):
            self.L["destroy"](self.c)
            self.c = None

    def __del__(self):
        self.close()

    def set_model(self, vp, vs, rho, eps, delta, theta):
        arrs = [_f32(a).reshape(-1) for a in (vp, vs, rho, eps, delta, theta)]
        for a in arrs:
            assert a.size == self.nz * self.nx
        dtm = ct.c_double(0)
        rc = self.L["set_model"](self.c, *[_ptr(a, ct.c_float) for a in arrs], ct.byref(dtm))
        self.dt_max = dtm.value
        if rc != 0:
            raise ValueError(self.L["last_error"](self.c).decode())

    def set_source(self, stype, posx, posz, src):
        posx = np.ascontiguousarray(posx, dtype=np.int32)
        posz = np.ascontiguousarray(posz, dtype=np.int32)
        src = _f32(src).reshape(len(posx), self.nt)
        self.L["set_source"](self.c, stype, len(posx), _ptr(posx, ct.c_int), _ptr(posz, ct.c_int),
                             _ptr(src, ct.c_float))

    def set_receivers(self, seistype, mt, posx, posz):
        posx = np.ascontiguousarray(posx, dtype=np.int32)
        posz = np.ascontiguousarray(posz, dtype=np.int32)
        self.nr = len(posx)
        self.ntsample = (self.nt - 1) // mt + 1
        self.L["set_receivers"](self.c, seistype, mt, len(posx), _ptr(posx, ct.c_int), _ptr(posz, ct.c_int))

    def set_energy(self, on=1):
        self.L["set_energy"](self.c, on)

    def run(self, k0=0, count=None):
        if count is None:
            count = self.nt - k0
        return self.L["run"](self.c, k0, count)

    def v_sweep(self):
        self.L["v_sweep"](self.c)

    def s_sweep(self):
        self.L["s_sweep"](self.c)

    def free_surface(self):
        self.L["free_surface"](self.c)

    def sources(self, it, on_stress):
        self.L["sources"](self.c, it, on_stress)

    def get(self, name):
        out = np.empty(self.Nx * self.Nz, dtype=np.float64)
        self.L["get_array"](self.c, AID[name], _ptr(out, ct.c_double))
        return out.reshape(self.Nx, self.Nz)

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert a.size == self.Nx * self.Nz
        self.L["set_array"](self.c, AID[name], _ptr(a, ct.c_double))

    def traces(self, comp):
        """comp: 1 = vx, 3 = vz, 4 = p (the reference's traces2d_output flags)."""
        out = np.empty(self.nr * self.ntsample, dtype=np.float64)
        rc = self.L["get_traces"](self.c, comp, _ptr(out, ct.c_double))
        assert rc == 0
        return out.reshape(self.nr, self.ntsample)

    def set_model_frozen(self, vp, vs, rho, eps, delta, theta):
        """Replace the model keeping the absorbing layers of the last set_model."""
        arrs = [_f32(a).reshape(-1) for a in (vp, vs, rho, eps, delta, theta)]
        for a in arrs:
            assert a.size == self.nz * self.nx
        self.L["set_model_frozen"](self.c, *[_ptr(a, ct.c_float) for a in arrs])

    def gradient(self, obs_vx=None, obs_vz=None, obs_p=None, want_gc=False):
        """J = 1/2 sum (d - obs)^2 and dJ/d(vp, eps, delta, theta[deg]) as (4, nx, nz),
        runs the forward from rest (all three forms).  want_gc also returns the
        imaging sums of the per-point arrays the loop reads, (8, Nx, Nz)."""
        J = ct.c_double(0)
        g = np.zeros(4 * self.nx * self.nz)
        gc = np.zeros(8 * self.Nx * self.Nz) if want_gc else None
        obs = [None if o is None else _f32(o).reshape(-1) for o in (obs_vx, obs_vz, obs_p)]
        rc = self.L["gradient"](self.c, *[None if o is None else _ptr(o, ct.c_float) for o in obs], ct.byref(J),
                                _ptr(g, ct.c_double), None if gc is None else _ptr(gc, ct.c_double))
        if rc != 0:
            raise ValueError(self.L["last_error"](self.c).decode())
        g = g.reshape(4, self.nx, self.nz)
        return (J.value, g, gc.reshape(8, self.Nx, self.Nz)) if want_gc else (J.value, g)

    def adjoint_source(self, y_vx=None, y_vz=None, y_p=None, ns=1):
        """F^T y: the adjoint of the source time functions, (ns, nt)."""
        ys = [None if y is None else np.ascontiguousarray(y, dtype=np.float64).reshape(-1) for y in (y_vx, y_vz, y_p)]
        out = np.zeros(ns * self.nt)
        rc = self.L["adjoint_source"](self.c, *[None if y is None else _ptr(y, ct.c_double) for y in ys],
                                      _ptr(out, ct.c_double))
        if rc != 0:
            raise ValueError(self.L["last_error"](self.c).decode())
        return out.reshape(ns, self.nt)

    def energy(self):
        E = np.empty(self.nt)
        Ek = np.empty(self.nt)
        Ep = np.empty(self.nt)
        rc = self.L["get_energy"](self.c, _ptr(E, ct.c_double), _ptr(Ek, ct.c_double), _ptr(Ep, ct.c_double))
        assert rc == 0
        return E, Ek, Ep


def set_threads(n, prec="f32"):
    return _lib(prec)["set_threads"](n)
