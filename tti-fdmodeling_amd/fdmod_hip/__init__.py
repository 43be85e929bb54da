"""Python front end of libfdmod_hip: one modelling context per object.

Mirrors the per-shot flow of the reference programs
(src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_*.c): create the padded grid,
set the model, the source and the receivers, run the time loop, read traces.
All work goes through the C-ABI in include/fdmod_hip.h.
"""
import ctypes as ct

import numpy as np

from ._lib import AID, FORMS, HISTORY, STORAGE, FdmodError, GradOpts, Params, SNAPSHOT_FN, _f32, _ptr, load

__all__ = ["Modeling", "Batch", "FdmodError", "FORMS", "load"]


def _history_mode(history):
    """fdmod_grad_opts.history from the Python spelling."""
    if isinstance(history, str):
        return HISTORY[history]
    return int(bool(history))


class Modeling:
    """A single-shot modelling context on the GPU.

    Arrays follow the reference layout: model inputs are (nx, nz) float32 with iz
    fastest; padded arrays returned by get() are (Nx, Nz); traces are
    (ntrace, ntsample) like Receiver2D.recv.
    """

    def __init__(self, form, nz, nx, nt, dx, dz, dt, ealThick=30, eal_alpha=-1.0, freesurface=0,
                 fpeak=10.0, scalar=1.0, mode=1):
        self.L = load()
        p = Params(FORMS[form] if isinstance(form, str) else int(form), nz, nx, nt, dx, dz, dt, ealThick,
                   eal_alpha, freesurface, fpeak, scalar, mode)
        h = ct.c_void_p()
        rc = self.L.fdmod_create(ct.byref(p), ct.byref(h))
        if rc != 0:
            raise FdmodError(rc, self.L.fdmod_strerror(rc).decode())
        self.c = h
        self.form, self.nz, self.nx, self.nt = form, nz, nx, nt
        d = (ct.c_int * 7)()
        self.L.fdmod_dims(self.c, d)
        self.Nz, self.Nx, self.e, self.izStart, self.izEnd, self.ixStart, self.ixEnd = list(d)
        self.nr = 0
        self.ntsample = 0
        self.ns = 0
        self._snap_cb = None

    def close(self):
        if getattr(self, "c", None):
            self.L.fdmod_destroy(self.c)
            self.c = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise FdmodError(rc, self.L.fdmod_last_error(self.c).decode())

    def set_model(self, vp, vs, rho, eps, delta, theta):
        arrs = [_f32(a).reshape(-1) for a in (vp, vs, rho, eps, delta, theta)]
        for a in arrs:
            if a.size != self.nz * self.nx:
                raise ValueError("model arrays must have nz*nx samples")
        dtm = ct.c_double(0)
        rc = self.L.fdmod_set_model(self.c, *[_ptr(a, ct.c_float) for a in arrs], ct.byref(dtm))
        self.dt_max = dtm.value
        self._check(rc)

    def set_source(self, stype, posx, posz, src):
        posx = np.ascontiguousarray(posx, dtype=np.int32)
        posz = np.ascontiguousarray(posz, dtype=np.int32)
        src = _f32(src).reshape(len(posx), self.nt)
        self._check(self.L.fdmod_set_source(self.c, stype, len(posx), _ptr(posx, ct.c_int), _ptr(posz, ct.c_int),
                                            _ptr(src, ct.c_float)))
        self.ns = len(posx)

    def set_receivers(self, seistype, mt, posx, posz):
        posx = np.ascontiguousarray(posx, dtype=np.int32)
        posz = np.ascontiguousarray(posz, dtype=np.int32)
        self._check(self.L.fdmod_set_receivers(self.c, seistype, mt, len(posx), _ptr(posx, ct.c_int),
                                               _ptr(posz, ct.c_int)))
        self.nr = len(posx)
        self.ntsample = (self.nt - 1) // mt + 1

    def set_snapshots(self, wsave, savew, snap_sample, sink):
        """sink(it, kind, array[n2, n1]) is called at every snapshot time."""

        def cb(user, it, kind, data, n1, n2):
            arr = np.ctypeslib.as_array(data, shape=(n2 * n1,)).reshape(n2, n1).copy()
            try:
                sink(it, kind.decode(), arr)
                return 0
            except Exception:  # noqa: BLE001 - reported to the library as a failed sink
                return 1

        self._snap_cb = SNAPSHOT_FN(cb)
        self._check(self.L.fdmod_set_snapshots(self.c, wsave, savew, snap_sample, self._snap_cb, None))

    def set_energy(self, on=1):
        self._check(self.L.fdmod_set_energy(self.c, on))

    def run(self, k0=0, count=None):
        """Run loop iterations k0 .. k0+count-1; returns device seconds."""
        if count is None:
            count = self.nt - k0
        sec = ct.c_double(0)
        self._check(self.L.fdmod_run(self.c, k0, count, ct.byref(sec)))
        return sec.value

    def reset(self):
        self._check(self.L.fdmod_reset(self.c))

    def run_energy_noalloc(self, k0=0, count=None):
        """Run loop iterations k0 .. k0+count-1 through the allocation-free entry point;
        returns the energy series (E, Ek, Ep) of those iterations and the device seconds."""
        if count is None:
            count = self.nt - k0
        n = max(0, min(k0 + count, self.nt) - k0)
        E, Ek, Ep = (np.zeros(max(n, 1), dtype=np.float32) for _ in range(3))
        sec = ct.c_double(0)
        self._check(self.L.fdmod_run_energy_noalloc(self.c, k0, count, _ptr(E, ct.c_float), _ptr(Ek, ct.c_float),
                                                    _ptr(Ep, ct.c_float), ct.byref(sec)))
        return E[:n], Ek[:n], Ep[:n], sec.value

    def traces(self, comp):
        """comp: 1 = vx, 3 = vz, 4 = p (traces2d_output flags)."""
        out = np.empty(self.nr * self.ntsample, dtype=np.float32)
        self._check(self.L.fdmod_get_traces(self.c, comp, _ptr(out, ct.c_float)))
        return out.reshape(self.nr, self.ntsample)

    def energy(self):
        E, Ek, Ep = (np.empty(self.nt, dtype=np.float32) for _ in range(3))
        self._check(self.L.fdmod_get_energy(self.c, _ptr(E, ct.c_float), _ptr(Ek, ct.c_float), _ptr(Ep, ct.c_float)))
        return E, Ek, Ep

    def get(self, name):
        out = np.empty(self.Nx * self.Nz, dtype=np.float32)
        self._check(self.L.fdmod_get_array(self.c, AID[name], _ptr(out, ct.c_float)))
        return out.reshape(self.Nx, self.Nz)

    def set(self, name, arr):
        a = _f32(arr).reshape(-1)
        if a.size != self.Nx * self.Nz:
            raise ValueError("padded arrays must have Nx*Nz samples")
        self._check(self.L.fdmod_set_array(self.c, AID[name], _ptr(a, ct.c_float)))

    def update_model(self, vp, vs, rho, eps, delta, theta):
        """Replace the model keeping the absorbing layers of the last set_model."""
        arrs = [_f32(a).reshape(-1) for a in (vp, vs, rho, eps, delta, theta)]
        for a in arrs:
            if a.size != self.nz * self.nx:
                raise ValueError("model arrays must have nz*nx samples")
        self._check(self.L.fdmod_update_model(self.c, *[_ptr(a, ct.c_float) for a in arrs]))

    def _series(self, arrs):
        out = []
        for a in arrs:
            if a is None:
                out.append(None)
                continue
            a = _f32(a).reshape(-1)
            if a.size != self.nr * self.ntsample:
                raise ValueError("data series must have ntrace*ntsample samples")
            out.append(a)
        return out, [None if a is None else _ptr(a, ct.c_float) for a in out]

    def gradient(self, obs_vx=None, obs_vz=None, obs_p=None, ncheckpoints=0, storage="f32", history=False):
        """J = 1/2 |d - obs|^2 over the given components and dJ/d(vp, eps, delta, theta[deg])
        as a float64 array (4, nx, nz); runs the forward from rest.
        storage "f32", "f16" or "f16all" (stored wavefields).  history: False or
        "checkpoints" (ncheckpoints states, segments recomputed), True or "full" (the
        strain of every step), "reconstruct" (boundary-saving reconstruction: the loop
        run backwards over the undamped interior from saved band states; ncheckpoints > 0
        adds that many restart states), "reconstruct_interior" (the same without the
        absorbing layers' imaging, which only the model's edge cells receive)."""
        keep, ptrs = self._series((obs_vx, obs_vz, obs_p))
        J = ct.c_double(0)
        g = np.empty(4 * self.nx * self.nz, dtype=np.float64)
        opts = GradOpts(int(ncheckpoints), STORAGE[storage], _history_mode(history))
        self._check(self.L.fdmod_gradient_opts(self.c, *ptrs, ct.byref(opts), ct.byref(J), _ptr(g, ct.c_double)))
        return J.value, g.reshape(4, self.nx, self.nz)

    def adjoint_source(self, y_vx=None, y_vz=None, y_p=None):
        """F^T y: the adjoint of the source time functions, float64 (ns, nt)."""
        keep, ptrs = self._series((y_vx, y_vz, y_p))
        ns = self.ns
        out = np.empty(ns * self.nt, dtype=np.float64)
        self._check(self.L.fdmod_adjoint_source(self.c, *ptrs, _ptr(out, ct.c_double)))
        return out.reshape(ns, self.nt)

    def adjoint_dot(self, y_vx=None, y_vz=None, y_p=None):
        """(<F s, y>, <s, F^T y>) reduced in FP64 on the device."""
        keep, ptrs = self._series((y_vx, y_vz, y_p))
        lhs, rhs = ct.c_double(0), ct.c_double(0)
        self._check(self.L.fdmod_adjoint_dot(self.c, *ptrs, ct.byref(lhs), ct.byref(rhs)))
        return lhs.value, rhs.value

    def time_sweeps(self, reps=20):
        """Mean device ms of one velocity-sweep and one stress-sweep launch, the two
        alternating as in the time loop."""
        v, s = ct.c_double(0), ct.c_double(0)
        self._check(self.L.fdmod_time_sweeps(self.c, reps, ct.byref(v), ct.byref(s)))
        return v.value, s.value


class Batch:
    """Many independent shots advanced together (one launch per sweep over all shots,
    time loop captured in a hipGraph).  Shot s has its own model width nx[s]; all other
    parameters are shared.  Per-shot results equal those of a single-shot Modeling."""

    def __init__(self, form, nz, nxs, nt, dx, dz, dt, ealThick=30, eal_alpha=-1.0, freesurface=0, fpeak=10.0,
                 scalar=1.0, mode=1):
        self.L = load()
        p = Params(FORMS[form] if isinstance(form, str) else int(form), nz, int(nxs[0]), nt, dx, dz, dt, ealThick,
                   eal_alpha, freesurface, fpeak, scalar, mode)
        nx = np.ascontiguousarray(nxs, dtype=np.int32)
        h = ct.c_void_p()
        rc = self.L.fdmod_batch_create(ct.byref(p), len(nx), _ptr(nx, ct.c_int), ct.byref(h))
        if rc != 0:
            raise FdmodError(rc, self.L.fdmod_strerror(rc).decode())
        self.b = h
        self.nz, self.nxs, self.nt = nz, list(nx), nt
        self.nr = [0] * len(nx)
        self.ntsample = [0] * len(nx)

    def close(self):
        if getattr(self, "b", None):
            self.L.fdmod_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise FdmodError(rc, self.L.fdmod_batch_last_error(self.b).decode())

    def set_model(self, shot, vp, vs, rho, eps, delta, theta):
        arrs = [_f32(a).reshape(-1) for a in (vp, vs, rho, eps, delta, theta)]
        for a in arrs:
            if a.size != self.nz * self.nxs[shot]:
                raise ValueError("model arrays must have nz*nx[shot] samples")
        dtm = ct.c_double(0)
        self._check(self.L.fdmod_batch_set_model(self.b, shot, *[_ptr(a, ct.c_float) for a in arrs], ct.byref(dtm)))

    def set_source(self, shot, stype, posx, posz, src):
        posx = np.ascontiguousarray(posx, dtype=np.int32)
        posz = np.ascontiguousarray(posz, dtype=np.int32)
        src = _f32(src).reshape(len(posx), self.nt)
        self._check(self.L.fdmod_batch_set_source(self.b, shot, stype, len(posx), _ptr(posx, ct.c_int),
                                                  _ptr(posz, ct.c_int), _ptr(src, ct.c_float)))

    def set_receivers(self, shot, seistype, mt, posx, posz):
        posx = np.ascontiguousarray(posx, dtype=np.int32)
        posz = np.ascontiguousarray(posz, dtype=np.int32)
        self._check(self.L.fdmod_batch_set_receivers(self.b, shot, seistype, mt, len(posx), _ptr(posx, ct.c_int),
                                                     _ptr(posz, ct.c_int)))
        self.nr[shot] = len(posx)
        self.ntsample[shot] = (self.nt - 1) // mt + 1

    def tune(self, reps=10):
        v, s = ct.c_int(0), ct.c_int(0)
        self._check(self.L.fdmod_batch_tune(self.b, reps, ct.byref(v), ct.byref(s)))
        return v.value, s.value

    def set_tiles(self, segLV, segLS):
        self._check(self.L.fdmod_batch_set_tiles(self.b, segLV, segLS))

    def run(self, k0=0, count=None):
        if count is None:
            count = self.nt - k0
        sec = ct.c_double(0)
        self._check(self.L.fdmod_batch_run(self.b, k0, count, ct.byref(sec)))
        return sec.value

    def reset(self):
        self._check(self.L.fdmod_batch_reset(self.b))

    def traces(self, shot, comp):
        out = np.empty(self.nr[shot] * self.ntsample[shot], dtype=np.float32)
        self._check(self.L.fdmod_batch_get_traces(self.b, shot, comp, _ptr(out, ct.c_float)))
        return out.reshape(self.nr[shot], self.ntsample[shot])

    def gradient(self, obs_vx=None, obs_vz=None, obs_p=None, ncheckpoints=0, storage="f32", history=False,
                 x0=None, nx_full=0):
        """Misfit gradients of all shots.  obs_* are lists of per-shot
        (ntrace, ntsample) arrays (None = component absent).  With x0 (per-shot first
        model column) the result is the FP64 sum over shots on the nx_full-column model,
        shape (4, nx_full, nz); otherwise a list of per-shot (4, nx[s], nz) arrays.
        Returns (total misfit, gradient)."""
        n = len(self.nxs)
        keep = []

        def ptrs(series):
            if series is None:
                return None
            arr = (ct.POINTER(ct.c_float) * n)()
            for s, a in enumerate(series):
                a = _f32(a).reshape(-1)
                if a.size != self.nr[s] * self.ntsample[s]:
                    raise ValueError("shot %d: data series must have ntrace*ntsample samples" % s)
                keep.append(a)
                arr[s] = _ptr(a, ct.c_float)
            return arr

        pv, pz, pp = ptrs(obs_vx), ptrs(obs_vz), ptrs(obs_p)
        opts = GradOpts(int(ncheckpoints), STORAGE[storage], _history_mode(history))
        J = ct.c_double(0)
        if x0 is not None:
            x0a = np.ascontiguousarray(x0, dtype=np.int32)
            out = np.empty(4 * int(nx_full) * self.nz, dtype=np.float64)
            self._check(self.L.fdmod_batch_gradient(self.b, pv, pz, pp, ct.byref(opts), _ptr(x0a, ct.c_int),
                                                    int(nx_full), ct.byref(J), _ptr(out, ct.c_double)))
            return J.value, out.reshape(4, int(nx_full), self.nz)
        tot = sum(4 * int(w) * self.nz for w in self.nxs)
        out = np.empty(tot, dtype=np.float64)
        self._check(self.L.fdmod_batch_gradient(self.b, pv, pz, pp, ct.byref(opts), None, 0, ct.byref(J),
                                                _ptr(out, ct.c_double)))
        res, o = [], 0
        for w in self.nxs:
            k = 4 * int(w) * self.nz
            res.append(out[o:o + k].reshape(4, int(w), self.nz))
            o += k
        return J.value, res
