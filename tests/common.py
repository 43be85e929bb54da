"""Shared helpers for the tests: workload builders for the oracle and the GPU path."""
import numpy as np

from fdmod_hip import synth


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def small_case(nz=60, nx=80, nt=300, form="bond", freesurface=0, ealThick=10, seed=1, theta=None,
               stype=1, seistype=3, mt=1, mode=1, water=0, sz=None, sx=None, rz=2, dx=10.0, dz=10.0, dt=1e-3,
               fpeak=15.0, nsrc=1):
    """A seeded small modelling case as a dict of constructor args + inputs."""
    m = synth.model_small(nz=nz, nx=nx, dx=dx, seed=seed)
    if theta is not None:
        m["theta"] = np.full((nx, nz), theta, dtype=np.float32)
    if water > 0:
        m["vp"][:, :water] = 1500.0
        m["vs"][:, :water] = 0.0
        m["rho"][:, :water] = 1000.0
        m["eps"][:, :water] = 0.0
        m["delta"][:, :water] = 0.0
    src = synth.ricker(fpeak, dt, nt)
    sx = nx // 2 if sx is None else sx
    sz = 2 if sz is None else sz
    if nsrc == 1:
        spx, spz, sw = [sx], [sz], src[None]
    else:
        spx = [sx + 3 * i for i in range(nsrc)]
        spz = [sz + (i % 3) for i in range(nsrc)]
        sw = np.stack([src * (1.0 + 0.1 * i) for i in range(nsrc)])
    ctor = dict(form=form, nz=nz, nx=nx, nt=nt, dx=dx, dz=dz, dt=dt, ealThick=ealThick, eal_alpha=-1.0,
                freesurface=freesurface, fpeak=fpeak, scalar=1.0, mode=mode)
    geom = dict(stype=stype, spx=spx, spz=spz, src=sw, seistype=seistype, mt=mt,
                rpx=np.arange(0, nx + 1, 1), rpz=np.full(nx + 1, rz))
    return ctor, m, geom


def run_case(cls, ctor, m, geom, **kw):
    o = cls(**ctor, **kw)
    o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    o.set_source(geom["stype"], geom["spx"], geom["spz"], geom["src"])
    o.set_receivers(geom["seistype"], geom["mt"], geom["rpx"], geom["rpz"])
    return o
