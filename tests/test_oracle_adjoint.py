"""Pins of the FP64 oracle adjoint and gradient (all three rotation forms).

The adjoint of the modelling operator is pinned by the dot-product test
<F s, y> = <s, F^T y>; the gradient of J = 1/2 |d - obs|^2 by central finite
differences of J with the absorbing layers frozen (the gradient's definition),
along random whole-model directions (edges included, so the transpose of the
edge padding is exercised) and by the h^2 convergence of the central difference.
"""
import numpy as np
import pytest

from fdmod_hip import synth
from common import run_case, small_case
from oracle import Oracle

NZ, NX, NT, DT = 40, 60, 300, 1e-3


def _model():
    return synth.model_small(nz=NZ, nx=NX, seed=3)


FORMS = ["bond", "tensor", "coord"]


def _mk(model, fs=0, stype=1, seis=2, mode=1, mt=1, form="bond"):
    o = Oracle(form, NZ, NX, NT, 10.0, 10.0, DT, ealThick=10, freesurface=fs, fpeak=15.0, mode=mode, prec="f64")
    o.set_model(model["vp"], model["vs"], model["rho"], model["eps"], model["delta"], model["theta"])
    o.set_source(stype, [30], [6], synth.ricker(15.0, DT, NT)[None])
    rx = np.arange(0, NX, 2)
    o.set_receivers(seis, mt, rx, np.full(len(rx), 3))
    return o


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs,stype,seis,mode,mt", [(0, 1, 2, 1, 1), (1, 1, 3, 1, 2), (0, 2, 3, 2, 1),
                                                   (1, 4, 1, 1, 1), (1, 7, 3, 1, 3), (0, 8, 2, 1, 1)])
def test_dot_product(fs, stype, seis, mode, mt, form):
    rng = np.random.default_rng(stype)
    o = _mk(_model(), fs, stype, seis, mode, mt, form)
    o.run()
    use = {1: seis in (1, 3), 3: seis in (1, 3), 4: seis in (2, 3)}
    d = {c: o.traces(c) for c in (1, 3, 4)}
    y = {c: rng.standard_normal(d[c].shape) for c in (1, 3, 4)}
    lhs = sum(float((d[c] * y[c]).sum()) for c in (1, 3, 4) if use[c])
    a = o.adjoint_source(*(y[c] if use[c] else None for c in (1, 3, 4)))
    src = synth.ricker(15.0, DT, NT)[None].astype(np.float32).astype(np.float64)
    rhs = float((src * a).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def _setup(fs, seis=2, form="bond"):
    m = _model()
    true = {k: v.copy() for k, v in m.items()}
    true["vp"] = (true["vp"] * 1.03).astype(np.float32)
    true["theta"] = (true["theta"] + 4).astype(np.float32)
    ot = _mk(true, fs, seis=seis, form=form)
    ot.run()
    obs = {c: ot.traces(c).astype(np.float32) for c in (1, 3, 4)}
    o = _mk(m, fs, seis=seis, form=form)
    kw = dict(obs_p=obs[4]) if seis == 2 else dict(obs_vx=obs[1], obs_vz=obs[3], obs_p=obs[4])
    J, g = o.gradient(**kw)
    comps = (4,) if seis == 2 else (1, 3, 4)

    def Jof(model):
        o.set_model_frozen(model["vp"], model["vs"], model["rho"], model["eps"], model["delta"], model["theta"])
        o.run()
        return sum(0.5 * float(((o.traces(c) - obs[c]) ** 2).sum()) for c in comps)

    return m, J, g, Jof


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs,seis", [(0, 2), (1, 3)])
def test_gradient_random_directions(fs, seis, form):
    m, J, g, Jof = _setup(fs, seis, form)
    assert abs(Jof(m) - J) <= 1e-12 * J
    rng = np.random.default_rng(10 + fs)
    for key, gi, sc in (("vp", 0, 0.5), ("eps", 1, 1e-4), ("delta", 2, 1e-4), ("theta", 3, 0.02)):
        dm = rng.standard_normal(m[key].shape)
        mp, mm = dict(m), dict(m)
        mp[key] = (m[key] + sc * dm).astype(np.float32)
        mm[key] = (m[key] - sc * dm).astype(np.float32)
        step = mp[key].astype(np.float64) - mm[key].astype(np.float64)
        fd = Jof(mp) - Jof(mm)
        an = float((g[gi] * step).sum())
        assert abs(fd - an) <= 1e-4 * abs(fd), (key, fd, an)


def test_gradient_central_difference_converges_as_h2():
    m, J, g, Jof = _setup(0)
    dm = np.zeros_like(m["vp"], dtype=np.float64)
    dm[20:40, 10:30] = 1.0
    errs = []
    for h in (20.0, 2.0):
        mp, mm = dict(m), dict(m)
        mp["vp"] = (m["vp"] + h * dm).astype(np.float32)
        mm["vp"] = (m["vp"] - h * dm).astype(np.float32)
        step = mp["vp"].astype(np.float64) - mm["vp"].astype(np.float64)
        fd = Jof(mp) - Jof(mm)
        errs.append(abs(fd - float((g[0] * step).sum())) / abs(fd))
    assert errs[1] < errs[0] / 50  # second-order: a 10x smaller step cuts the error ~100x


def test_gradient_vanishes_at_true_model():
    # observed data = modelled data up to their float32 rounding: J and the gradient
    # are at rounding level compared with those of a perturbed model
    _, J1, g1, _ = _setup(0)
    m = _model()
    o = _mk(m, 0)
    o.run()
    J, g = o.gradient(obs_p=o.traces(4).astype(np.float32))
    assert J < 1e-10 * J1
    assert np.abs(g).max() < 1e-5 * np.abs(g1).max()


@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_checkpointed_history_gives_the_same_gradient(form, monkeypatch):
    # a history budget below one step forces checkpoints every ceil(sqrt(nt)) steps with
    # each segment recomputed: the recomputed forward is the same arithmetic, so the
    # gradient is bit for bit the one from the stored history
    ctor, m, geom = small_case(nz=24, nx=30, nt=90, form=form, freesurface=1, ealThick=6, water=2)
    obs = {}
    t = run_case(Oracle, ctor, m, geom, prec="f64")
    t.set_model(m["vp"] * np.float32(1.02), m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    t.run()
    obs = [t.traces(c).astype(np.float32) for c in (1, 3, 4)]
    o = run_case(Oracle, ctor, m, geom, prec="f64")
    monkeypatch.delenv("ORACLE_HIST_BYTES", raising=False)
    J1, g1 = o.gradient(*obs)
    monkeypatch.setenv("ORACLE_HIST_BYTES", "1")
    J2, g2 = o.gradient(*obs)
    assert J1 == J2 and np.array_equal(g1, g2)
    assert np.abs(g1).max() > 0
