"""Batched shots (one launch per sweep over all shots, hipGraph-captured loop) give
per-shot traces bit-identical to the single-shot path and to the oracle."""
import numpy as np
import pytest

from fdmod_hip import Batch, Modeling, synth
from oracle import Oracle

pytestmark = pytest.mark.gpu

NZ, NT, DT = 40, 150, 1e-3


def _shots(multi1=False):
    m = synth.model_small(nz=NZ, nx=200, seed=7)
    m["theta"] = np.where(np.arange(NZ)[None, :] > 20, 25.0, -10.0).astype(np.float32) + 0 * m["theta"]
    crops = [(10, 50), (40, 73), (90, 61), (100, 88)]  # (x0, width): uneven apertures
    stypes = [1, 2, 8, 1]
    shots = []
    for s, (x0, w) in enumerate(crops):
        mc = synth.crop(m, x0, w)
        # multi1: the last shot has two stype-1 sources (not added by the stress sweep)
        ns = 2 if stypes[s] == 8 or (multi1 and s == 3) else 1
        sx = [w // 2 + 2 * i for i in range(ns)]
        sz = [3 + i for i in range(ns)]
        wav = np.stack([synth.ricker(15.0, DT, NT) * (1 + 0.3 * i) for i in range(ns)])
        shots.append(dict(m=mc, w=w, stype=stypes[s], sx=sx, sz=sz, src=wav, rx=np.arange(0, w, 3),
                          rz=np.full(len(range(0, w, 3)), 2)))
    return shots


def _single(cls, form, sh, seis=3, **kw):
    o = cls(form, NZ, sh["w"], NT, 10.0, 10.0, DT, ealThick=8, freesurface=1, fpeak=15.0, **kw)
    m = sh["m"]
    o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    o.set_source(sh["stype"], sh["sx"], sh["sz"], sh["src"])
    o.set_receivers(seis, 2, sh["rx"], sh["rz"])
    return o


def _batch(form, shots, seis=3):
    b = Batch(form, NZ, [sh["w"] for sh in shots], NT, 10.0, 10.0, DT, ealThick=8, freesurface=1, fpeak=15.0)
    for s, sh in enumerate(shots):
        m = sh["m"]
        b.set_model(s, m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
        b.set_source(s, sh["stype"], sh["sx"], sh["sz"], sh["src"])
        b.set_receivers(s, seis, 2, sh["rx"], sh["rz"])
    return b


# shot groups (run one after another inside each captured chunk): all shots in one
# group, every shot its own group, and two groups of two (grids of 9728..14592 points)
@pytest.mark.parametrize("group_pts", [None, "1", "30000"])
@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_batch_equals_single_shots(form, group_pts, monkeypatch):
    shots = _shots()
    print("batch", form, group_pts, flush=True)
    if group_pts:
        monkeypatch.setenv("FDMOD_BATCH_PTS", group_pts)
    else:
        monkeypatch.delenv("FDMOD_BATCH_PTS", raising=False)
    b = _batch(form, shots)
    b.run(0, 70)   # two captured chunks + eager remainder, then a second call
    b.run(70, 80)
    for s, sh in enumerate(shots):
        g = _single(Modeling, form, sh)
        g.run()
        o = _single(Oracle, form, sh)
        o.run()
        for comp in (1, 3, 4):
            tb = b.traces(s, comp)
            assert np.array_equal(tb, g.traces(comp)), (s, comp)
            assert np.array_equal(tb, o.traces(comp).astype(np.float32)), (s, comp)


# pressure-only receivers: the Bond and coordinate batched stress sweeps record them
# (every second iteration, mt = 2) together with the fused stype-1 sources
@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_batch_pressure_receivers_equal_single_shots(form):
    shots = _shots()
    print("batch p receivers", form, flush=True)
    b = _batch(form, shots, seis=2)
    b.run(0, 70)
    b.run(70, 80)
    for s, sh in enumerate(shots):
        g = _single(Modeling, form, sh, seis=2)
        g.run()
        o = _single(Oracle, form, sh, seis=2)
        o.run()
        tb = b.traces(s, 4)
        assert np.array_equal(tb, g.traces(4)), s
        assert np.array_equal(tb, o.traces(4).astype(np.float32)), s


# a shot with two stype-1 sources next to shots whose sweeps add theirs: its pressure
# receivers inside the footprints sample after the separate injection
@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_batch_pressure_receivers_after_unfused_stress_sources(form):
    shots = _shots(multi1=True)
    print("batch p receivers, two stype-1 sources in shot 3", form, flush=True)
    b = _batch(form, shots, seis=2)
    b.run(0, 70)
    b.run(70, 80)
    for s, sh in enumerate(shots):
        o = _single(Oracle, form, sh, seis=2)
        o.run()
        assert np.array_equal(b.traces(s, 4), o.traces(4).astype(np.float32)), s


def test_autotuner_and_tile_shapes_keep_results():
    shots = _shots()
    b = _batch("bond", shots)
    v, s = b.tune(3)
    assert v in (0, 22, 33, 44, 55, 66, 88) and s in (4, 5, 8, 11, 22)
    b.run()
    ref = [b.traces(i, 4).copy() for i in range(len(shots))]
    # 0: per-group velocity segments (the default)
    for lv, ls in ((22, 11), (88, 44), (7, 5), (0, 11)):
        b.reset()
        b.set_tiles(lv, ls)
        b.run()
        for i in range(len(shots)):
            assert np.array_equal(b.traces(i, 4), ref[i]), (lv, ls, i)


@pytest.mark.parametrize("storage,history", [("f32", True), ("f16", False), ("f32", "reconstruct"),
                                             ("f32", "reconstruct_interior")])
def test_batch_gradient_equals_single_shot_gradients(storage, history):
    shots = _shots()
    b = _batch("bond", shots)
    ns = (NT - 1) // 2 + 1
    rng = np.random.default_rng(11)
    obs_p = [rng.standard_normal((len(sh["rx"]), ns)).astype(np.float32) * 1e-6 for sh in shots]
    J, per = b.gradient(obs_p=obs_p, storage=storage, history=history)
    x0 = [10, 40, 90, 100]
    J2, full = b.gradient(obs_p=obs_p, storage=storage, history=history, x0=x0, nx_full=200)
    ref = np.zeros_like(full)
    Jsum = 0.0
    for s, sh in enumerate(shots):
        g = _single(Modeling, "bond", sh)
        Js, gs = g.gradient(obs_p=obs_p[s], storage=storage, history=history)
        assert np.array_equal(per[s], gs), s
        ref[:, x0[s]:x0[s] + sh["w"]] += gs
        Jsum += Js
    assert J == Jsum and J2 == Jsum
    assert np.array_equal(full, ref)


def test_batch_gradient_groups_of_one(monkeypatch):
    # one shot at a time: the pooled gradient memory is lent to shots of different
    # widths in turn and must grow to each
    monkeypatch.setenv("FDMOD_GRAD_GROUP", "1")
    shots = _shots()
    b = _batch("bond", shots)
    ns = (NT - 1) // 2 + 1
    rng = np.random.default_rng(12)
    obs_p = [rng.standard_normal((len(sh["rx"]), ns)).astype(np.float32) * 1e-6 for sh in shots]
    for _ in range(2):
        J, per = b.gradient(obs_p=obs_p, storage="f16", history=True)
        for s, sh in enumerate(shots):
            g = _single(Modeling, "bond", sh)
            Js, gs = g.gradient(obs_p=obs_p[s], storage="f16", history=True)
            assert np.array_equal(per[s], gs), s
