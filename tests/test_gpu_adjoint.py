"""GPU adjoint and misfit gradient (all three rotation forms) against the FP64 oracle adjoint,
the device dot-product test, finite differences of the GPU misfit along 16 random
directions, checkpoint-count invariance and the torch.autograd.Function wrapper.

Tolerances: the GPU runs FP32 stencils (FP64 reductions and imaging sums), the
oracle runs FP64 throughout.  Adjoint sources and gradients agree to a relative L2
of 1e-4; the dot product closes to 5e-5 (FP32 propagation over 300 steps, FP64 sums;
1e-5 for most cases, 1.0e-5 for the coordinate form with the stype-8 source).
"""
import numpy as np
import pytest

from common import rel_l2
from fdmod_hip import Modeling, synth
from oracle import Oracle

pytestmark = pytest.mark.gpu

NZ, NX, NT, DT = 40, 60, 300, 1e-3


def _model(seed=3):
    return synth.model_small(nz=NZ, nx=NX, seed=seed)


FORMS = ["bond", "tensor", "coord"]


def _mk(cls, model, fs=0, stype=1, seis=2, mode=1, mt=1, form="bond", dx=10.0, dz=10.0, **kw):
    o = cls(form, NZ, NX, NT, dx, dz, DT, ealThick=10, freesurface=fs, fpeak=15.0, mode=mode, **kw)
    o.set_model(model["vp"], model["vs"], model["rho"], model["eps"], model["delta"], model["theta"])
    o.set_source(stype, [30], [6], synth.ricker(15.0, DT, NT)[None])
    rx = np.arange(0, NX, 2)
    o.set_receivers(seis, mt, rx, np.full(len(rx), 3))
    return o


def _use(seis):
    return {1: seis in (1, 3), 3: seis in (1, 3), 4: seis in (2, 3)}


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs,stype,seis,mode,mt", [(0, 1, 2, 1, 1), (1, 1, 3, 1, 2), (0, 2, 3, 2, 1),
                                                   (1, 4, 1, 1, 1), (1, 7, 3, 1, 3), (0, 8, 2, 1, 1)])
def test_dot_product_and_adjoint_source(fs, stype, seis, mode, mt, form):
    print("dot", form, fs, stype, seis, mode, mt, flush=True)
    rng = np.random.default_rng(stype)
    g = _mk(Modeling, _model(), fs, stype, seis, mode, mt, form)
    use = _use(seis)
    shape = (g.nr, g.ntsample)
    y = {c: rng.standard_normal(shape).astype(np.float32) for c in (1, 3, 4)}
    ys = [y[c] if use[c] else None for c in (1, 3, 4)]
    lhs, rhs = g.adjoint_dot(*ys)
    assert abs(lhs - rhs) <= 5e-5 * abs(lhs), (lhs, rhs)
    a_gpu = g.adjoint_source(*ys)
    o = _mk(Oracle, _model(), fs, stype, seis, mode, mt, form, prec="f64")
    a_ref = o.adjoint_source(*[None if v is None else v.astype(np.float64) for v in ys])
    assert rel_l2(a_gpu, a_ref) <= 1e-4


# three stype-1 sources with overlapping footprints (injected after the stress sweep) and
# pressure receivers inside them: the adjoint of each source time function against the
# FP64 oracle, and the dot product
@pytest.mark.parametrize("form", FORMS)
def test_adjoint_source_of_overlapping_stress_sources(form):
    print("dot, three stype-1 sources", form, flush=True)
    m = _model()
    src = synth.ricker(15.0, DT, NT)
    sw = np.stack([src * (1.0 + 0.2 * i) for i in range(3)])
    spx, spz = [26, 29, 32], [4, 5, 6]

    def mk(cls, **kw):
        o = cls(form, NZ, NX, NT, 10.0, 10.0, DT, ealThick=10, freesurface=0, fpeak=15.0, **kw)
        o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
        o.set_source(1, spx, spz, sw)
        rx = np.arange(0, NX, 2)
        o.set_receivers(2, 1, rx, np.full(len(rx), 3))
        return o

    g = mk(Modeling)
    rng = np.random.default_rng(21)
    y = rng.standard_normal((g.nr, g.ntsample)).astype(np.float32)
    lhs, rhs = g.adjoint_dot(None, None, y)
    assert abs(lhs - rhs) <= 5e-5 * abs(lhs), (lhs, rhs)
    a_gpu = g.adjoint_source(None, None, y)
    o = mk(Oracle, prec="f64")
    a_ref = o.adjoint_source(None, None, y.astype(np.float64), ns=3)
    assert a_gpu.shape == (3, NT)
    assert rel_l2(a_gpu, a_ref) <= 1e-4


def _obs(fs, seis, mode=1, mt=1, form="bond", **kw):
    m = _model()
    true = {k: v.copy() for k, v in m.items()}
    true["vp"] = (true["vp"] * 1.03).astype(np.float32)
    true["theta"] = (true["theta"] + 4).astype(np.float32)
    ot = _mk(Oracle, true, fs, seis=seis, mode=mode, mt=mt, form=form, prec="f64", **kw)
    ot.run()
    use = _use(seis)
    return m, [ot.traces(c).astype(np.float32) if use[c] else None for c in (1, 3, 4)]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs,seis,mode,mt", [(0, 2, 1, 1), (1, 3, 1, 2), (1, 1, 2, 1)])
def test_gradient_matches_fp64_oracle(fs, seis, mode, mt, form):
    print("gradient", form, fs, seis, mode, mt, flush=True)
    m, obs = _obs(fs, seis, mode, mt, form)
    g = _mk(Modeling, m, fs, seis=seis, mode=mode, mt=mt, form=form)
    J, grad = g.gradient(*obs)
    o = _mk(Oracle, m, fs, seis=seis, mode=mode, mt=mt, form=form, prec="f64")
    Jr, gr = o.gradient(*obs)
    assert abs(J - Jr) <= 1e-5 * Jr
    for i in range(4):
        assert rel_l2(grad[i], gr[i]) <= 1e-4, i
    # the traces left in the context are those of the run from rest: a fresh context's
    comp = 4 if seis != 1 else 1
    fresh = _mk(Modeling, m, fs, seis=seis, mode=mode, mt=mt, form=form)
    fresh.run()
    assert np.array_equal(g.traces(comp), fresh.traces(comp))


# dx != dz (8 m x 12.5 m cells): the adjoint of every invdx / invdz placement, of the
# coordinate form's invdx on d(sxx)/dz (CR:633) and of the source weights
@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs,stype,seis", [(0, 1, 3), (1, 4, 3)])
def test_adjoint_and_gradient_dx_ne_dz(form, fs, stype, seis):
    print("dx != dz", form, fs, stype, seis, flush=True)
    sp = dict(dx=8.0, dz=12.5)
    rng = np.random.default_rng(21)
    g = _mk(Modeling, _model(), fs, stype, seis, form=form, **sp)
    use = _use(seis)
    ys = [rng.standard_normal((g.nr, g.ntsample)).astype(np.float32) if use[c] else None for c in (1, 3, 4)]
    lhs, rhs = g.adjoint_dot(*ys)
    assert abs(lhs - rhs) <= 5e-5 * abs(lhs), (lhs, rhs)
    o = _mk(Oracle, _model(), fs, stype, seis, form=form, prec="f64", **sp)
    a_ref = o.adjoint_source(*[None if v is None else v.astype(np.float64) for v in ys])
    assert rel_l2(g.adjoint_source(*ys), a_ref) <= 1e-4
    m, obs = _obs(fs, seis, form=form, **sp)
    g = _mk(Modeling, m, fs, seis=seis, form=form, **sp)
    J, grad = g.gradient(*obs)
    o = _mk(Oracle, m, fs, seis=seis, form=form, prec="f64", **sp)
    Jr, gr = o.gradient(*obs)
    assert abs(J - Jr) <= 1e-5 * Jr
    errs = [rel_l2(grad[i], gr[i]) for i in range(4)]
    print("  gradient vs FP64", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= 1e-4, errs


@pytest.mark.parametrize("form", FORMS)
def test_gradient_repeat_runs_bitwise(form):
    # every gradient mode is deterministic: the FP64 imaging sums and reductions add in a
    # fixed order, so a second call (new context or the same one) gives the same bits
    m, obs = _obs(1, 2, form=form)
    g = _mk(Modeling, m, 1, seis=2, form=form)
    h = _mk(Modeling, m, 1, seis=2, form=form)
    for kw in (dict(), dict(storage="f16", history=True), dict(storage="f16all"), dict(history="reconstruct")):
        J1, g1 = g.gradient(*obs, **kw)
        J2, g2 = g.gradient(*obs, **kw)
        J3, g3 = h.gradient(*obs, **kw)
        assert J1 == J2 == J3, kw
        assert np.array_equal(g1, g2) and np.array_equal(g1, g3), kw


def test_gradient_independent_of_checkpoint_count():
    # the recomputed forward is bit-identical; only the grouping of the FP64 imaging
    # additions (blocks of steps close at segment ends) changes with the count
    m, obs = _obs(1, 2)
    g = _mk(Modeling, m, 1, seis=2)
    ref = g.gradient(*obs, ncheckpoints=0)
    for ncp in (1, 7, NT):
        J, grad = g.gradient(*obs, ncheckpoints=ncp)
        assert J == ref[0], ncp
        for i in range(4):
            assert rel_l2(grad[i], ref[1][i]) <= 1e-12, (ncp, i)


@pytest.mark.parametrize("form", FORMS)
def test_gradient_storage_modes(form):
    # FP32 full history = checkpointed recomputation up to the grouping of FP64 sums;
    # FP16 storage (per-block scales) within 1e-3 of the FP64 oracle
    m, obs = _obs(1, 2, form=form)
    g = _mk(Modeling, m, 1, seis=2, form=form)
    ref = g.gradient(*obs)
    J, grad = g.gradient(*obs, storage="f32", history=True)
    assert J == ref[0] and all(rel_l2(grad[i], ref[1][i]) <= 1e-12 for i in range(4))
    o = _mk(Oracle, m, 1, seis=2, form=form, prec="f64")
    Jr, gr = o.gradient(*obs)
    for storage, hist in (("f16", True), ("f16", False), ("f16all", False)):
        J, grad = g.gradient(*obs, storage=storage, history=hist)
        assert J == ref[0]  # the forward itself is FP32 in every mode
        errs = [rel_l2(grad[i], gr[i]) for i in range(4)]
        print(form, storage, "history" if hist else "checkpoints", ["%.2e" % e for e in errs], flush=True)
        # FP16 checkpoints restart each recomputed segment from rounded states (normal
        # stresses stored as sum and difference): here at most 7.8e-4 (tensor form, theta),
        # the other modes below 4e-4
        assert max(errs) <= 1e-3, errs


# Pressure receivers on one row at consecutive columns (the gradient workloads' geometry):
# the two adjoint sweeps and the imaging add the pressure residuals themselves, with no
# injection launch.  The same receivers listed right to left take the separate injection
# kernel; both gradients must agree to rounding, and with the FP64 oracle.
@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("storage,hist", [("f32", False), ("f16", True)])
def test_fused_line_injection_matches_separate_injection(form, storage, hist):
    print("fused injection", form, storage, hist, flush=True)
    m = _model()
    true = {k: v.copy() for k, v in m.items()}
    true["vp"] = (true["vp"] * 1.03).astype(np.float32)
    true["theta"] = (true["theta"] + 4).astype(np.float32)
    rx, rz = np.arange(NX), np.full(NX, 3)

    def mk(cls, model, xs, **kw):
        o = cls(form, NZ, NX, NT, 10.0, 10.0, DT, ealThick=10, freesurface=0, fpeak=15.0, **kw)
        o.set_model(model["vp"], model["vs"], model["rho"], model["eps"], model["delta"], model["theta"])
        o.set_source(1, [30], [6], synth.ricker(15.0, DT, NT)[None])
        o.set_receivers(2, 1, xs, rz)
        return o

    ot = mk(Oracle, true, rx, prec="f64")
    ot.run()
    obs = ot.traces(4).astype(np.float32)
    J1, g1 = mk(Modeling, m, rx).gradient(None, None, obs, storage=storage, history=hist)
    J2, g2 = mk(Modeling, m, rx[::-1].copy()).gradient(None, None, obs[::-1].copy(), storage=storage, history=hist)
    assert abs(J1 - J2) <= 1e-12 * J1, (J1, J2)
    for i in range(4):
        assert rel_l2(g1[i], g2[i]) <= 1e-12, (i, rel_l2(g1[i], g2[i]))
    Jr, gr = mk(Oracle, m, rx, prec="f64").gradient(None, None, obs)
    assert abs(J1 - Jr) <= 1e-5 * Jr
    errs = [rel_l2(g1[i], gr[i]) for i in range(4)]
    print("  vs FP64", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= (1e-4 if storage == "f32" else 1e-3), errs


def test_gradient_finite_differences_16_directions():
    m, obs = _obs(0, 2)
    g = _mk(Modeling, m, 0, seis=2)
    J0, grad = g.gradient(*obs)

    def J_of(model):
        g.update_model(model["vp"], model["vs"], model["rho"], model["eps"], model["delta"], model["theta"])
        g.run()
        r = g.traces(4).astype(np.float64) - obs[2]
        return 0.5 * float((r * r).sum())

    assert abs(J_of(m) - J0) <= 1e-6 * J0
    rng = np.random.default_rng(42)
    keys = [("vp", 0, 2.0), ("eps", 1, 2e-3), ("delta", 2, 2e-3), ("theta", 3, 0.5)]
    for d in range(16):
        key, gi, sc = keys[d % 4]
        dm = rng.standard_normal(m[key].shape)
        mp, mm = dict(m), dict(m)
        mp[key] = (m[key] + sc * dm).astype(np.float32)
        mm[key] = (m[key] - sc * dm).astype(np.float32)
        step = mp[key].astype(np.float64) - mm[key].astype(np.float64)
        fd = J_of(mp) - J_of(mm)
        an = float((grad[gi] * step).sum())
        assert abs(fd - an) <= 5e-3 * abs(fd), (d, key, fd, an)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from fdmod_hip.torch_misfit import misfit

    m, obs = _obs(0, 2)
    g = _mk(Modeling, m, 0, seis=2)
    J_ref, grad_ref = g.gradient(*obs)
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    ts = {k: torch.tensor(m[k], device=dev, requires_grad=True) for k in ("vp", "eps", "delta", "theta")}
    J = misfit(ts["vp"], ts["eps"], ts["delta"], ts["theta"], g, m["vs"], m["rho"], {"p": obs[2]})
    (3.0 * J).backward()
    assert abs(J.item() - J_ref) <= 1e-12 * J_ref
    for i, k in enumerate(("vp", "eps", "delta", "theta")):
        got = ts[k].grad.double().cpu().numpy()
        assert ts[k].grad.device.type == dev
        assert rel_l2(got, 3.0 * grad_ref[i]) <= 1e-6, k
    with torch.no_grad():
        J2 = misfit(ts["vp"], ts["eps"], ts["delta"], ts["theta"], g, m["vs"], m["rho"], {"p": obs[2]})
    assert abs(J2.item() - J_ref) <= 1e-6 * J_ref
    # the storage and history options reach the library
    for k in ts.values():
        k.grad = None
    J3 = misfit(ts["vp"], ts["eps"], ts["delta"], ts["theta"], g, m["vs"], m["rho"], {"p": obs[2]},
                storage="f16", history=True)
    J3.backward()
    J_f16, grad_f16 = g.gradient(*obs, storage="f16", history=True)
    assert J3.item() == J_f16
    assert rel_l2(ts["vp"].grad.double().cpu().numpy(), grad_f16[0]) <= 1e-6


def test_dot_product_at_gradient_workload_size():
    # the adjoint-gradient workload's grid: a 1126 x 2000 crop, nt = 4000, dt = 0.8 ms
    cfg = synth.CONFIGS["large"]
    m = synth.crop(synth.model_large(), 2000, 2000)
    nt = 4000
    print("dot at 1126x2000, nt=4000", flush=True)
    g = Modeling("bond", cfg["nz"], 2000, nt, cfg["dx"], cfg["dz"], cfg["dt"], ealThick=cfg["ealThick"],
                 freesurface=0, fpeak=cfg["fpeak"])
    g.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    g.set_source(1, [1000], [cfg["sz"]], synth.ricker(cfg["fpeak"], cfg["dt"], nt)[None])
    rx = np.arange(2000)
    g.set_receivers(2, 1, rx, np.full(2000, cfg["rz"]))
    y = np.random.default_rng(5).standard_normal((2000, nt)).astype(np.float32)
    lhs, rhs = g.adjoint_dot(None, None, y)
    assert np.isfinite(lhs) and abs(lhs - rhs) <= 1e-4 * abs(lhs), (lhs, rhs)


def _obs_src(fs, stype, seis, mode, form):
    m = _model()
    true = {k: v.copy() for k, v in m.items()}
    true["vp"] = (true["vp"] * 1.03).astype(np.float32)
    true["theta"] = (true["theta"] + 4).astype(np.float32)
    ot = _mk(Oracle, true, fs, stype=stype, seis=seis, mode=mode, form=form, prec="f64")
    ot.run()
    use = _use(seis)
    return m, [ot.traces(c).astype(np.float32) if use[c] else None for c in (1, 3, 4)]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs,stype,seis,mode", [(0, 1, 2, 1), (1, 1, 3, 1), (1, 2, 3, 2), (0, 4, 1, 1)])
def test_gradient_boundary_saving_reconstruction(fs, stype, seis, mode, form):
    # The loop run backwards over the undamped interior (saved band states, saved
    # history outside the interior) rebuilds the forward wavefield to FP32 rounding:
    # the gradient matches the checkpointed one and the FP64 oracle's.
    print("reconstruct", form, fs, stype, seis, mode, flush=True)
    m, obs = _obs_src(fs, stype, seis, mode, form)
    g = _mk(Modeling, m, fs, stype=stype, seis=seis, mode=mode, form=form)
    Jc, gc = g.gradient(*obs)
    J, gr = g.gradient(*obs, history="reconstruct")
    assert J == Jc  # same forward arithmetic
    errs = [rel_l2(gr[i], gc[i]) for i in range(4)]
    print("  vs checkpoints", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= 1e-5, errs
    J, ga = g.gradient(*obs, ncheckpoints=7, history="reconstruct")  # with restart states
    assert J == Jc and max(rel_l2(ga[i], gc[i]) for i in range(4)) <= 1e-5
    o = _mk(Oracle, m, fs, stype=stype, seis=seis, mode=mode, form=form, prec="f64")
    Jr, gref = o.gradient(*obs)
    for i in range(4):
        assert rel_l2(gr[i], gref[i]) <= 1e-4, i


def test_reconstruction_rejects_fp16_storage():
    from fdmod_hip import FdmodError

    m, obs = _obs(0, 2)
    g = _mk(Modeling, m, 0, seis=2)
    with pytest.raises(FdmodError, match="FP32"):
        g.gradient(*obs, storage="f16", history="reconstruct")


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fs", [0, 1])
def test_gradient_reconstruction_interior(fs, form):
    # Without the absorbing layers' imaging the gradient changes only where the
    # transpose of the edge padding adds those layers: the first and last model
    # columns, the bottom row and, without a free surface, the top row.
    print("reconstruct_interior", form, fs, flush=True)
    m, obs = _obs_src(fs, 1, 2, 1, form)
    g = _mk(Modeling, m, fs, seis=2, form=form)
    Jc, gc = g.gradient(*obs)
    J, gi = g.gradient(*obs, history="reconstruct_interior")
    assert J == Jc
    inner = (slice(1, -1), slice(0 if fs else 1, -1))
    errs = [rel_l2(gi[i][inner], gc[i][inner]) for i in range(4)]
    edge = [rel_l2(gi[i], gc[i]) for i in range(4)]
    print("  inner", ["%.2e" % e for e in errs], "whole", ["%.2e" % e for e in edge], flush=True)
    assert max(errs) <= 1e-5, errs
