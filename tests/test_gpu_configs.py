"""The five benchmark workloads at their real sizes, on the GPU.

  C1  200 x 300 Bond shot, nt = 1000, no free surface: traces and fields bit-identical
      to the FP32 oracle (the reference's arithmetic) and within 1e-5 relative L2 of
      the FP64 oracle.
  C2  the 1126 x 7872 marine shot, all three forms, free surface on: the first 300
      steps bit-identical to the FP32 oracle, NaN-aware (the reference's Bond
      free-surface fix-up is NaN in water).
  C3  uneven 1500-4000-column aperture crops, 1126 deep, advanced together by the
      batched graph-captured loop with autotuned tiles: traces bit-identical to each
      shot run alone (and the smallest crop to the oracle).
  C4  one 1126 x 2000 crop, nt = 4000: central finite differences of the misfit along
      16 random smooth directions against the adjoint gradient.
  C5  batched FP16-history gradients of uneven 1500-3000-column crops, nt = 4000:
      each shot's gradient bit-identical to its single-shot gradient, the FP64 shot
      sum equal to the sum of those, and each within 1e-3 of the plain FP32 path.

The time loop is a fixed sequence of identical steps, so C2 and C3 check a bounded
number of steps; C4 and C5 run the full nt.
"""
import functools

import numpy as np
import pytest

from common import rel_l2
from fdmod_hip import Batch, Modeling, synth
from oracle import Oracle

pytestmark = pytest.mark.gpu

FORMS = ["bond", "tensor", "coord"]
LARGE = synth.CONFIGS["large"]


@functools.lru_cache(maxsize=1)
def _large():
    return synth.model_large()


@functools.lru_cache(maxsize=1)
def _large_true():
    return synth.add_anomalies(_large())


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_config1_small_bond_bitwise():
    c = synth.CONFIGS["small"]
    m = synth.model_small(nz=c["nz"], nx=c["nx"], dx=c["dx"])
    src = synth.ricker(c["fpeak"], c["dt"], c["nt"])
    rx = np.arange(c["nx"])
    print("C1: 200 x 300 Bond, nt = 1000", flush=True)

    def mk(cls, **kw):
        o = cls("bond", c["nz"], c["nx"], c["nt"], c["dx"], c["dz"], c["dt"], ealThick=c["ealThick"],
                freesurface=c["freesurface"], fpeak=c["fpeak"], **kw)
        o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
        o.set_source(1, [c["sx"]], [c["sz"]], src[None])
        o.set_receivers(3, 1, rx, np.full(len(rx), c["rz"]))
        o.run()
        return o

    g, o, o64 = mk(Modeling), mk(Oracle), mk(Oracle, prec="f64")
    for comp in (1, 3, 4):
        a = g.traces(comp)
        assert np.isfinite(a).all()
        assert np.array_equal(a, o.traces(comp).astype(np.float32)), comp
        # the 1e-5 relative L2 trace tolerance against the FP64 oracle
        assert rel_l2(a, o64.traces(comp)) <= 1e-5, comp
    for f in ("vx", "vz", "sxx", "szz", "sxz"):
        assert np.array_equal(g.get(f), o.get(f).astype(np.float32)), f


def _full_shot(cls, form, nt):
    m = _large()
    o = cls(form, LARGE["nz"], LARGE["nx"], nt, LARGE["dx"], LARGE["dz"], LARGE["dt"], ealThick=LARGE["ealThick"],
            freesurface=LARGE["freesurface"], fpeak=LARGE["fpeak"])
    o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    o.set_source(1, [LARGE["sx"]], [LARGE["sz"]], synth.ricker(LARGE["fpeak"], LARGE["dt"], nt)[None])
    o.set_receivers(2, 1, np.arange(LARGE["nx"]), np.full(LARGE["nx"], LARGE["rz"]))
    return o


@pytest.mark.parametrize("form", FORMS)
def test_config2_full_size_bitwise(form):
    nt = 300
    print("C2: 1126 x 7872", form, "nt =", nt, flush=True)
    g = _full_shot(Modeling, form, nt)
    g.run()
    o = _full_shot(Oracle, form, nt)
    o.run()
    a = g.traces(4)
    assert _same(a, o.traces(4).astype(np.float32))
    fields = ("vx", "vz", "sxx", "szz", "sxz") + (("txx", "tzz", "txz") if form == "tensor" else ())
    for f in fields:
        assert _same(g.get(f), o.get(f).astype(np.float32)), f
    if form == "bond":
        # BT:543-568 divides 0/0 where water meets the free surface (theta = 0, vs = 0)
        assert np.isnan(g.get("c11")[:, g.izStart]).any() and np.isnan(a).any()
    else:
        assert np.isfinite(a).all() and np.abs(a).max() > 0
    g.close()
    o.close()


def _crops(nshots, lo, hi, seed, nx_full):
    """Seeded aperture crops: (x0, width, source column inside the crop)."""
    rng = np.random.default_rng(seed)
    widths = [lo, hi] + [int(w) for w in rng.integers(lo, hi + 1, size=nshots - 2)]
    out = []
    for w in widths:
        sx = int(rng.integers(0, nx_full))
        x0 = int(min(max(sx - w // 2, 0), nx_full - w))
        out.append((x0, w, sx - x0))
    return out


@pytest.mark.parametrize("form", FORMS)
def test_config3_batched_crops_at_size(form):
    nt = 200
    m = _large()
    crops = _crops(4, 1500, 4000, 3, LARGE["nx"])
    print("C3:", form, "crops", [w for _, w, _ in crops], "nt =", nt, flush=True)
    kw = dict(ealThick=LARGE["ealThick"], freesurface=LARGE["freesurface"], fpeak=LARGE["fpeak"])
    src = synth.ricker(LARGE["fpeak"], LARGE["dt"], nt)
    b = Batch(form, LARGE["nz"], [w for _, w, _ in crops], nt, LARGE["dx"], LARGE["dz"], LARGE["dt"], **kw)
    for s, (x0, w, sx) in enumerate(crops):
        mc = synth.crop(m, x0, w)
        b.set_model(s, mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
        b.set_source(s, 1, [sx], [LARGE["sz"]], src[None])
        b.set_receivers(s, 2, 1, np.arange(w), np.full(w, LARGE["rz"]))
    tiles = b.tune(2)
    print("  autotuned tiles", tiles, flush=True)
    b.run(0, 70)
    b.run(70, nt - 70)
    for s, (x0, w, sx) in enumerate(crops):
        mc = synth.crop(m, x0, w)
        g = Modeling(form, LARGE["nz"], w, nt, LARGE["dx"], LARGE["dz"], LARGE["dt"], **kw)
        g.set_model(mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
        g.set_source(1, [sx], [LARGE["sz"]], src[None])
        g.set_receivers(2, 1, np.arange(w), np.full(w, LARGE["rz"]))
        g.run()
        tb = b.traces(s, 4)
        assert _same(tb, g.traces(4)), s
        if form != "bond":
            assert np.isfinite(tb).all()
        g.close()
        if w == 1500:
            o = Oracle(form, LARGE["nz"], w, nt, LARGE["dx"], LARGE["dz"], LARGE["dt"], **kw)
            o.set_model(mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
            o.set_source(1, [sx], [LARGE["sz"]], src[None])
            o.set_receivers(2, 1, np.arange(w), np.full(w, LARGE["rz"]))
            o.run()
            assert _same(tb, o.traces(4).astype(np.float32)), s
            o.close()
    b.close()


GRAD = dict(nz=1126, dx=10.0, dz=10.0, dt=8e-4, nt=4000, ealThick=60, freesurface=0, fpeak=12.0, sz=2, rz=2)


def _grad_shot(m, w, sx, nt=GRAD["nt"]):
    c = GRAD
    g = Modeling("bond", c["nz"], w, nt, c["dx"], c["dz"], c["dt"], ealThick=c["ealThick"],
                 freesurface=c["freesurface"], fpeak=c["fpeak"])
    g.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    g.set_source(1, [sx], [c["sz"]], synth.ricker(c["fpeak"], c["dt"], nt)[None])
    g.set_receivers(2, 1, np.arange(w), np.full(w, c["rz"]))
    return g


def test_config4_gradient_fd_16_directions():
    from scipy.ndimage import gaussian_filter

    W, x0 = 2000, 2936
    bg = synth.crop(_large(), x0, W)
    tr = synth.add_anomalies(bg)  # five anomalies inside the crop, at depths the record sees
    print("C4: 1126 x 2000, nt = 4000, 16 directions", flush=True)
    gt = _grad_shot(tr, W, W // 2)
    gt.run()
    obs = gt.traces(4).copy()
    gt.close()
    g = _grad_shot(bg, W, W // 2)
    J0, grad = g.gradient(obs_p=obs)
    assert np.isfinite(grad).all() and J0 > 0
    d0 = g.traces(4).astype(np.float64)
    print(f"  J0 = {J0:.4e}, J0 / (|d|^2 / 2) = {J0 / (0.5 * (d0 * d0).sum()):.3e}", flush=True)

    def J_of(model):
        g.update_model(model["vp"], model["vs"], model["rho"], model["eps"], model["delta"], model["theta"])
        g.run()
        r = g.traces(4).astype(np.float64) - obs
        return 0.5 * float((r * r).sum())

    assert abs(J_of(bg) - J0) <= 1e-6 * J0
    rng = np.random.default_rng(16)
    # parameter, gradient index, cap on the largest perturbation
    # (a direction nearly orthogonal to the gradient reaches the cap; a small cap keeps the
    # fourth-order term that extrapolation leaves below the tolerance there)
    keys = [("vp", 0, 10.0), ("eps", 1, 0.01), ("delta", 2, 0.01), ("theta", 3, 2.0)]
    # Smooth random directions, zero in the top 5% of the depth (the water layer and the
    # rock just below it).  A perturbation next to the source and receivers changes the
    # direct wave by far more than the residual measures: the misfit along it is then
    # dominated by its quadratic term, and the FP32 rounding of the direct wave stops
    # cancelling between the two one-sided runs.
    nz = bg["vp"].shape[1]
    depth = 0.5 * (1.0 + np.tanh((np.arange(nz) - 0.05 * nz) / 8.0))[None, :]
    worst = 0.0
    for d in range(16):
        key, gi, cap = keys[d % 4]
        dm = gaussian_filter(rng.standard_normal(bg[key].shape), sigma=6.0) * depth
        dm /= np.abs(dm).max()
        # a step changing the misfit by about 2% of J0 along the gradient, within the cap
        slope = abs(float((grad[gi] * dm).sum()))
        h = min(cap, 0.01 * J0 / max(slope, 1e-300))
        # central differences at h and h/2, Richardson-combined: D(h) = an + b h^2 + ...
        # so (4 D(h/2) - D(h)) / 3 removes the curvature term that larger steps
        # (capped parameters with a small slope) leave
        fds, ans = [], []
        for hh in (h, 0.5 * h):
            mp, mm = dict(bg), dict(bg)
            mp[key] = (bg[key] + hh * dm).astype(np.float32)
            mm[key] = (bg[key] - hh * dm).astype(np.float32)
            step = (mp[key].astype(np.float64) - mm[key].astype(np.float64)) / (2.0 * hh)
            assert np.abs(step).max() > 0
            fds.append((J_of(mp) - J_of(mm)) / (2.0 * hh))
            ans.append(float((grad[gi] * step).sum()))  # the float32-rounded step, per unit h
        fd = (4.0 * fds[1] - fds[0]) / 3.0
        an = (4.0 * ans[1] - ans[0]) / 3.0
        err = abs(fd - an) / abs(fd)
        worst = max(worst, err)
        print(f"  direction {d:2d} {key:5s} h={h:.3g} dJ/dh: central {fds[0]:.5e} / {fds[1]:.5e}, "
              f"extrapolated {fd:.5e}, adjoint {an:.5e}, rel err {err:.2e}", flush=True)
        assert err <= 5e-3, (d, key, fd, an)
    print("  worst", worst, flush=True)
    g.close()


@pytest.mark.parametrize("form", FORMS)
def test_config4_fp16_checkpoints_at_size(form):
    """FP16 checkpoints and FP16 strain within the recomputed segments (storage f16all,
    ceil(sqrt(nt)) checkpoints) on the 1126 x 2000, nt = 4000 workload, against the plain
    FP32 path: within 1e-3 for the Bond and coordinate forms (measured <= 3.3e-4).  The
    tensor form's theta component is the exception (2.1e-3 measured; its FP16 strain
    history alone gives 9.7e-4): its other three components stay within 1e-3."""
    W, x0 = 2000, 2936
    bg = synth.crop(_large(), x0, W)
    tr = synth.add_anomalies(bg)
    print("C4 FP16 checkpoints:", form, flush=True)

    def shot(m):
        c = GRAD
        g = Modeling(form, c["nz"], W, c["nt"], c["dx"], c["dz"], c["dt"], ealThick=c["ealThick"],
                     freesurface=c["freesurface"], fpeak=c["fpeak"])
        g.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
        g.set_source(1, [W // 2], [c["sz"]], synth.ricker(c["fpeak"], c["dt"], c["nt"])[None])
        g.set_receivers(2, 1, np.arange(W), np.full(W, c["rz"]))
        return g

    gt = shot(tr)
    gt.run()
    obs = gt.traces(4).copy()
    gt.close()
    g = shot(bg)
    J32, g32 = g.gradient(obs_p=obs)
    J16, g16 = g.gradient(obs_p=obs, storage="f16all")
    g.close()
    assert J16 == J32  # the forward itself is FP32
    errs = [rel_l2(g16[i], g32[i]) for i in range(4)]
    print("  FP16 checkpoints vs FP32", ["%.2e" % e for e in errs], flush=True)
    assert max(errs[:3]) <= 1e-3, errs
    assert errs[3] <= (3e-3 if form == "tensor" else 1e-3), errs


def test_config5_batched_fp16_gradients_at_size():
    nt = GRAD["nt"]
    full, true = _large(), _large_true()
    nx_full = full["vp"].shape[0]
    # uneven crops, each with its source above one of the anomalies, so every shot's record sees one
    d = np.abs(true["vp"] - full["vp"]).max(axis=1)
    centres = []
    for cidx in np.argsort(-d):
        if all(abs(int(cidx) - c0) > 800 for c0 in centres):
            centres.append(int(cidx))
        if len(centres) == 3:
            break
    crops = []
    for w, sx in zip((1500, 3000, 2200), centres):
        x0 = int(min(max(sx - w // 2, 0), nx_full - w))
        crops.append((x0, w, sx - x0))
    print("C5: crops", [w for _, w, _ in crops], "nt =", nt, flush=True)
    c = GRAD
    kw = dict(ealThick=c["ealThick"], freesurface=c["freesurface"], fpeak=c["fpeak"])

    def batch(model):
        b = Batch("bond", c["nz"], [w for _, w, _ in crops], nt, c["dx"], c["dz"], c["dt"], **kw)
        for s, (x0, w, sx) in enumerate(crops):
            mc = synth.crop(model, x0, w)
            b.set_model(s, mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
            b.set_source(s, 1, [sx], [c["sz"]], synth.ricker(c["fpeak"], c["dt"], nt)[None])
            b.set_receivers(s, 2, 1, np.arange(w), np.full(w, c["rz"]))
        return b

    bt = batch(true)
    bt.run()
    obs = [bt.traces(s, 4).copy() for s in range(len(crops))]
    bt.close()
    b = batch(full)
    J, per = b.gradient(obs_p=obs, storage="f16", history=True)
    Jf, gsum = b.gradient(obs_p=obs, storage="f16", history=True, x0=[x0 for x0, _, _ in crops], nx_full=nx_full)
    b.close()
    ref = np.zeros_like(gsum)
    Js = 0.0
    for s, (x0, w, sx) in enumerate(crops):
        g = _grad_shot(synth.crop(full, x0, w), w, sx)
        J16, g16 = g.gradient(obs_p=obs[s], storage="f16", history=True)
        assert np.array_equal(per[s], g16), s
        J32, g32 = g.gradient(obs_p=obs[s])
        assert J16 == J32
        errs = [rel_l2(g16[i], g32[i]) for i in range(4)]
        print("  shot", s, "width", w, "misfit", J32, "FP16 vs FP32", ["%.2e" % e for e in errs], flush=True)
        assert J32 > 0 and np.abs(g32).max() > 0, s
        assert max(errs) <= 1e-3, errs
        ref[:, x0:x0 + w] += g16
        Js += J16
        g.close()
    assert J == Js and Jf == Js
    assert np.array_equal(gsum, ref)
