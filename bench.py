#!/usr/bin/env python
"""Benchmark: plain forward RSG shot on one MI355X, 1126 x 7872 model, FP32.

One "step" is one time step of the modelling loop (velocity sweep, stress sweep,
free surface, source, receivers) over the full padded grid.  The workload is the
full-size marine TTI model (nz=1126, nx=7872, dx=dz=10 m, dt=0.8 ms, ealThick=60,
free surface on, one 12 Hz Ricker source, 7872 pressure receivers), built from a
fixed seed; inputs are resident in HBM before the timed region.

value = updated grid points per step x steps / wall time  (Gpts/s), where the
updated points are (Nx-10) * (Nz-5-izStart) as in the reference loop bounds.

Also reported:
  roofline      the stress sweep (the larger of the two sweeps) timed with HIP
                events between the launches of loop iterations on the library's
                stream: algorithmic bytes / mean launch time
  forms         Gpts/s of all three rotation forms
  cpu_baseline  the FP32 CPU restatement of the reference's OpenMP loop (the oracle,
                a port) on the host cores, on a bounded number of steps of the same
                workload, plus a bitwise comparison of the GPU wavefield after the
                same steps
  batched       64 uneven aperture crops in one batched, graph-captured loop against
                the same shots run one after another
  gradients     misfit gradients wrt (vp, eps, delta, theta): one shot on a 1126 x 2000
                crop (plain FP32 checkpointed, and FP16 strain history), and 32 shots
                (batched FP16 against the plain single-shot path); the FP64 oracle
                gradient on a reduced replica is the CPU reference and the error bar,
                with the observed data shared, and with each path's own data
"""
import argparse
import json
import os
import sys
import time

# The CPU baseline's OpenMP placement (one thread per physical core, spread over the
# cores); set before any library in this process starts an OpenMP runtime.
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")
# The CPUs this process may use, read before an OpenMP runtime binds the main thread to
# its first place (after which the thread's own mask holds one CPU).
_CPUS = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tti-fdmodeling_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per updated point per step (state R+W and parameter words, FP32):
#   velocity sweep: 5 state reads + 2 writes + params; stress sweep: 5 reads + 3 (6 tensor) writes + params
SWEEP_BYTES = {
    "bond": {"v": 4 * (7 + 3), "s": 4 * (8 + 8)},
    "coord": {"v": 4 * (7 + 5), "s": 4 * (8 + 8)},
    "tensor": {"v": 4 * (7 + 3), "s": 4 * (11 + 9)},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(nt):
    from fdmod_hip import synth

    cfg = dict(synth.CONFIGS["large"])
    cfg["nt"] = nt
    m = synth.model_large(nz=cfg["nz"], nx=cfg["nx"], dx=cfg["dx"])
    src = synth.ricker(cfg["fpeak"], cfg["dt"], nt)
    return cfg, m, src


def make(cls, form, cfg, m, src):
    o = cls(form, cfg["nz"], cfg["nx"], cfg["nt"], cfg["dx"], cfg["dz"], cfg["dt"], ealThick=cfg["ealThick"],
            eal_alpha=-1.0, freesurface=cfg["freesurface"], fpeak=cfg["fpeak"], scalar=1.0, mode=1)
    o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    o.set_source(1, [cfg["sx"]], [cfg["sz"]], src[None])
    o.set_receivers(2, 1, np.arange(cfg["nx"]), np.full(cfg["nx"], cfg["rz"]))
    return o


def gpu_sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def time_form(form, cfg, m, src, steps, warmup, g=None):
    from fdmod_hip import Modeling

    if g is None:
        g = make(Modeling, form, cfg, m, src)
    pts = (g.ixEnd - g.ixStart) * (g.izEnd - g.izStart)
    if warmup:
        g.run(0, warmup)
    g.reset()
    gpu_sync()
    t0 = time.perf_counter()
    dev_s = g.run(0, steps)
    gpu_sync()
    wall = time.perf_counter() - t0
    return g, pts, wall, dev_s


def host_cpu():
    """CPU model, and the logical CPUs, physical cores and sockets this process may run
    on (its affinity mask, read through sysfs topology)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                if k.strip() == "model name":
                    model = v.strip()
                    break
    except OSError:
        pass
    cpus = _CPUS
    cores, sockets = set(), set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            pkg = open(base + "physical_package_id").read().strip()
            core = open(base + "core_id").read().strip()
        except OSError:
            pkg, core = "0", str(c)
        cores.add((pkg, core))
        sockets.add(pkg)
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores), "sockets": len(sockets)}


def cpu_baseline(cfg, m, src, form, steps, g_fields):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    host = host_cpu()
    cores = int(os.environ.get("FDMOD_CPU_THREADS", "0")) or host["physical_cores"]
    oracle.set_threads(cores)
    o = make(oracle.Oracle, form, cfg, m, src)
    sec = o.run(0, steps)
    pts = (o.ixEnd - o.ixStart) * (o.izEnd - o.izStart)
    res = {"value": round(pts * steps / sec / 1e9, 4), "unit": "Gpts/s", "cores": cores, "kind": "port",
           "sample": f"first {steps} time steps of the same {form} shot (loop time only, omp_get_wtime)",
           "seconds": round(sec, 3), "host": host,
           "openmp": {"threads": cores, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                      "OMP_PLACES": os.environ.get("OMP_PLACES")}}
    same = True
    worst = 0.0
    for name, ga in g_fields.items():
        oa = o.get(name).astype(np.float32)
        eq = (ga == oa) | (np.isnan(ga) & np.isnan(oa))
        same &= bool(eq.all())
        d = np.abs(ga.astype(np.float64) - oa)
        d = d[np.isfinite(d)]
        if d.size:
            worst = max(worst, float(d.max()))
    parity = {"steps": steps, "fields": sorted(g_fields), "bitwise_equal": same, "max_abs_diff": worst}
    return res, parity


def batched_vs_loop(form, m, cfg, steps, nshots=64, seed=3):
    """64 shots on the full model, each on its own aperture crop of 1500-4000 columns
    (seeded, centred on the source and clamped to the model), receivers in every crop
    column: all shots advanced together (batched launches, captured hipGraph loop,
    autotuned tiles) against the same shots run one after another on the single-shot
    path.  Both time `steps` time steps per shot; inputs resident before timing."""
    from fdmod_hip import Batch, Modeling, synth

    rng = np.random.default_rng(seed)
    nx_full = m["vp"].shape[0]
    centres = np.linspace(300, nx_full - 300, nshots).astype(int)
    widths = rng.integers(1500, 4001, size=nshots)
    crops = []
    for c, w in zip(centres, widths):
        x0 = int(np.clip(c - w // 2, 0, nx_full - w))
        crops.append((x0, int(w), int(c - x0)))
    src = synth.ricker(cfg["fpeak"], cfg["dt"], steps)
    kw = dict(ealThick=cfg["ealThick"], freesurface=cfg["freesurface"], fpeak=cfg["fpeak"])
    b = Batch(form, cfg["nz"], [w for _, w, _ in crops], steps, cfg["dx"], cfg["dz"], cfg["dt"], **kw)
    pts = 0
    for s, (x0, w, sx) in enumerate(crops):
        mc = synth.crop(m, x0, w)
        b.set_model(s, mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
        b.set_source(s, 1, [sx], [cfg["sz"]], src[None])
        b.set_receivers(s, 2, 1, np.arange(w), np.full(w, cfg["rz"]))
        izs = cfg["ealThick"] + 5 if cfg["freesurface"] else 5
        pts += (w + 2 * (cfg["ealThick"] + 5) - 10) * (cfg["nz"] + 2 * (cfg["ealThick"] + 5) - 5 - izs)
    log(f"bench: batched {nshots} shots, autotuning tiles")
    tiles = b.tune(5)
    b.run(0, min(32, steps))
    b.reset()
    gpu_sync()
    t0 = time.perf_counter()
    b.run(0, steps)
    gpu_sync()
    wb = time.perf_counter() - t0
    b.close()
    log(f"bench: the same {nshots} shots one after another")
    wl = 0.0
    for s, (x0, w, sx) in enumerate(crops):
        mc = synth.crop(m, x0, w)
        g = Modeling(form, cfg["nz"], w, steps, cfg["dx"], cfg["dz"], cfg["dt"], **kw)
        g.set_model(mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
        g.set_source(1, [sx], [cfg["sz"]], src[None])
        g.set_receivers(2, 1, np.arange(w), np.full(w, cfg["rz"]))
        gpu_sync()
        t0 = time.perf_counter()
        g.run(0, steps)
        gpu_sync()
        wl += time.perf_counter() - t0
        g.close()
    return {"workload": f"{nshots} shots, {form} form, aperture crops 1500-4000 columns of the 1126x7872 model "
                        f"(seed {seed}), 1126 deep, free surface, receivers in every crop column",
            "steps": steps, "updated_points_per_step": pts, "tiles_segLV_segLS": list(tiles),
            "gpts_batched": round(pts * steps / wb / 1e9, 3), "gpts_single_shot_loop": round(pts * steps / wl / 1e9, 3),
            "speedup": round(wl / wb, 3)}


GRAD_CFG = dict(nz=1126, dx=10.0, dz=10.0, dt=8e-4, nt=4000, ealThick=60, freesurface=0, fpeak=12.0, sz=2, rz=2)


def _grad_shot(cls, m, w, sx, nt, form="bond", **kw):
    c = GRAD_CFG
    g = cls(form, c["nz"], w, nt, c["dx"], c["dz"], c["dt"], ealThick=c["ealThick"], freesurface=c["freesurface"],
            fpeak=c["fpeak"], **kw)
    g.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    from fdmod_hip import synth

    g.set_source(1, [sx], [c["sz"]], synth.ricker(c["fpeak"], c["dt"], nt)[None])
    g.set_receivers(2, 1, np.arange(w), np.full(w, c["rz"]))
    return g


def _rel(a, b):
    return [float(np.linalg.norm(a[i] - b[i]) / max(np.linalg.norm(b[i]), 1e-300)) for i in range(4)]


GRAD_MODES = {"plain_f32_checkpoints": dict(), "f16_history": dict(storage="f16", history=True),
              "f16_checkpoints": dict(storage="f16all"),
              "f32_reconstruct": dict(history="reconstruct"),
              "f32_reconstruct_interior": dict(history="reconstruct_interior")}


def gradient_single(full, nt, modes=tuple(GRAD_MODES), forms=("bond", "tensor", "coord")):
    """Adjoint gradient workload: one shot on a 1126 x 2000 crop, nt steps.  Bond form:
    the plain FP32 checkpointed gradient, the FP16 full-history gradient and the FP32
    boundary-saving reconstruction (exact, and without the absorbing layers' imaging);
    tensor and coordinate forms: the plain and the FP16-history gradients."""
    from fdmod_hip import Modeling, synth

    W, x0 = 2000, 2936
    bg = synth.crop(full, x0, W)
    tr = synth.add_anomalies(bg)  # five anomalies inside the crop
    out = {"workload": f"one shot, 1126 x {W} crop of the seeded model (columns {x0}..{x0 + W - 1}), true model "
                       f"= background + 5 Gaussian anomalies inside the crop at 0.8-2.8 km depth, nt={nt}, dt=0.8 ms, "
                       f"no free surface, {W} pressure receivers; gradient wrt vp, eps, delta, theta"}
    for form in forms:
        gt = _grad_shot(Modeling, tr, W, W // 2, nt, form=form)
        gt.run()
        obs = gt.traces(4).copy()
        gt.close()
        g = _grad_shot(Modeling, bg, W, W // 2, nt, form=form)
        res, fo = {}, {}
        for name in modes:
            if form != "bond" and name.startswith("f32_reconstruct"):
                continue
            kw = GRAD_MODES[name]
            g.gradient(obs_p=obs, **kw)  # warm-up (allocations)
            gpu_sync()
            t0 = time.perf_counter()
            J, gr = g.gradient(obs_p=obs, **kw)
            wall = time.perf_counter() - t0
            res[name] = gr
            fo[name] = {"seconds": round(wall, 4), "gradients_per_hour": round(3600.0 / wall, 1), "misfit": J}
        ref = res.get("plain_f32_checkpoints")
        if ref is not None and "f16_history" in res:
            fo["f16_vs_f32_rel_l2"] = [round(x, 7) for x in _rel(res["f16_history"], ref)]
        if ref is not None and "f16_checkpoints" in res:
            fo["f16_checkpoints_vs_f32_rel_l2"] = [round(x, 7) for x in _rel(res["f16_checkpoints"], ref)]
        if ref is not None and "f32_reconstruct" in res:
            fo["reconstruct_vs_f32_rel_l2"] = [float("%.3g" % x) for x in _rel(res["f32_reconstruct"], ref)]
        if ref is not None and "f32_reconstruct_interior" in res:
            inner = (slice(None), slice(1, -1), slice(1, -1))  # off the model's edge cells
            fo["reconstruct_interior_vs_f32_rel_l2_inner"] = [
                float("%.3g" % x) for x in _rel(res["f32_reconstruct_interior"][inner], ref[inner])]
        g.close()
        if form == "bond":
            out.update(fo)  # the Bond form's figures stay at the top level
        else:
            out[form] = fo
    return out


def gradient_batched(full, true_full, nt, nshots, seed=5):
    """Batched gradients: nshots shots with crops W ~ U{1500..3000} around uniformly
    drawn sources, FP16 strain history, FP64 sum over shots into the full-model
    gradient; against the plain FP32 checkpointed single-shot path on the same shots."""
    from fdmod_hip import Batch, Modeling, synth

    nx_full = full["vp"].shape[0]
    rng = np.random.default_rng(seed)
    shots = []
    for _ in range(nshots):
        w = int(rng.integers(1500, 3001))
        sx = int(rng.integers(0, nx_full))
        x0 = int(min(max(sx - w // 2, 0), nx_full - w))
        shots.append((x0, w, sx - x0))
    log(f"bench: observed data for {nshots} shots (batched forward, true model)")
    kw = dict(ealThick=GRAD_CFG["ealThick"], freesurface=GRAD_CFG["freesurface"], fpeak=GRAD_CFG["fpeak"])
    c = GRAD_CFG
    bt = Batch("bond", c["nz"], [w for _, w, _ in shots], nt, c["dx"], c["dz"], c["dt"], **kw)
    for s, (x0, w, sx) in enumerate(shots):
        mc = synth.crop(true_full, x0, w)
        bt.set_model(s, mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
        bt.set_source(s, 1, [sx], [c["sz"]], synth.ricker(c["fpeak"], c["dt"], nt)[None])
        bt.set_receivers(s, 2, 1, np.arange(w), np.full(w, c["rz"]))
    bt.run()
    obs = [bt.traces(s, 4).copy() for s in range(nshots)]
    bt.close()
    b = Batch("bond", c["nz"], [w for _, w, _ in shots], nt, c["dx"], c["dz"], c["dt"], **kw)
    for s, (x0, w, sx) in enumerate(shots):
        mc = synth.crop(full, x0, w)
        b.set_model(s, mc["vp"], mc["vs"], mc["rho"], mc["eps"], mc["delta"], mc["theta"])
        b.set_source(s, 1, [sx], [c["sz"]], synth.ricker(c["fpeak"], c["dt"], nt)[None])
        b.set_receivers(s, 2, 1, np.arange(w), np.full(w, c["rz"]))
    # Both legs are timed warm: each path's device workspace (the batch's pooled
    # gradient memory, each single-shot context's checkpoints and segment history) is
    # allocated by an untimed first call, as it is once per context in an inversion.
    log(f"bench: batched FP16 gradients of {nshots} shots (warm-up call, then timed)")
    x0s = [x0 for x0, _, _ in shots]
    b.gradient(obs_p=obs, storage="f16", history=True, x0=x0s, nx_full=nx_full)
    gpu_sync()
    t0 = time.perf_counter()
    Jb, gb = b.gradient(obs_p=obs, storage="f16", history=True, x0=x0s, nx_full=nx_full)
    wb = time.perf_counter() - t0
    b.close()
    log(f"bench: plain FP32 checkpointed gradients of the same {nshots} shots, one after another")
    gp = np.zeros_like(gb)
    wp, Jp = 0.0, 0.0
    for s, (x0, w, sx) in enumerate(shots):
        g = _grad_shot(Modeling, synth.crop(full, x0, w), w, sx, nt)
        g.gradient(obs_p=obs[s])  # warm-up: allocates the context's gradient workspace
        gpu_sync()
        t0 = time.perf_counter()
        J, gr = g.gradient(obs_p=obs[s])
        wp += time.perf_counter() - t0
        gp[:, x0:x0 + w] += gr
        Jp += J
        g.close()
    pts = sum((w + 2 * (c["ealThick"] + 5) - 10) * (c["nz"] + 2 * (c["ealThick"] + 5) - 10) for _, w, _ in shots)
    return {"workload": f"{nshots} shots, Bond form, crops W ~ U{{1500..3000}} of the seeded 1126 x {nx_full} model "
                        f"(seed {seed}) around uniformly drawn sources, true model = background + 5 Gaussian "
                        f"anomalies at 0.8-2.8 km depth, nt={nt}, dt=0.8 ms, no free surface, pressure receivers in every crop column",
            "updated_points_per_step": pts,
            "timing": "both legs warm (workspace allocated by an untimed first call)",
            "batched_f16": {"seconds": round(wb, 3), "gradients_per_hour": round(3600.0 * nshots / wb, 1),
                            "storage": "FP16 strain history, FP32 checkpoint-free, FP64 imaging and shot sum"},
            "plain_f32_single_shot": {"seconds": round(wp, 3), "gradients_per_hour": round(3600.0 * nshots / wp, 1),
                                      "storage": "FP32, ceil(sqrt(nt)) checkpoints with recomputation"},
            "speedup": round(wp / wb, 3),
            "misfit_rel_diff": abs(Jb - Jp) / Jp,
            "batched_vs_plain_rel_l2": [round(x, 7) for x in _rel(gb, gp)]}


def cpu_gradient_c4(nt=4000, form="bond"):
    """The FP64 oracle gradient on the gradient workload's own geometry (1126 x 2000 crop,
    nt = 4000, checkpointed recomputation, OpenMP on the host cores) and the GPU
    gradients' relative L2 errors against it, per parameter (vp, eps, delta, theta), with
    the observed data shared (modelled by the FP64 oracle) and with each path's own data
    (the GPU's FP32 modelling for the GPU gradients); the FP32 oracle (the reference's
    arithmetic) is measured the same way."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from fdmod_hip import Modeling, synth

    host = host_cpu()
    cores = int(os.environ.get("FDMOD_CPU_THREADS", "0")) or host["physical_cores"]
    os.environ.setdefault("ORACLE_PROGRESS", "500")  # the FP64 legs run for minutes each
    oracle.set_threads(cores)
    oracle.set_threads(cores, "f64")
    W, x0 = 2000, 2936
    bg = synth.crop(synth.model_large(), x0, W)
    tr = synth.add_anomalies(bg)
    log("bench: FP64 oracle forward on the true model (1126 x 2000, nt = %d, %s form)" % (nt, form))
    t0 = time.perf_counter()
    ot = _grad_shot(oracle.Oracle, tr, W, W // 2, nt, form=form, prec="f64")
    ot.run()
    obs64 = ot.traces(4).astype(np.float32)
    ot.close()
    log("bench: FP64 oracle gradient")
    o = _grad_shot(oracle.Oracle, bg, W, W // 2, nt, form=form, prec="f64")
    J64, g64 = o.gradient(obs_p=obs64)
    o.close()
    w64 = time.perf_counter() - t0
    log("bench: FP32 oracle gradients (shared and own data)")
    o32 = _grad_shot(oracle.Oracle, bg, W, W // 2, nt, form=form, prec="f32")
    _, g32 = o32.gradient(obs_p=obs64)
    ot32 = _grad_shot(oracle.Oracle, tr, W, W // 2, nt, form=form, prec="f32")
    ot32.run()
    _, g32o = o32.gradient(obs_p=ot32.traces(4).astype(np.float32))
    ot32.close()
    o32.close()
    errs = {"reference_fp32_oracle": [round(x, 7) for x in _rel(g32, g64)]}
    own = {"reference_fp32_oracle": [round(x, 7) for x in _rel(g32o, g64)]}
    gt = _grad_shot(Modeling, tr, W, W // 2, nt, form=form)
    gt.run()
    obs_gpu = gt.traces(4).copy()
    gt.close()
    g = _grad_shot(Modeling, bg, W, W // 2, nt, form=form)
    for name in ("plain_f32_checkpoints", "f16_history", "f16_checkpoints"):
        _, gg = g.gradient(obs_p=obs64, **GRAD_MODES[name])
        errs[name] = [round(x, 7) for x in _rel(gg, g64)]
        _, gg = g.gradient(obs_p=obs_gpu, **GRAD_MODES[name])
        own[name] = [round(x, 7) for x in _rel(gg, g64)]
    g.close()
    return {"sample": f"FP64 oracle gradient on the 1126 x {W} crop, nt={nt}, {form} form, no free surface, true model "
                      "= background + 5 anomalies inside the crop; ceil(sqrt(nt)) checkpoints, OpenMP adjoint",
            "seconds_fp64_forward_and_gradient": round(w64, 1), "cores": cores, "misfit_fp64": J64,
            "gpu_rel_l2_vs_fp64": errs, "gpu_rel_l2_vs_fp64_own_data": own}


def cpu_gradient_replica(seed=7):
    """FP64 oracle gradient (the CPU reference for the adjoint) on a reduced replica of
    the gradient workload, timed, and the GPU FP32 and FP16 gradients on the same
    replica compared with it."""
    from fdmod_hip import Modeling, synth
    import oracle
    from oracle import Oracle

    cores = int(os.environ.get("FDMOD_CPU_THREADS", "0")) or host_cpu()["physical_cores"]
    oracle.set_threads(cores, "f64")
    oracle.set_threads(cores)
    nz, W, nt = 141, 250, 800
    full = synth.model_large(nz=nz, nx=W, seed=seed)
    # a change the data see within nt steps: +3% vp in the rock, plus two anomalies
    true = synth.add_anomalies(full, n=2, zrange=(0.35, 0.85))
    true["vp"] = np.where(full["vs"] > 0, true["vp"] * 1.03, true["vp"]).astype(np.float32)
    kw = dict(ealThick=30, freesurface=0, fpeak=12.0)

    def mk(cls, m, **extra):
        g = cls("bond", nz, W, nt, 10.0, 10.0, 8e-4, **kw, **extra)
        g.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
        g.set_source(1, [W // 2], [2], synth.ricker(12.0, 8e-4, nt)[None])
        g.set_receivers(2, 1, np.arange(W), np.full(W, 2))
        return g

    ot = mk(Oracle, true, prec="f64")
    ot.run()
    obs = ot.traces(4).astype(np.float32)
    o = mk(Oracle, full, prec="f64")
    t0 = time.perf_counter()
    Jr, gr = o.gradient(obs_p=obs)
    wc = time.perf_counter() - t0
    # the reference's own FP32 arithmetic (the FP32 oracle) against FP64: the floor the
    # GPU's FP32 stencils share
    o32 = mk(Oracle, full, prec="f32")
    J32, g32 = o32.gradient(obs_p=obs)
    g = mk(Modeling, full)
    errs = {"reference_fp32_oracle": [round(x, 7) for x in _rel(g32, gr)]}
    paths = (("plain_f32_checkpoints", dict()), ("f16_history", dict(storage="f16", history=True)),
             ("f16_checkpoints", dict(storage="f16")))
    for name, kwg in paths:
        J, gg = g.gradient(obs_p=obs, **kwg)
        errs[name] = [round(x, 7) for x in _rel(gg, gr)]
    # each pipeline with observed data modelled in its own arithmetic (the FP64 reference
    # keeps its FP64 data): the FP32 rounding of the direct wave, which the true and the
    # background model share, then cancels in the residual as it does in an FP32 inversion
    ot32 = mk(Oracle, true, prec="f32")
    ot32.run()
    _, g32o = o32.gradient(obs_p=ot32.traces(4).astype(np.float32))
    own = {"reference_fp32_oracle": [round(x, 7) for x in _rel(g32o, gr)]}
    gt = mk(Modeling, true)
    gt.run()
    obs_gpu = gt.traces(4).copy()
    gt.close()
    for name, kwg in paths:
        J, gg = g.gradient(obs_p=obs_gpu, **kwg)
        own[name] = [round(x, 7) for x in _rel(gg, gr)]
    g.close()
    return {"sample": f"FP64 oracle gradient (OpenMP adjoint) on a {nz} x {W} replica, nt={nt}, ealThick=30, "
                      "Bond form, no free surface; data from the FP64 oracle on the model with +3% rock vp and "
                      "two anomalies", "seconds": round(wc, 3),
            "gradients_per_hour": round(3600.0 / wc, 2), "kind": "port", "cores": cores,
            "gpu_rel_l2_vs_fp64": errs,
            "gpu_rel_l2_vs_fp64_own_data": own}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=200)
    # headline form: the coordinate-rotation program, whose traces on the marine model are
    # finite (the Bond program's free-surface fix-up is NaN in water, as in the reference)
    ap.add_argument("--form", default="coord", choices=["bond", "coord", "tensor"])
    ap.add_argument("--cpu-steps", type=int, default=1500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=6000)
    ap.add_argument("--grad-shots", type=int, default=32)
    ap.add_argument("--grad-steps", type=int, default=4000)
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--grad-only", action="store_true", help="only the single-shot gradient workload (profiling)")
    ap.add_argument("--grad-modes", default=",".join(GRAD_MODES), help="single-shot gradient modes (comma list)")
    ap.add_argument("--grad-forms", default="bond,tensor,coord", help="forms of the single-shot gradient workload")
    ap.add_argument("--fp64-c4", action="store_true",
                    help="with --grad-only: the FP64 oracle gradient on the 1126 x 2000 workload (minutes of CPU)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU shot sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    modes = tuple(x for x in args.grad_modes.split(",") if x)
    gforms = tuple(x for x in args.grad_forms.split(",") if x)
    if any(x not in GRAD_MODES for x in modes):
        ap.error(f"--grad-modes: choose from {','.join(GRAD_MODES)}")
    if args.grad_only:
        from fdmod_hip import synth

        log("bench: gradient workload only")
        full = synth.model_large()
        out = {"single_shot": gradient_single(full, args.grad_steps, modes, gforms)}
        if args.fp64_c4:
            out["cpu_fp64_c4"] = {fm: cpu_gradient_c4(args.grad_steps, fm) for fm in gforms}
        print(json.dumps({"gradients": out}))
        return
    nt = max(args.steps, args.cpu_steps, 1)
    log(f"bench: building the 1126x7872 workload (nt={nt})")
    cfg, m, src = workload(nt)

    from fdmod_hip import Modeling

    # The per-sweep timing runs first: its loop iterations also bring the GPU out of idle
    # before the warm-up and timed steps (a 20-step run straight after the setup ran at
    # 55.5 Gpts/s, after 200 steps of warm-up at 58.7).  Its first 100 iterations are
    # discarded: started on an idle GPU they read the velocity sweep 10% slow.
    g = make(Modeling, args.form, cfg, m, src)
    log("bench: per-sweep HIP-event timing")
    g.time_sweeps(100)
    g.reset()
    v_ms, s_ms = g.time_sweeps(200)
    g.reset()
    log(f"bench: timing {args.form}, {args.steps} steps after {args.warmup} warm-up steps")
    g, pts, wall, dev_s = time_form(args.form, cfg, m, src, args.steps, args.warmup, g)
    value = pts * args.steps / wall / 1e9
    p_tr = g.traces(4)
    finite = bool(np.isfinite(p_tr[:, :args.steps]).all())
    # The same shot with the host buffers inside the timed region (the library's boundary
    # takes host arrays): model upload and setup, source and receiver upload, the loop and
    # the traces' download.  Reported beside value, never as it.
    log("bench: host-inclusive run (upload, setup, loop, trace download)")
    gpu_sync()
    t0 = time.perf_counter()
    gh = make(Modeling, args.form, cfg, m, src)
    gh.run(0, args.steps)
    gh.traces(4)
    gpu_sync()
    wall_h = time.perf_counter() - t0
    gh.close()
    host_inclusive = {"gpts": round(pts * args.steps / wall_h / 1e9, 3), "seconds": round(wall_h, 4),
                      "includes": "model upload and setup (MEAL, Cij, flags), source and receiver upload, "
                                  f"{args.steps} loop iterations, pressure-trace download"}
    sb = SWEEP_BYTES[args.form]
    achieved = sb["s"] * pts / (s_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": f"k_ssweep<{args.form}>", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_point": sb["s"], "points_per_launch": pts, "ms_per_launch": round(s_ms, 5),
                "v_sweep": {"ms_per_launch": round(v_ms, 5), "bytes_per_point": sb["v"],
                            "achieved": round(sb["v"] * pts / (v_ms * 1e-3) / 1e9, 1)},
                "step_bytes_per_point": sb["v"] + sb["s"],
                "step_achieved": round((sb["v"] + sb["s"]) * pts * args.steps / wall / 1e9, 1)}
    tf = os.path.join(ROOT, "profiles", f"traffic_ssweep_{args.form}_1126x7872.json")
    if os.path.exists(tf):
        with open(tf) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")

    batched = None
    if args.batch_steps > 0:
        batched = batched_vs_loop(args.form, m, cfg, args.batch_steps)

    gradients = None
    if not args.no_grad:
        from fdmod_hip import synth

        log("bench: gradient workloads (true model = background + 5 anomalies)")
        true_full = synth.add_anomalies(m)
        gradients = {"single_shot": gradient_single(m, args.grad_steps, modes, gforms)}
        if args.grad_shots > 0:
            gradients["batched"] = gradient_batched(m, true_full, args.grad_steps, args.grad_shots)

    forms = {}
    if not args.no_forms:
        for fm in ("bond", "coord", "tensor"):
            if fm == args.form:
                forms[fm] = round(value, 3)
                continue
            log(f"bench: timing form {fm} (1000 steps)")
            n = min(1000, args.steps)
            gg, pp, ww, _ = time_form(fm, cfg, m, src, n, 50)
            forms[fm] = round(pp * n / ww / 1e9, 3)
            gg.close()

    cpu = None
    parity = None
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"bench: GPU wavefield after {args.cpu_steps} steps for the bitwise check")
        g.reset()
        g.run(0, args.cpu_steps)
        fields = {k: g.get(k) for k in ("vx", "vz", "sxx", "szz", "sxz")}
        log(f"bench: CPU baseline (oracle FP32 port), {args.cpu_steps} steps")
        cpu, parity = cpu_baseline(cfg, m, src, args.form, args.cpu_steps, fields)
        parity["finite"] = all(bool(np.isfinite(v).all()) for v in fields.values())
        parity = {args.form: parity}
        # the other two forms over 400 steps (the Bond form's marine wavefield is NaN
        # from the free-surface fix-up in water, as in the reference: compared NaN-aware)
        n2 = max(1, min(args.cpu_steps, 400))
        names_of = {"coord": ("vx", "vz", "sxx", "szz", "sxz"), "bond": ("vx", "vz", "sxx", "szz", "sxz"),
                    "tensor": ("vx", "vz", "sxx", "szz", "sxz", "txx", "tzz", "txz")}
        for fm in ("coord", "tensor", "bond"):
            if fm == args.form:
                continue
            log(f"bench: {fm}-form bitwise check, {n2} steps")
            gc = make(__import__("fdmod_hip").Modeling, fm, cfg, m, src)
            gc.run(0, n2)
            fc = {k: gc.get(k) for k in names_of[fm]}
            fc_finite = all(bool(np.isfinite(v).all()) for v in fc.values())
            gc.close()
            _, pc = cpu_baseline(cfg, m, src, fm, n2, fc)
            pc["finite"] = fc_finite
            parity[fm] = pc
        if gradients is not None:
            log("bench: FP64 oracle gradient on a reduced replica")
            gradients["cpu_fp64_reference"] = cpu_gradient_replica()

    out = {
        "metric": "Gpts/s/step batched vs single-shot; gradients/hr at <=1e-3 rel err vs FP64",
        "value": round(value, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded marine TTI model, Ricker source)",
        "config": {"workload": f"plain forward single shot, {args.form} form, nz=1126 nx=7872 dx=dz=10m "
                               f"dt=0.8ms ealThick=60 free surface, 7872 p receivers",
                   "grid_padded": [g.Nz, g.Nx], "updated_points_per_step": pts, "parallelism": "single GPU"},
        "device_seconds": round(dev_s, 4),
        "host_inclusive": host_inclusive,
        "traces_finite": finite,
        "note": ("the reference's Bond free-surface fix-up (BT:543-568) divides 0/0 in water, so this marine "
                 "model's Bond traces are NaN in the reference too and here bit for bit; forms_gpts gives the "
                 "coordinate and tensor forms, whose traces are finite") if args.form == "bond" and not finite
                else None,
        "roofline": roofline,
        "forms_gpts": forms,
        "batched": batched,
        "gradients": gradients,
        "cpu_baseline": cpu,
        "parity_full_size": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
