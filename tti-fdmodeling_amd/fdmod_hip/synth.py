"""Seeded synthetic workloads: TTI models, Ricker sources and receiver lines.

Arrays use the reference's file layout: shape (nx, nz), iz fastest, float32 --
exactly what the model files hold (nx records of nz samples,
src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_bond_transformation.c:345-355).
"""
import numpy as np


def ricker(fpeak, dt, nt, A=1.0, delay=None):
    """Ricker wavelet with the semantics of riker_wlt (src/lib/wavelet/wavelet.c:81-99):
    t = it*dt - delay, zero once t > delay, first sample forced to zero."""
    if delay is None:
        delay = np.float32(1.5) / np.float32(fpeak)
    pi = np.float32(3.141592654)
    t = np.arange(nt, dtype=np.float32) * np.float32(dt) - np.float32(delay)
    a = (pi * np.float32(fpeak) * t) ** 2
    w = np.float32(A) * (np.float32(1) - np.float32(2) * a) * np.exp(-a)
    w = np.where(t <= np.float32(delay), w, np.float32(0)).astype(np.float32)
    w[0] = 0.0
    return w


def _gardner(vp):
    return (310.0 * np.power(vp, 0.25)).astype(np.float32)


def model_small(nz=200, nx=300, dx=10.0, seed=1):
    """Small layered TTI model: vp rising linearly with depth 1800->3500 m/s, cut by 3
    dipping planar interfaces with +-200 m/s jumps; vs = vp/1.73; rho = 310 vp^0.25;
    eps ~ U(0.10, 0.20) and delta ~ U(0.05, 0.10) per layer; theta = 30 deg."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx, dtype=np.float64)[:, None] * dx
    z = np.arange(nz, dtype=np.float64)[None, :] * dx
    zmax = (nz - 1) * dx
    vp = 1800.0 + 1700.0 * z / zmax + 0.0 * x
    layer = np.zeros((nx, nz), dtype=np.int64)
    for k in range(3):
        z0 = (k + 1) * zmax / 4.0 + rng.uniform(-0.05, 0.05) * zmax
        slope = np.tan(np.deg2rad(rng.uniform(-15.0, 15.0)))
        jump = rng.choice([-200.0, 200.0])
        below = z > (z0 + slope * (x - 0.5 * (nx - 1) * dx))
        vp = vp + jump * below
        layer += below
    eps_l = rng.uniform(0.10, 0.20, size=4)
    del_l = rng.uniform(0.05, 0.10, size=4)
    vp = vp.astype(np.float32)
    return dict(vp=vp, vs=(vp / np.float32(1.73)).astype(np.float32), rho=_gardner(vp),
                eps=eps_l[layer].astype(np.float32), delta=del_l[layer].astype(np.float32),
                theta=np.full((nx, nz), 30.0, dtype=np.float32))


def model_large(nz=1126, nx=7872, dx=10.0, seed=2, water=300.0):
    """Large marine TTI model: smooth vp 1500-4800 m/s under a water layer (vs=0,
    eps=delta=0), 8 folded interfaces, one salt-like body (vp 4500, vs 2600);
    eps 0-0.25 and delta 0-0.12 rising with depth; theta = interface dip clipped to
    +-45 deg and smoothed with a Gaussian of sigma 50 m."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    xs = np.arange(nx, dtype=np.float64) * dx
    zs = np.arange(nz, dtype=np.float64) * dx
    Z = np.broadcast_to(zs[None, :], (nx, nz))
    zmax = zs[-1]
    iw = zs < water
    # folded interfaces z_k(x) and their dips
    nif = 8
    base = np.linspace(water + 600.0, zmax - 800.0, nif)
    zi = np.empty((nif, nx))
    dip = np.empty((nif, nx))
    for k in range(nif):
        amp = rng.uniform(150.0, 600.0)
        lam = rng.uniform(2500.0, 12000.0)
        ph = rng.uniform(0, 2 * np.pi)
        amp2 = rng.uniform(20.0, 150.0)
        lam2 = rng.uniform(1500.0, 5000.0)
        ph2 = rng.uniform(0, 2 * np.pi)
        zi[k] = (base[k] + amp * np.sin(2 * np.pi * xs / lam + ph) + amp2 * np.sin(2 * np.pi * xs / lam2 + ph2))
        dzdx = (amp * 2 * np.pi / lam * np.cos(2 * np.pi * xs / lam + ph)
                + amp2 * 2 * np.pi / lam2 * np.cos(2 * np.pi * xs / lam2 + ph2))
        dip[k] = np.rad2deg(np.arctan(dzdx))
    zi = np.sort(zi, axis=0)
    layer = np.zeros((nx, nz), dtype=np.int64)
    for k in range(nif):
        layer += (Z > zi[k][:, None])
    frac = np.clip((Z - water) / (zmax - water), 0.0, 1.0)
    vp = 1800.0 + 2800.0 * frac ** 0.8
    jumps = rng.uniform(-150.0, 250.0, size=nif + 1)
    jumps[0] = 0.0
    vp = vp + np.cumsum(jumps)[layer]
    vp = np.clip(vp, 1600.0, 4800.0)
    vs = vp / 1.73
    eps = np.clip(0.25 * frac + rng.uniform(-0.03, 0.03, size=nif + 1)[layer], 0.0, 0.25)
    dlt = np.clip(0.12 * frac + rng.uniform(-0.02, 0.02, size=nif + 1)[layer], 0.0, 0.12)
    # theta: dip of the interface just above each point (layer 0 takes the first one)
    th = dip[np.maximum(layer - 1, 0), np.arange(nx)[:, None]]
    # salt body: an ellipse in the middle third of the section
    cx, cz = xs[nx // 2] + rng.uniform(-0.1, 0.1) * xs[-1], water + 0.35 * (zmax - water)
    ax, az = 0.12 * xs[-1], 0.12 * (zmax - water)
    salt = ((xs[:, None] - cx) / ax) ** 2 + ((zs[None, :] - cz) / az) ** 2 < 1.0
    vp = np.where(salt, 4500.0, vp)
    vs = np.where(salt, 2600.0, vs)
    eps = np.where(salt, 0.0, eps)
    dlt = np.where(salt, 0.0, dlt)
    # water layer
    vp[:, iw] = 1500.0
    vs[:, iw] = 0.0
    eps[:, iw] = 0.0
    dlt[:, iw] = 0.0
    th[:, iw] = 0.0
    th = np.clip(th, -45.0, 45.0)
    th = gaussian_filter(th, sigma=50.0 / dx, mode="nearest")
    th[:, iw] = 0.0
    rho = np.where(iw[None, :], 1000.0, 310.0 * np.power(vp, 0.25))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(vp=f32(vp), vs=f32(vs), rho=f32(rho), eps=f32(eps), delta=f32(dlt), theta=f32(th))


def add_anomalies(model, dx=10.0, n=5, sigma=150.0, seed=4, dvp=0.08, deps=0.05, xrange=(0.15, 0.85),
                  zrange=(0.07, 0.25)):
    """True model for gradient tests: background + n Gaussian anomalies (sigma in m,
    +-dvp relative in vp, +-deps absolute in epsilon) in the rock below the water layer,
    centred at seeded positions within the fractions xrange of the width and zrange of
    the depth.  The default depths (0.8-2.8 km on the 1126-deep model) are ones whose
    reflections reach the surface receivers within nt = 4000 steps of 0.8 ms; anomalies
    deeper than about 3.4 km leave those records unchanged to FP32 rounding."""
    rng = np.random.default_rng(seed)
    nx, nz = model["vp"].shape
    X = np.arange(nx)[:, None] * dx
    Z = np.arange(nz)[None, :] * dx
    g_vp = np.zeros((nx, nz))
    g_ep = np.zeros((nx, nz))
    rock = model["vs"] > 0
    for _ in range(n):
        cx = rng.uniform(*xrange) * (nx - 1) * dx
        cz = rng.uniform(*zrange) * (nz - 1) * dx
        g = np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * sigma ** 2))
        g_vp += rng.choice([-1.0, 1.0]) * dvp * g
        g_ep += rng.choice([-1.0, 1.0]) * deps * g
    out = {k: v.copy() for k, v in model.items()}
    out["vp"] = np.where(rock, model["vp"] * (1.0 + g_vp), model["vp"]).astype(np.float32)
    out["eps"] = np.where(rock, np.clip(model["eps"] + g_ep, 0.0, 0.5), model["eps"]).astype(np.float32)
    return out


def crop(model, x0, w):
    """Aperture crop: columns [x0, x0+w) of every model array (cutbin2d's role,
    src/main/wavelet/cutbin2d.c:30-94)."""
    return {k: np.ascontiguousarray(v[x0:x0 + w]) for k, v in model.items()}


# Named configurations, sizes as used by the tests and the benchmark.
CONFIGS = {
    "small": dict(nz=200, nx=300, dx=10.0, dz=10.0, dt=1e-3, nt=1000, ealThick=50, freesurface=0,
                  fpeak=15.0, sz=2, sx=150, rz=2),
    "large": dict(nz=1126, nx=7872, dx=10.0, dz=10.0, dt=8e-4, nt=8000, ealThick=60, freesurface=1,
                  fpeak=12.0, sz=2, sx=3936, rz=2),
}
