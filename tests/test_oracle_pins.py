"""Pinning the CPU oracle against mathematics that does not restate its code.

The reference has no tests or golden vectors and cannot be built here (it needs
Seismic Unix's cwp.h/par.h/su.h and libraries), so each oracle function is pinned
by an independent check: a numpy formulation, a closed form, a hand value, or an
invariant that a plausible mistake (dropped term, wrong sign or index, swapped
operands) breaks.
"""
import numpy as np
import pytest

from common import rel_l2, run_case, small_case
from oracle import Oracle

C = np.array([1.24362739770907, -0.114021164233934, 0.0272965993810630, -0.00699199493927320,
              0.00132022981661381], dtype=np.float32).astype(np.float64)
GAIN = float(np.sum((2 * np.arange(1, 6) - 1) * C))  # sum (2k-1) C_k


def _rand_model(nz, nx, seed=0, water=0, theta=(-60, 60)):
    rng = np.random.default_rng(seed)
    vp = rng.uniform(2000, 3000, (nx, nz)).astype(np.float32)
    vs = (vp / rng.uniform(1.6, 2.0, (nx, nz))).astype(np.float32)
    rho = rng.uniform(2000, 2500, (nx, nz)).astype(np.float32)
    eps = rng.uniform(0.05, 0.2, (nx, nz)).astype(np.float32)
    de = rng.uniform(0.0, 0.1, (nx, nz)).astype(np.float32)
    th = rng.uniform(theta[0], theta[1], (nx, nz)).astype(np.float32)
    if water:
        vp[:, :water], vs[:, :water], eps[:, :water], de[:, :water] = 1500, 0, 0, 0
    return vp, vs, rho, eps, de, th


def test_gain_value():
    # the LS coefficient set is not exactly normalised: sum (2k-1) C_k = 1.000985...
    assert abs(GAIN - 1.000985005687) < 1e-7  # float32-rounded coefficients


# ---------------------------------------------------------------- padding
def test_pad_matches_numpy_edge_padding():
    nz, nx = 7, 9
    m = _rand_model(nz, nx, seed=1)
    o = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=3, prec="f32")
    o.set_model(*m)
    # eps and delta are not modified by the setup, so they show the padding alone
    for name, a in (("eps", m[3]), ("delta", m[4])):
        assert np.array_equal(o.get(name).astype(np.float32), np.pad(a, o.e, mode="edge"))


# ---------------------------------------------------------------- stability
def test_stability_limit_hand_value():
    # dt_max = 1/(vp sqrt((1/dx^2+1/dz^2)(C1+C3+C5))): 1.306 ms at vp = 4800, dx = dz = 10
    nz, nx = 6, 6
    vp = np.full((nx, nz), 4800.0, np.float32)
    ones = np.ones((nx, nz), np.float32)
    o = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1.0e-3, ealThick=2)
    o.set_model(vp, vp / 2, 2000 * ones, 0 * ones, 0 * ones, 0 * ones)
    assert abs(o.dt_max - 1.30604e-3) < 2e-8
    o2 = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1.31e-3, ealThick=2)
    with pytest.raises(ValueError):
        o2.set_model(vp, vp / 2, 2000 * ones, 0 * ones, 0 * ones, 0 * ones)


# ---------------------------------------------------------------- MEAL
@pytest.mark.parametrize("ealThick,R", [(5, 1e-4), (15, 1e-5)])
def test_meal_profile(ealThick, R):
    """n = ealThick+5 layers: R = 10^-((log10 n - 1)/log10 2 + 3 + 0.5*2), so n=10 -> 1e-4 and
    n=20 -> 1e-5.  At the outermost layer w = 2^1 - 1 = 1, so D*Beta = 3 vp ln(1/R) / (2 n h)."""
    nz, nx, dz, vpv, vsv, fpeak = 12, 14, 10.0, 2500.0, 1200.0, 15.0
    ones = np.ones((nx, nz), np.float32)
    o = Oracle("coord", nz, nx, 4, dz, dz, 1e-3, ealThick=ealThick, fpeak=fpeak, prec="f64")
    o.set_model(vpv * ones, vsv * ones, 2000 * ones, 0 * ones, 0 * ones, 0 * ones)
    n = o.e
    D, B = o.get("D"), o.get("Beta")
    mid = o.Nx // 2
    d0 = 3.0 * vpv * np.log(1.0 / R) / (2 * n * dz)
    assert abs(D[mid, 0] * B[mid, 0] / d0 - 1) < 1e-5              # top edge, one side only
    assert abs(D[mid, o.Nz - 1] * B[mid, o.Nz - 1] / d0 - 1) < 1e-5  # bottom edge
    # Beta at the edge: 1 + |2 vs/(h fpeak)/10 - 1|
    assert abs(B[mid, 0] - (1 + abs(2 * vsv / (dz * fpeak) / 10 - 1))) < 1e-6
    # corner: both sides equal, D*Beta = hypot -> sqrt(2) d0 and Beta-1 -> sqrt(2) (beta-1)
    assert abs(D[0, 0] * B[0, 0] / (np.sqrt(2) * d0) - 1) < 1e-5
    # interior exactly undamped; profile increases toward the edge
    assert np.all(D[n:-n, n:-n] == 0) and np.all(B[n:-n, n:-n] == 1)
    col = D[mid, :n] * B[mid, :n]
    assert np.all(np.diff(col) < 0)
    c1, c2 = o.get("coef1"), o.get("coef2")
    assert np.all(c1[n:-n, n:-n] == 1.0) and np.all(c2[n:-n, n:-n] == np.float64(np.float32(1e-3)))


@pytest.mark.parametrize("eal_alpha,alpha", [(-1.0, 2.5), (1.5, 1.5)])
def test_meal_inner_layers_closed_form(eal_alpha, alpha):
    """Every layer l = 1..n of the profile (meal.c:23-29, 74-77; n = ealThick + 5): on a side
    column at mid depth D*Beta = d0x w(l) and Beta - 1 = |beta_x - 1| w(l), on the top row
    at mid width the same with dz, where w(l) = 2^((l/n)^alpha) - 1 (alpha = 2.5 for the
    default eal_alpha = -1), d0 = 3 vp ln(1/R) / (2 n h), beta = 2 vs / (h fpeak) / 10.
    Layer l of the left side is column n - l, of the top row is row n - l."""
    nz, nx, dx, dz, vpv, vsv, fpeak = 24, 26, 8.0, 12.5, 2500.0, 1200.0, 15.0
    ones = np.ones((nx, nz), np.float32)
    o = Oracle("bond", nz, nx, 4, dx, dz, 1e-3, ealThick=15, eal_alpha=eal_alpha, fpeak=fpeak, prec="f64")
    o.set_model(vpv * ones, vsv * ones, 2000 * ones, 0 * ones, 0 * ones, 0 * ones)
    n = o.e
    assert n == 20
    R = 1e-5  # 10^-((log10 20 - 1)/log10 2 + 3 + 1)
    D, B = o.get("D"), o.get("Beta")
    for l in range(1, n + 1):
        w = 2.0 ** ((l / n) ** alpha) - 1.0
        for h, (ix, iz) in ((dx, (n - l, o.Nz // 2)), (dz, (o.Nx // 2, n - l))):
            d0 = 3.0 * vpv * np.log(1.0 / R) / (2 * n * h)
            beta = 2.0 * vsv / (h * fpeak) / 10.0
            assert abs(D[ix, iz] * B[ix, iz] - d0 * w) <= 1e-9 * d0, (l, h)
            assert abs(B[ix, iz] - 1.0 - abs(beta - 1.0) * w) <= 1e-9, (l, h)


def test_meal_free_surface_leaves_top_undamped():
    nz, nx = 12, 14
    ones = np.ones((nx, nz), np.float32)
    o = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=4, freesurface=1)
    o.set_model(2500 * ones, 1200 * ones, 2000 * ones, 0 * ones, 0 * ones, 0 * ones)
    D = o.get("D")
    n = o.e
    assert np.all(D[n:-n, :n] == 0)      # no top layer
    assert np.all(D[:n, :n] == 0)        # side layers start at izStart
    assert np.all(D[:n, n:] > 0)


def test_interior_coefs_bitwise_f32():
    ctor, m, geom = small_case(nz=30, nx=40, nt=4)
    o = run_case(Oracle, ctor, m, geom)
    n = o.e
    c1, c2 = o.get("coef1").astype(np.float32), o.get("coef2").astype(np.float32)
    assert np.all(c1[n:-n, n:-n] == np.float32(1.0))
    assert np.all(c2[n:-n, n:-n] == np.float32(1e-3))


# ---------------------------------------------------------------- Cij
def _rotate_vti(c, s, C11, C13, C33, C55):
    """Rotate the 2-D VTI stiffness tensor with Q = [[c, s], [-s, c]], the rotation the
    tensor-rotation form applies to stress (sigma = Q tau Q^T), and read Voigt terms."""
    T = np.zeros(C11.shape + (2, 2, 2, 2))
    T[..., 0, 0, 0, 0] = C11
    T[..., 1, 1, 1, 1] = C33
    T[..., 0, 0, 1, 1] = C13
    T[..., 1, 1, 0, 0] = C13
    for ijkl in [(0, 1, 0, 1), (1, 0, 0, 1), (0, 1, 1, 0), (1, 0, 1, 0)]:
        T[(Ellipsis,) + ijkl] = C55
    Q = np.stack([np.stack([c, s], -1), np.stack([-s, c], -1)], -2)
    R = np.einsum("...ia,...jb,...kc,...ld,...abcd->...ijkl", Q, Q, Q, Q, T)
    return dict(c11=R[..., 0, 0, 0, 0], c13=R[..., 0, 0, 1, 1], c15=R[..., 0, 0, 0, 1],
                c33=R[..., 1, 1, 1, 1], c35=R[..., 1, 1, 0, 1], c55=R[..., 0, 1, 0, 1])


def test_bond_cij_is_rotated_tensor():
    nz, nx = 8, 9
    m = _rand_model(nz, nx, seed=3)
    ob = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    ot = Oracle("tensor", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    ob.set_model(*m)
    ot.set_model(*m)
    th = ot.get("theta")
    r = _rotate_vti(np.cos(th), np.sin(th), *(ot.get(k) for k in ("c11", "c13", "c33", "c55")))
    scale = np.abs(ot.get("c11")).max()
    for k, v in r.items():
        assert np.abs(ob.get(k) - v).max() / scale < 1e-12, k


def test_thomsen_inversion():
    """Thomsen: eps = (C11-C33)/(2 C33), delta = ((C13+C55)^2-(C33-C55)^2)/(2 C33 (C33-C55))."""
    nz, nx = 8, 9
    m = _rand_model(nz, nx, seed=4)
    o = Oracle("tensor", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    o.set_model(*m)
    c11, c13, c33, c55 = (o.get(k) for k in ("c11", "c13", "c33", "c55"))
    e = o.e
    eps = (c11 - c33) / (2 * c33)
    de = ((c13 + c55) ** 2 - (c33 - c55) ** 2) / (2 * c33 * (c33 - c55))
    assert np.abs(eps[e:-e, e:-e] - m[3]).max() < 1e-6
    assert np.abs(de[e:-e, e:-e] - m[4]).max() < 1e-6
    # C33 = rho vp^2 and C55 = rho vs^2 after MEAL (interior Beta = 1)
    assert np.allclose(c33[e:-e, e:-e], m[2].astype(np.float64) * m[0].astype(np.float64) ** 2, rtol=1e-6)


def test_theta_zero_bond_is_vti():
    nz, nx = 6, 7
    m = list(_rand_model(nz, nx, seed=5))
    m[5] = np.zeros_like(m[5])
    ob = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    ot = Oracle("tensor", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    ob.set_model(*m)
    ot.set_model(*m)
    for k in ("c11", "c13", "c33", "c55"):
        assert np.array_equal(ob.get(k), ot.get(k))
    assert np.all(ob.get("c15") == 0) and np.all(ob.get("c35") == 0)


def test_theta_90_swaps_c11_c33():
    nz, nx = 6, 7
    m = list(_rand_model(nz, nx, seed=6))
    m[5] = np.full_like(m[5], 90.0)
    ob = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    ot = Oracle("tensor", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=0, prec="f64")
    ob.set_model(*m)
    ot.set_model(*m)
    assert np.allclose(ob.get("c11"), ot.get("c33"), rtol=1e-12)
    assert np.allclose(ob.get("c33"), ot.get("c11"), rtol=1e-12)


# ---------------------------------------------------------------- staggered parameters
def test_invrho_and_trig_averages_numpy():
    nz, nx = 7, 8
    m = _rand_model(nz, nx, seed=7)
    o = Oracle("coord", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=1)
    o.set_model(*m)
    r = o.get("rho").astype(np.float32)
    s = o.get("sin").astype(np.float32)
    c = o.get("cos").astype(np.float32)
    ir = (1.0 / r.astype(np.float64)).astype(np.float32)
    ssum = r[:-1, :-1] + r[1:, :-1] + r[:-1, 1:] + r[1:, 1:]
    ir[:-1, :-1] = (4.0 / ssum.astype(np.float64)).astype(np.float32)
    assert np.array_equal(o.get("invrho").astype(np.float32), ir)
    for arr, name in ((s, "sinv"), (c, "cosv")):
        exp = arr.copy()
        exp[:-1, :-1] = (0.25 * (arr[:-1, :-1] + arr[1:, :-1] + arr[:-1, 1:] + arr[1:, 1:]).astype(np.float64))
        assert np.array_equal(o.get(name).astype(np.float32), exp.astype(np.float32)), name
    th = o.get("theta")
    assert np.allclose(s, np.sin(th), atol=1e-7) and np.allclose(c, np.cos(th), atol=1e-7)


# ---------------------------------------------------------------- Bond free-surface fix-up
def test_bond_fixup_vti_hand_case_and_water_nan():
    """theta=0, vs>0: c15 = c35 = 0, so the reference's expressions give
    c35' = -c13/c33 and c11' = c11 + c13^2/c33; the other five become 0.  In water at
    theta = 0 (c55 = 0) the fix-up divides 0/0 and c11' is NaN (kept, as in the reference)."""
    nz, nx = 10, 9
    m = list(_rand_model(nz, nx, seed=8, water=3))
    m[5] = np.zeros_like(m[5])
    ob = Oracle("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=2, freesurface=1, prec="f64")
    ot = Oracle("tensor", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=2, freesurface=1, prec="f64")
    ob.set_model(*m)
    ot.set_model(*m)
    e = ob.e
    row = ob.get("c11")[:, e]
    assert np.isnan(row).all()  # the surface row lies in the water
    m2 = list(_rand_model(nz, nx, seed=8))
    m2[5] = np.zeros_like(m2[5])
    ob.__init__("bond", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=2, freesurface=1, prec="f64")
    ot.__init__("tensor", nz, nx, 4, 10.0, 10.0, 1e-3, ealThick=2, freesurface=1, prec="f64")
    ob.set_model(*m2)
    ot.set_model(*m2)
    c11, c13, c33 = (ot.get(k)[:, e] for k in ("c11", "c13", "c33"))
    assert np.allclose(ob.get("c11")[:, e], c11 + c13 ** 2 / c33, rtol=1e-12)
    for k in ("c13", "c15", "c33", "c35", "c55"):
        assert np.all(ob.get(k)[:, e] == 0), k


# ---------------------------------------------------------------- RSG derivatives
def _sweep_setup(form="bond", nz=40, nx=44, dx=10.0, dz=10.0):
    m = _rand_model(nz, nx, seed=9, theta=(0, 0))
    o = Oracle(form, nz, nx, 4, dx, dz, 1e-3, ealThick=2, prec="f64")
    o.set_model(*m)
    for f in ("vx", "vz", "sxx", "szz", "sxz"):
        o.set(f, np.zeros((o.Nx, o.Nz)))
    return o


def _interior(o):
    return (slice(o.ixStart, o.ixEnd), slice(o.izStart, o.izEnd))


@pytest.mark.parametrize("dx,dz", [(10.0, 10.0), (8.0, 12.5)])
def test_rsg_exact_on_linear_fields(dx, dz):
    """On f = a x + b z the staggered pair gives d/dx = GAIN*a and d/dz = GAIN*b."""
    o = _sweep_setup(dx=dx, dz=dz)
    X = np.arange(o.Nx)[:, None] * dx
    Z = np.arange(o.Nz)[None, :] * dz
    a, b = 0.7, -1.3
    o.set("sxx", a * X + b * Z + 0 * Z)   # vx <- d(sxx)/dx
    o.set("sxz", -b * X + a * Z + 0 * Z)  # vx += d(sxz)/dz, vz <- d(sxz)/dx
    o.v_sweep()
    k = o.get("coef2") * o.get("invrho")
    sl = _interior(o)
    assert np.allclose((o.get("vx") / k)[sl], GAIN * (a + a), rtol=1e-10)
    assert np.allclose((o.get("vz") / k)[sl], GAIN * (-b), rtol=1e-10)


def test_coordinate_form_dz_of_sxx_carries_invdx():
    """CR:633: the coordinate form's d(sxx)/dz is scaled by invdx, not invdz.  With the
    velocity-point trig set to theta = 90 degrees (cosv = 0, sinv = 1) the velocity update
    is vx = k (-sxxdz + sxzdx): on sxx = b z (sxz = 0) the diagonal pair gives
    0.5 (d1 - d2) = GAIN b dz, so vx / k = -GAIN b dz / dx (-GAIN b with invdz) -- at
    dx = 8, dz = 12.5 the two differ by 56%."""
    dx, dz = 8.0, 12.5
    o = _sweep_setup(form="coord", dx=dx, dz=dz)
    Z = np.arange(o.Nz)[None, :] * dz + 0.0 * np.arange(o.Nx)[:, None]
    b = -1.3
    o.set("cosv", np.zeros((o.Nx, o.Nz)))
    o.set("sinv", np.ones((o.Nx, o.Nz)))
    o.set("sxx", b * Z)
    o.v_sweep()
    k = o.get("coef2") * o.get("invrho")
    sl = _interior(o)
    assert np.allclose((o.get("vx") / k)[sl], -GAIN * b * dz / dx, rtol=1e-10)
    assert np.allclose(o.get("vz")[sl], 0.0, atol=1e-12)


def _symbol(kappa, h):
    return (2.0 / h) * np.sum(C * np.sin((2 * np.arange(1, 6) - 1) * kappa * h / 2))


def test_rsg_plane_wave_symbol_and_staggering():
    """Velocities live at (x+dx/2, z+dz/2) relative to stresses.  For a plane wave
    sin(kx x + kz z) the discrete derivative equals the closed-form diagonal symbol
    evaluated at the staggered position (exactly, up to rounding)."""
    dx = dz = 10.0
    o = _sweep_setup(dx=dx, dz=dz)
    X = np.arange(o.Nx)[:, None] * dx
    Z = np.arange(o.Nz)[None, :] * dz
    kx, kz = 2 * np.pi / 97.0, 2 * np.pi / 143.0
    o.set("sxx", np.sin(kx * X + kz * Z))
    o.v_sweep()
    k = o.get("coef2") * o.get("invrho")
    got = (o.get("vx") / k)[_interior(o)]
    # the diagonal stencils see u = x+z and w = x-z directions with spacing sqrt(2)h;
    # in terms of (kx, kz): d/dx = 0.5*(S(ku) + S(kw)) style combination -> evaluate directly
    h = dx
    ku, kw = kx + kz, kx - kz
    Xs, Zs = (X + dx / 2)[_interior(o)[0]], (Z + dz / 2)[:, _interior(o)[1]]
    ph = kx * Xs + kz * Zs
    exp = 0.5 * (_symbol(ku, h) + _symbol(kw, h)) * np.cos(ph) * h / dx
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-12)


def test_rsg_velocity_to_stress_staggering():
    dx = dz = 10.0
    o = _sweep_setup(dx=dx, dz=dz)
    X = np.arange(o.Nx)[:, None] * dx
    Z = np.arange(o.Nz)[None, :] * dz
    kx, kz = 2 * np.pi / 89.0, 2 * np.pi / 131.0
    # velocity samples sit at +h/2; zero all stiffness except c11 so sxx = dt*c11*dvx/dx
    o.set("vx", np.sin(kx * (X + dx / 2) + kz * (Z + dz / 2)))
    for name in ("c13", "c15", "c35", "c33", "c55"):
        o.set(name, np.zeros((o.Nx, o.Nz)))
    o.s_sweep()
    sl = _interior(o)
    got = (o.get("sxx") / (o.get("coef2") * o.get("c11")))[sl]
    ku, kw = kx + kz, kx - kz
    exp = 0.5 * (_symbol(ku, dx) + _symbol(kw, dx)) * np.cos(kx * X + kz * Z)[sl]
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- sources
def test_source_weights_hand_values():
    """w = exp(alpha r^2) dt/(dx dz), alpha = ln(0.8)/(dx^2+dz^2): 1 at the centre, 0.8 on the
    diagonal neighbour, 0.8^0.5 on an axis neighbour (dx = dz)."""
    ctor, m, geom = small_case(nz=30, nx=30, nt=4, stype=5, ealThick=2)
    geom["src"] = np.ones_like(geom["src"])
    o = run_case(Oracle, {**ctor}, m, geom)
    o.sources(0, 0)
    vx = o.get("vx")
    ix, iz = geom["spx"][0] + o.e, geom["spz"][0] + o.e
    w0 = 1e-3 / 100.0
    assert abs(vx[ix, iz] - w0) / w0 < 1e-6
    assert abs(vx[ix + 1, iz + 1] - 0.8 * w0) / w0 < 1e-6
    assert abs(vx[ix + 1, iz] - np.sqrt(0.8) * w0) / w0 < 1e-6
    assert np.count_nonzero(vx) == 121


@pytest.mark.parametrize("stype", [5, 6, 7])
def test_force_sources_hand_values_dx_ne_dz(stype):
    """stype 5/6/7 (BT:740-812): w = exp(alpha (rx^2 + rz^2)) dt/(dx dz) with
    alpha = ln(0.8)/(dx^2 + dz^2), so with dx = 8, dz = 12.5 the diagonal neighbour gets
    0.8 w0 and the axis neighbours 0.8^(dx^2/(dx^2+dz^2)) w0 (x) and 0.8^(dz^2/(dx^2+dz^2))
    w0 (z); 5 loads vx, 6 loads vz, 7 loads both with sqrt(2)/2."""
    dx, dz = 8.0, 12.5
    ctor, m, geom = small_case(nz=30, nx=30, nt=4, stype=stype, ealThick=2, dx=dx, dz=dz)
    geom["src"] = np.full_like(geom["src"], 3.0)
    o = run_case(Oracle, ctor, m, geom, prec="f64")
    o.sources(0, 0)
    vx, vz = o.get("vx"), o.get("vz")
    ix, iz = geom["spx"][0] + o.e, geom["spz"][0] + o.e
    w0 = 3.0 * float(np.float32(1e-3)) / (dx * dz)
    q = float(np.float32(0.8))  # logf(0.8f), BT:647
    gx, gz = q ** (dx * dx / (dx * dx + dz * dz)), q ** (dz * dz / (dx * dx + dz * dz))
    amp = {5: (1.0, 0.0), 6: (0.0, 1.0), 7: (np.sqrt(0.5), np.sqrt(0.5))}[stype]
    for f, a in ((vx, amp[0]), (vz, amp[1])):
        if a == 0.0:
            assert np.count_nonzero(f) == 0
            continue
        assert np.count_nonzero(f) == 121
        for (ox, oz), fac in (((0, 0), 1.0), ((1, 1), q), ((-1, 1), q), ((1, 0), gx), ((0, -1), gz),
                              ((5, 5), q ** 25), ((-5, 2), gx ** 25 * gz ** 4)):
            assert abs(f[ix + ox, iz + oz] - a * fac * w0) <= 1e-10 * w0, (stype, ox, oz)


def test_source_curl_and_divergence_symmetry():
    ctor, m, geom = small_case(nz=30, nx=30, nt=4, stype=2, ealThick=2)
    geom["src"] = np.ones_like(geom["src"])
    o = run_case(Oracle, ctor, m, geom)
    o.sources(0, 0)
    ix, iz = geom["spx"][0] + o.e, geom["spz"][0] + o.e
    vx, vz = o.get("vx"), o.get("vz")
    blk = (slice(ix - 5, ix + 6), slice(iz - 5, iz + 6))
    # divergence source: vx odd in x and even in z; vz = transpose pattern
    assert np.allclose(vx[blk], -vx[blk][::-1, :]) and np.allclose(vx[blk], vx[blk][:, ::-1])
    assert np.allclose(vz[blk], vx[blk].T)
    ctor3, m3, geom3 = small_case(nz=30, nx=30, nt=4, stype=3, ealThick=2)
    geom3["src"] = np.ones_like(geom3["src"])
    o3 = run_case(Oracle, ctor3, m3, geom3)
    o3.sources(0, 0)
    # curl source is the divergence pattern rotated by 90 degrees: (vx, vz) -> (-vz, vx)
    assert np.allclose(o3.get("vx")[blk], -vz[blk]) and np.allclose(o3.get("vz")[blk], vx[blk])


def test_source_stype8_four_points():
    ctor, m, geom = small_case(nz=30, nx=30, nt=4, stype=8, ealThick=2)
    geom["src"] = np.full_like(geom["src"], 2.0)
    o = run_case(Oracle, ctor, m, geom)
    o.sources(0, 0)
    ix, iz = geom["spx"][0] + o.e, geom["spz"][0] + o.e
    q = 0.125 * np.sqrt(2) * 2.0
    vx, vz = o.get("vx"), o.get("vz")
    exp = {(0, 0): (-q, q), (-1, 0): (-q, -q), (-1, -1): (q, -q), (0, -1): (q, q)}
    for (dxo, dzo), (ex, ez) in exp.items():
        assert abs(vx[ix + dxo, iz + dzo] - ex) < 1e-6 and abs(vz[ix + dxo, iz + dzo] - ez) < 1e-6
    assert np.count_nonzero(vx) == 4


# ---------------------------------------------------------------- whole loop
def test_linearity_exact_power_of_two():
    ctor, m, geom = small_case(nz=40, nx=50, nt=120)
    a = run_case(Oracle, ctor, m, geom)
    a.run()
    geom2 = dict(geom, src=geom["src"] * 4.0)
    b = run_case(Oracle, ctor, m, geom2)
    b.run()
    for comp in (1, 3, 4):
        # exact except where values pass through the subnormal range (wavefront tails)
        assert rel_l2(4.0 * a.traces(comp), b.traces(comp)) < 1e-7


def test_determinism():
    ctor, m, geom = small_case(nz=40, nx=50, nt=120, form="tensor", freesurface=1)
    a = run_case(Oracle, ctor, m, geom)
    b = run_case(Oracle, ctor, m, geom)
    a.run()
    b.run()
    assert np.array_equal(a.traces(4), b.traces(4))


def test_free_surface_invariants():
    ctor, m, geom = small_case(nz=40, nx=50, nt=60, form="bond", freesurface=1, sz=6)
    o = run_case(Oracle, ctor, m, geom)
    o.run(0, 40)
    o.free_surface()  # idempotent on a state that already satisfies it
    z0 = o.izStart
    sxx, szz, sxz = (o.get(k).astype(np.float32) for k in ("sxx", "szz", "sxz"))
    cols = slice(o.ixStart, o.ixEnd)
    assert np.all(szz[cols, z0] == 0) and np.all(sxz[cols, z0] == 0)
    assert np.abs(sxx[cols, z0]).max() > 0
    for k in range(1, 6):
        assert np.array_equal(szz[cols, z0 - k], -szz[cols, z0 + k])
        assert np.array_equal(sxz[cols, z0 - k], -sxz[cols, z0 + k])
        assert np.array_equal(sxx[cols, z0 - k], np.float32(2) * sxx[cols, z0] - sxx[cols, z0 + k])


def test_receivers_sample_fields():
    ctor, m, geom = small_case(nz=40, nx=50, nt=30, seistype=3)
    o = run_case(Oracle, ctor, m, geom)
    o.run()
    e = o.e
    gx, gz = geom["rpx"] + e, geom["rpz"] + e
    p = 0.5 * (o.get("sxx") + o.get("szz")).astype(np.float32)
    assert np.array_equal(o.traces(4)[:, -1], (np.float32(0.5) * (o.get("sxx")[gx, gz].astype(np.float32)
                                                                 + o.get("szz")[gx, gz].astype(np.float32))))
    assert np.array_equal(o.traces(1)[:, -1], o.get("vx")[gx, gz])
    assert np.array_equal(o.traces(3)[:, -1], o.get("vz")[gx, gz])
    assert p.shape == (o.Nx, o.Nz)


@pytest.mark.parametrize("theta", [0.0, 30.0, -55.0])
def test_cross_form_pressure_agrees_at_constant_theta(theta):
    """At constant theta, with dx = dz and no free surface, the three forms are the same
    discrete system in rotated variables, and p = (sxx+szz)/2 is rotation invariant."""
    res = {}
    for form in ("bond", "tensor", "coord"):
        ctor, m, geom = small_case(nz=50, nx=60, nt=250, form=form, theta=theta, sz=20)
        o = run_case(Oracle, ctor, m, geom, prec="f64")
        o.run()
        res[form] = o.traces(4)
    assert rel_l2(res["tensor"], res["bond"]) < 1e-11
    assert rel_l2(res["coord"], res["bond"]) < 1e-11


def test_fp32_vs_fp64_oracle_error_small():
    ctor, m, geom = small_case(nz=50, nx=60, nt=300, form="bond")
    a = run_case(Oracle, ctor, m, geom, prec="f32")
    b = run_case(Oracle, ctor, m, geom, prec="f64")
    a.run()
    b.run()
    assert rel_l2(a.traces(4), b.traces(4)) < 1e-5


def test_acoustic_limit_matches_2d_greens_function():
    """vs = 0, eps = delta = 0 turns the system into the acoustic wave equation, and the
    stype-1 source injects dp/dt = 0.5 src(t) g(x)/(dx dz) over the 11x11 Gaussian footprint
    g = 0.8^(r^2/(dx^2+dz^2)).  Then p = d/dt sum_i 0.5 src * g_i G2(|x - x_i|, t) with the
    2-D Green's function G2 = H(ct - r)/(2 pi c sqrt(c^2 t^2 - r^2)), evaluated in frequency
    as (-i/4) H0^(2)(w r/c)/c^2.  Waveform and amplitude must agree to dispersion error."""
    from scipy.special import hankel2

    from fdmod_hip import synth

    nz, nx, h, dt, nt, c, rho = 160, 160, 10.0, 1e-3, 700, 2000.0, 1000.0
    ones = np.ones((nx, nz), np.float32)
    o = Oracle("bond", nz, nx, nt, h, h, dt, ealThick=40, fpeak=15.0, prec="f64")
    o.set_model(c * ones, 0 * ones, rho * ones, 0 * ones, 0 * ones, 0 * ones)
    src = synth.ricker(15.0, dt, nt)
    sx, sz = 40, 80
    o.set_source(1, [sx], [sz], src[None])
    offs = np.array([30, 40, 50])  # cells along x, arrivals before any boundary return
    o.set_receivers(2, 1, sx + offs, np.full(len(offs), sz))
    o.run()
    p = o.traces(4)
    a = np.log(0.8) / (2 * h * h)
    nf = 8192
    Sw = np.fft.rfft(0.5 * src.astype(np.float64), nf)
    w = 2 * np.pi * np.fft.rfftfreq(nf, dt)
    for i, off in enumerate(offs):
        P = np.zeros_like(w, dtype=complex)
        for ox in range(-5, 6):
            for oz in range(-5, 6):
                g = np.exp(a * ((ox * h) ** 2 + (oz * h) ** 2))
                rr = np.hypot((off - ox) * h, oz * h)
                P[1:] += g * (-0.25j) * hankel2(0, w[1:] * rr / c) / c ** 2
        ana = np.fft.irfft(1j * w * P * Sw, nf)[:nt]
        num = p[i]
        cc = np.dot(ana, num) / np.linalg.norm(ana) / np.linalg.norm(num)
        amp = np.linalg.norm(num) / np.linalg.norm(ana)
        assert cc > 0.997, (off, cc)
        assert abs(amp - 1) < 0.01, (off, amp)


# ---------------------------------------------------------------- energy diagnostic
def _undamped(o):
    o.set("coef1", np.ones((o.Nx, o.Nz)))
    o.set("coef2", np.full((o.Nx, o.Nz), float(np.float32(1e-3))))


@pytest.mark.parametrize("state,ek,ep", [
    # (vx, vz, sxx, szz, sxz), hand values per cell of Ek/(dx dz) and Ep/(dx dz) with
    # 1/invrho = 2, c11 = 4, c55 = 1: Ek = (vx^2 + vz^2) / invrho / 2; the reference's Ep
    # (BT:936-946) reduces to ((c11 + 2 c55)(sxx^2 - szz^2) / (4 c55 (c11 + c55)) + sxz^2 / c55) / 2
    ((1.0, 2.0, 0.0, 0.0, 0.0), 5.0, 0.0),
    ((0.0, 0.0, 3.0, 0.0, 0.0), 0.0, 1.35),
    ((0.0, 0.0, 0.0, 1.0, 0.0), 0.0, -0.15),     # szz alone counts negatively, as the reference's sum does
    ((0.0, 0.0, 0.0, 0.0, 0.5), 0.0, 0.125),
    ((1.0, 2.0, 3.0, 1.0, 0.5), 5.0, 1.325),     # the sxx*szz cross terms cancel
])
def test_energy_hand_values(state, ek, ep):
    nz, nx, dx, dz = 9, 11, 8.0, 12.5
    ctor, m, geom = small_case(nz=nz, nx=nx, nt=2, ealThick=2, dx=dx, dz=dz)
    geom["src"] = np.zeros_like(geom["src"])
    o = run_case(Oracle, ctor, m, geom, prec="f64")
    o.set_energy(1)
    _undamped(o)
    full = lambda v: np.full((o.Nx, o.Nz), v)
    o.set("invrho", full(0.5))
    o.set("c11", full(4.0))
    o.set("c55", full(1.0))
    for name, v in zip(("vx", "vz", "sxx", "szz", "sxz"), state):
        o.set(name, full(v))
    o.run(0, 1)  # a constant state has zero derivatives: one undamped step leaves it as it is
    E, Ek, Ep = o.energy()
    cells = nx * nz * dx * dz
    assert abs(Ek[0] - ek * cells) <= 1e-12 * cells
    assert abs(Ep[0] - ep * cells) <= 1e-12 * cells
    assert abs(E[0] - (ek + ep) * cells) <= 1e-12 * cells


def test_energy_kinetic_part_bounded_by_the_conserved_energy():
    """No damping (coef1 = 1, coef2 = dt everywhere), no source, a smooth initial pulse in
    an isotropic medium: the kinetic energy the diagnostic reports oscillates below the
    initial total elastic energy (computed here from the fields with the isotropic
    compliance), and the wavefield keeps that total to a few percent."""
    nz, nx, nt = 40, 44, 300
    ctor, m, geom = small_case(nz=nz, nx=nx, nt=nt, ealThick=2)
    m = {k: v.copy() for k, v in m.items()}
    m["eps"][:] = 0.0
    m["delta"][:] = 0.0
    m["theta"][:] = 0.0
    geom["src"] = np.zeros_like(geom["src"])
    o = run_case(Oracle, ctor, m, geom, prec="f64")
    o.set_energy(1)
    _undamped(o)
    X, Z = np.meshgrid(np.arange(o.Nx), np.arange(o.Nz), indexing="ij")
    bump = np.exp(-((X - o.Nx / 2) ** 2 + (Z - o.Nz / 2) ** 2) / (2 * 4.0 ** 2))
    o.set("sxx", 1e3 * bump)
    o.set("szz", 4e2 * bump)
    o.set("vx", 1e-4 * np.roll(bump, 3, axis=0))
    # the total over the whole grid: without damping the waves cross into the padding,
    # which the diagnostic (a sum over the model region) leaves out
    sl = (slice(None), slice(None))
    c11, c13, c55 = o.get("c11")[sl], o.get("c13")[sl], o.get("c55")[sl]
    rho = 1.0 / o.get("invrho")[sl]

    def total():
        vx, vz = o.get("vx")[sl], o.get("vz")[sl]
        a, b, c = o.get("sxx")[sl], o.get("szz")[sl], o.get("sxz")[sl]
        w = (c11 * a * a - 2 * c13 * a * b + c11 * b * b) / (c11 * c11 - c13 * c13) + c * c / c55
        return 0.5 * float((rho * (vx * vx + vz * vz) + w).sum()) * 100.0

    E0 = total()
    for k in range(nt):
        o.run(k, 1)
        if k % 50 == 49:
            assert abs(total() / E0 - 1.0) < 0.05, k
    _, Ek, _ = o.energy()
    assert np.all(Ek >= 0) and Ek.max() <= 1.02 * E0 and Ek.max() >= 0.3 * E0
    assert np.abs(np.diff(np.sign(np.diff(Ek)))).sum() > 2  # it oscillates, it does not just decay
