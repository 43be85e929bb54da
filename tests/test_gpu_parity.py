"""GPU parity: libfdmod_hip (through the C-ABI) against the FP32 oracle.

The kernels evaluate the reference's float expressions with FP contraction
off and the setup in the reference's mixed precision, so parameters and traces
are expected to be bit-identical to the FP32 restatement of the reference.
"""
import numpy as np
import pytest

from common import rel_l2, run_case, small_case
from fdmod_hip import Modeling
from oracle import Oracle

pytestmark = pytest.mark.gpu

PARAMS = {"bond": ["coef1", "coef2", "D", "Beta", "invrho", "c11", "c13", "c15", "c33", "c35", "c55", "vp", "vs",
                   "rho", "theta"],
          "tensor": ["coef1", "coef2", "invrho", "c11", "c13", "c33", "c55", "sin2", "cos2", "sin2t"],
          "coord": ["coef1", "coef2", "invrho", "c11", "c13", "c33", "c55", "sin", "cos", "sinv", "cosv"]}


def _pair(**kw):
    ctor, m, geom = small_case(**kw)
    print("case", kw, flush=True)
    o = run_case(Oracle, ctor, m, geom)
    g = run_case(Modeling, ctor, m, geom)
    return o, g, geom


@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_setup_parameters_bitwise(form):
    o, g, _ = _pair(form=form, theta=None, freesurface=1, water=5)
    for name in PARAMS[form]:
        a, b = g.get(name), o.get(name).astype(np.float32)
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), f"{name}: {np.count_nonzero(~same)} of {a.size} differ"


@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
@pytest.mark.parametrize("fs", [0, 1])
def test_traces_bitwise(form, fs):
    o, g, _ = _pair(form=form, freesurface=fs, nt=400)
    o.run()
    g.run()
    for comp in (1, 3, 4):
        a, b = g.traces(comp), o.traces(comp).astype(np.float32)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), f"comp {comp}: rel l2 {rel_l2(a, b):.3e}"
    fields = ("vx", "vz", "sxx", "szz", "sxz") + (("txx", "tzz", "txz") if form == "tensor" else ())
    for f in fields:  # tau includes the source the sweep adds on the free surface's image rows
        assert np.array_equal(g.get(f), o.get(f).astype(np.float32)), f


# pressure-only receivers, recorded every second iteration by the stress sweep itself
@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_pressure_receivers_bitwise(form):
    o, g, _ = _pair(form=form, freesurface=1, nt=300, seistype=2, mt=2, rz=4)
    o.run()
    g.run()
    a, b = g.traces(4), o.traces(4).astype(np.float32)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b), f"rel l2 {rel_l2(a, b):.3e}"


@pytest.mark.parametrize("stype", [2, 3, 4, 5, 6, 7, 8])
def test_velocity_sources_bitwise(stype):
    o, g, _ = _pair(form="tensor", stype=stype, nt=200, nsrc=3)
    o.run()
    g.run()
    for comp in (1, 3, 4):
        assert np.array_equal(g.traces(comp), o.traces(comp).astype(np.float32)), comp


def test_mode2_mt_bitwise():
    o, g, _ = _pair(form="coord", mode=2, mt=3, nt=200, nsrc=2)
    o.run()
    g.run()
    for comp in (1, 3, 4):
        assert np.array_equal(g.traces(comp), o.traces(comp).astype(np.float32)), comp


# dx != dz: every invdx / invdz placement, the MEAL thicknesses in x and z, the source
# weights (rx, rz) and the coordinate form's invdx on d(sxx)/dz (CR:633) differ from
# their swapped forms only when the spacings differ
@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
@pytest.mark.parametrize("fs,stype", [(0, 1), (1, 1), (1, 4), (0, 7)])
def test_traces_bitwise_dx_ne_dz(form, fs, stype):
    water = 3 if fs and form != "bond" else 0  # water under the Bond fix-up is NaN (BT:543-568)
    o, g, _ = _pair(form=form, freesurface=fs, stype=stype, nt=300, dx=8.0, dz=12.5, water=water)
    for name in PARAMS[form]:
        a, b = g.get(name), o.get(name).astype(np.float32)
        assert np.array_equal(a, b, equal_nan=True), name
    o.run()
    g.run()
    for comp in (1, 3, 4):
        a, b = g.traces(comp), o.traces(comp).astype(np.float32)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), f"comp {comp}: rel l2 {rel_l2(a, b):.3e}"
    fields = ("vx", "vz", "sxx", "szz", "sxz") + (("txx", "tzz", "txz") if form == "tensor" else ())
    for f in fields:
        assert np.array_equal(g.get(f), o.get(f).astype(np.float32)), f


# snapshots (isoelastic.c:240-289): savew = 1 keeps the absorbing layers (the whole
# padded grid), savew = 0 crops them; vx and vz as they are, p = 0.5 (sxx + szz)
@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
@pytest.mark.parametrize("fs", [0, 1])
def test_pressure_receivers_after_unfused_stress_sources(form, fs):
    # three stype-1 sources with overlapping footprints, which the stress sweep does not
    # add itself, and pressure receivers through the footprints: the receivers sample
    # after the injection, in the reference's step order (BT:893-924)
    o, g, _ = _pair(form=form, freesurface=fs, nt=300, seistype=2, nsrc=3, sz=3, rz=3)
    o.run()
    g.run()
    assert np.array_equal(g.traces(4), o.traces(4).astype(np.float32))


@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
def test_edge_source_with_receivers_in_its_footprint(form):
    # one stype-1 source within five cells of the grid's left edge (not added by the
    # stress sweep): one launch adds it and records the receivers, seven of them inside
    # its footprint
    o, g, _ = _pair(form=form, nt=300, seistype=3, ealThick=3, sx=1, sz=3, rz=3)
    o.run()
    g.run()
    for comp in (1, 3, 4):
        assert np.array_equal(g.traces(comp), o.traces(comp).astype(np.float32)), comp


@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
@pytest.mark.parametrize("fs", [0, 1])
def test_pressure_receivers_on_vertical_lines(form, fs):
    # two full-depth receiver columns (61 receivers each, more than a strip's 64 lanes
    # hold per segment with the surface line) and a surface line: the stress sweep walks
    # each column's receiver list in the loop
    ctor, m, geom = small_case(form=form, freesurface=fs, nt=300, seistype=2, mt=1, sx=40, sz=2)
    nz, nx = ctor["nz"], ctor["nx"]
    zs = np.arange(0, nz + 1)
    geom["rpx"] = np.concatenate([np.full(nz + 1, 30), np.full(nz + 1, 44), np.arange(0, nx + 1)])
    geom["rpz"] = np.concatenate([zs, zs, np.full(nx + 1, 2)])
    print("case vsp", form, fs, len(geom["rpx"]), "receivers", flush=True)
    o = run_case(Oracle, ctor, m, geom)
    g = run_case(Modeling, ctor, m, geom)
    o.run()
    g.run()
    assert np.array_equal(g.traces(4), o.traces(4).astype(np.float32))


@pytest.mark.parametrize("form", ["bond", "tensor", "coord"])
@pytest.mark.parametrize("fs", [0, 1])
@pytest.mark.parametrize("seis", [2, 3])
def test_stress_source_across_a_strip_boundary(form, fs, seis):
    # the source's 11-row footprint straddles the boundary between the first two 64-row
    # strips (padded row izStart + 64), and the receivers lie on the source's row: each
    # strip adds its own rows of taps
    izs = (10 + 5) if fs else 5
    sz = izs + 64 - 15  # padded row izStart + 64 = the source cell
    o, g, _ = _pair(form=form, freesurface=fs, nz=80, nt=300, seistype=seis, sz=sz, rz=sz)
    o.run()
    g.run()
    for comp in ((4,) if seis == 2 else (1, 3, 4)):
        assert np.array_equal(g.traces(comp), o.traces(comp).astype(np.float32)), comp
    for f in ("vx", "vz", "sxx", "szz", "sxz"):
        assert np.array_equal(g.get(f), o.get(f).astype(np.float32)), f


@pytest.mark.parametrize("savew", [0, 1])
def test_snapshots_bitwise(savew):
    nt, every = 120, 40
    ctor, m, geom = small_case(nz=30, nx=40, nt=nt, form="coord", freesurface=1)
    g = run_case(Modeling, ctor, m, geom)
    snaps = {}
    g.set_snapshots(3, savew, every, lambda it, kind, a: snaps.__setitem__((it, kind), a))
    g.run()
    o = run_case(Oracle, ctor, m, geom)
    e = 0 if savew else o.e
    crop = (slice(e, o.Nx - e), slice(e, o.Nz - e))
    for k in range(nt):
        o.run(k, 1)
        if k % every:
            continue
        vx, vz = o.get("vx").astype(np.float32), o.get("vz").astype(np.float32)
        p = np.float32(0.5) * (o.get("sxx").astype(np.float32) + o.get("szz").astype(np.float32))
        for kind, ref in (("vx", vx), ("vz", vz), ("p", p)):
            got = snaps[(k, kind)]
            assert got.shape == ref[crop].shape, (k, kind, got.shape)
            assert np.array_equal(got, ref[crop]), (k, kind)
    assert sorted({it for it, _ in snaps}) == list(range(0, nt, every))
