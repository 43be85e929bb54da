"""fdmod_run_energy_noalloc, the allocation-free run, through the C-ABI: the same
energy series and traces as the allocating path (fdmod_set_energy + fdmod_run +
fdmod_get_energy), bit for bit, over two calls in a row -- the second on a smaller shot,
so the shared static scratch holds the first call's contents when it starts -- and its
rejections (more iterations than its scratch holds, snapshots)."""
import numpy as np
import pytest

from common import run_case, small_case
from fdmod_hip import FdmodError, Modeling

pytestmark = pytest.mark.gpu

MAX_STEPS = 16384  # FDMOD_NOALLOC_MAX_STEPS


def _reference(form, fs, **kw):
    ctor, m, geom = small_case(form=form, freesurface=fs, **kw)
    ref = run_case(Modeling, ctor, m, geom)
    ref.set_energy(1)
    ref.run()
    return ctor, m, geom, ref, ref.energy()


def _check(g, ref, series, parts):
    for i in range(3):
        got = np.concatenate([p[i] for p in parts])
        assert got.shape == series[i].shape
        assert np.array_equal(got, series[i]), i
    assert np.isfinite(series[0]).all() and np.abs(series[0]).max() > 0
    for comp in (1, 3, 4):
        assert np.array_equal(g.traces(comp), ref.traces(comp)), comp


def test_noalloc_energy_equals_the_allocating_path_twice():
    print("noalloc: 60 x 80 coordinate shot, 300 iterations in calls of 200 and 100", flush=True)
    ctor, m, geom, ref, series = _reference("coord", 1, nz=60, nx=80, nt=300)
    g = run_case(Modeling, ctor, m, geom)
    _check(g, ref, series, [g.run_energy_noalloc(0, 200), g.run_energy_noalloc(200, 100)])
    # (no free surface: the Bond fix-up zeroes c55 on the surface row, where the energy
    # density divides by it)
    print("noalloc: then a smaller 40 x 50 Bond shot, 120 iterations in one call", flush=True)
    ctor, m, geom, ref, series = _reference("bond", 0, nz=40, nx=50, nt=120)
    g = run_case(Modeling, ctor, m, geom)
    _check(g, ref, series, [g.run_energy_noalloc(0, 120)])


def test_noalloc_rejections():
    ctor, m, geom = small_case(nz=30, nx=40, nt=50)
    g = run_case(Modeling, ctor, m, geom)
    with pytest.raises(FdmodError, match="at most"):
        g.run_energy_noalloc(0, MAX_STEPS + 1)
    g.set_snapshots(1, 0, 10, lambda it, kind, a: None)
    with pytest.raises(FdmodError, match="snapshots"):
        g.run_energy_noalloc(0, 10)
