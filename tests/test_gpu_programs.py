"""End-to-end runs of the drop-in programs on the GPU against the FP32 oracle:
SU traces (data and headers), side-effect parameter dumps, snapshots, energy files
and the timing lines."""
import os
import re
import subprocess

import numpy as np
import pytest

import sufile
from oracle import Oracle
from test_host_programs import BIN, PROGS, base_args, run, write_case

pytestmark = pytest.mark.gpu

FORM_OF = {PROGS[0]: "bond", PROGS[1]: "tensor", PROGS[2]: "coord"}


@pytest.mark.parametrize("prog", PROGS)
def test_program_matches_oracle(tmp_path, prog):
    nz, nx, nt = 30, 40, 120
    m = write_case(tmp_path, nz=nz, nx=nx, nt=nt)
    args = [a for a in base_args(nz, nx, nt) if not a.startswith("wsave")]
    args += ["wsave=3", "snap_sample=40", "energy=1", "freesurface=1", "mt=2", "shotID=7", "fx=100", "fz=5"]
    print("program", prog, flush=True)
    r = run(prog, *args, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    out = r.stdout.decode()
    assert re.search(r"Finite difference modeling takes [0-9.]+ s\.", out)
    assert re.search(r"Total time takes [0-9.]+ s\.", out)

    form = FORM_OF[prog]
    o = Oracle(form, nz, nx, nt, 10.0, 10.0, 1e-3, ealThick=5, freesurface=1, fpeak=15.0)
    o.set_model(m["vp"], m["vs"], m["rho"], m["eps"], m["delta"], m["theta"])
    (shdr, wav), = sufile.read(tmp_path / "src.su")
    o.set_source(1, [20], [2], wav[None])
    rec = sufile.read(tmp_path / "rec.su")
    rx, rz = rec[0][1].astype(int), rec[1][1].astype(int)
    o.set_receivers(3, 2, rx, rz)
    o.set_energy(1)
    snaps = {}
    for k in range(nt):
        o.run(k, 1)
        if k % 40 == 0:
            e = o.e
            snaps[k] = {n: o.get(n).astype(np.float32)[e:-e, e:-e] for n in ("vx", "vz", "sxx", "szz")}

    ns = (nt - 1) // 2 + 1
    for comp, flag in (("vx", 1), ("vz", 3), ("p", 4)):
        path = tmp_path / f"traces_{comp}_0007_{ns:05d}_{len(rx):05d}_02.00ms.su"
        tr = sufile.read(path)
        data = np.stack([d for _, d in tr])
        assert np.array_equal(data, o.traces(flag).astype(np.float32)), comp
        h0 = tr[3][0]
        assert h0["ns"] == ns and h0["dt"] == 2000 and h0["fldr"] == 7 and h0["tracr"] == 4
        assert h0["sx"] == int((100 + 20 * 10) * 1000) and h0["gx"] == int((100 + rx[3] * 10) * 1000)
        assert h0["sdel"] == int((5 + 2 * 10) * 1000) and h0["offset"] == h0["gx"] - h0["sx"]
        assert h0["trid"] == rx[3]

    Nz, Nx = o.Nz, o.Nx
    names = ["D", "Beta", "coef1", "coef2", "invrho", "c11", "c13", "c33", "c55"] + (["c15", "c35"] if form == "bond" else [])
    for n in names:
        a = np.fromfile(tmp_path / f"{n}_{Nz:05d}_{Nx:05d}.bin", dtype=np.float32).reshape(Nx, Nz)
        b = o.get(n).astype(np.float32)
        assert np.array_equal(a, b) or np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(
            np.nan_to_num(a), np.nan_to_num(b)), n

    for it, f in snaps.items():
        ms = int(1000 * it * np.float32(1e-3))
        vx = np.fromfile(tmp_path / f"snap_vx_{nz:05d}_{nx:05d}_{ms:05d}ms.bin", dtype=np.float32).reshape(nx, nz)
        p = np.fromfile(tmp_path / f"snap_p_{nz:05d}_{nx:05d}_{ms:05d}ms.bin", dtype=np.float32).reshape(nx, nz)
        assert np.array_equal(vx, f["vx"])
        assert np.array_equal(p, np.float32(0.5) * (f["sxx"] + f["szz"]))

    E = np.loadtxt(tmp_path / "Eng.txt")
    Eo, _, _ = o.energy()
    fin = np.isfinite(Eo) & (np.abs(Eo) > 0)
    assert np.allclose(E[fin], Eo[fin], rtol=2e-3, atol=0)
