"""The drop-in host programs and SU utilities: file formats, parameters and every
failure the reference reports with err() (missing parameters, unreadable files,
unstable dt, positions outside the model, receiver-file mismatch).  All of these
checks run before the first device call, so they are tested here without a GPU."""
import os
import subprocess

import numpy as np
import pytest

import sufile
from fdmod_hip import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# FDMOD_HOST_BIN / FDMOD_HOST_LIB select other builds of the host programs and helpers
# (the sanitizer build of oracle/sanitize_check.sh)
BIN = os.environ.get("FDMOD_HOST_BIN") or os.path.join(ROOT, "tti-fdmodeling_amd", "bin")
HOSTLIB = os.environ.get("FDMOD_HOST_LIB") or os.path.join(ROOT, "tti-fdmodeling_amd", "_build", "libfdmod_host.so")
PROGS = ["sufdmod2dTTIelastic_bond_transformation", "sufdmod2dTTIelastic_tensor_rotation",
         "sufdmod2dTTIelastic_coordinate_rotation"]


def run(prog, *args, cwd=None, stdin=None, timeout=60):
    return subprocess.run([os.path.join(BIN, prog)] + list(args), cwd=cwd, stdin=stdin, capture_output=True,
                          timeout=timeout)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not all(os.path.exists(os.path.join(BIN, p)) for p in PROGS + ["sumksrc2d", "sumkrcev2d", "cutbin2d"]):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tti-fdmodeling_amd"), "programs"], check=True)


def test_sumksrc2d_header_and_ricker(tmp_path):
    r = run("sumksrc2d", "nt=500", "dt=0.001", "sdepth=2", "sx=150", "fpeak=15")
    assert r.returncode == 0, r.stderr
    p = tmp_path / "s.su"
    p.write_bytes(r.stdout)
    (hdr, data), = sufile.read(p)
    assert hdr["sdepth"] == 2000 and hdr["sx"] == 150000 and hdr["scalel"] == -1000 and hdr["scalco"] == -1000
    assert hdr["ns"] == 500 and hdr["dt"] == 1000 and hdr["ntr"] == 1 and abs(hdr["d1"] - 0.001) < 1e-9
    ref = synth.ricker(15.0, 1e-3, 500)
    assert np.allclose(data, ref, rtol=1e-5, atol=1e-7)
    # Ricker: peak of 1 at t = delay = 1.5/f, zero first sample
    assert data[0] == 0.0 and abs(data.max() - 1.0) < 1e-3 and abs(np.argmax(data) * 1e-3 - 0.1) < 1.5e-3


@pytest.mark.parametrize("mode,expect", [(0, list(range(0, 21))), (1, list(range(0, 11))), (2, list(range(10, 21)))])
def test_sumkrcev2d_spread(tmp_path, mode, expect):
    r = run("sumkrcev2d", "sx=10", "gdel=3", "offsetmax=100", "nx=21", f"mode={mode}")
    assert r.returncode == 0, r.stderr
    p = tmp_path / "r.su"
    p.write_bytes(r.stdout)
    (hx, x), (hz, z) = sufile.read(p)
    assert list(x.astype(int)) == expect
    assert np.all(z == 3) and hx["ntr"] == 2 and hx["gdel"] == 3000


def test_sumkrcev2d_offset_window(tmp_path):
    r = run("sumkrcev2d", "sx=50", "offsetmin=5", "offsetmax=20", "dtr=5", "nx=200", "mode=0")
    p = tmp_path / "r.su"
    p.write_bytes(r.stdout)
    (hx, x), _ = sufile.read(p)
    assert list(x.astype(int)) == [30, 35, 40, 45, 55, 60, 65, 70]


def test_cutbin2d_matches_slicing(tmp_path):
    a = np.arange(12 * 9, dtype=np.float32).reshape(9, 12)  # n2=9 records of n1=12
    src = tmp_path / "a.bin"
    a.tofile(src)
    with open(src, "rb") as f:
        r = run("cutbin2d", "n1=12", "n2=9", "fx1=2", "fx2=3", "nx1=5", "nx2=4", stdin=f)
    assert r.returncode == 0, r.stderr
    out = np.frombuffer(r.stdout, dtype=np.float32).reshape(4, 5)
    assert np.array_equal(out, a[3:7, 2:7])


def write_case(d, nz=30, nx=40, nt=50, sx=20, sz=2, dt=1e-3, theta=20.0):
    m = synth.model_small(nz=nz, nx=nx, seed=3)
    m["theta"][:] = theta
    for k, name in (("vp", "vp"), ("vs", "vs"), ("rho", "rho"), ("eps", "eps"), ("delta", "del"), ("theta", "th")):
        m[k].tofile(d / f"{name}.bin")
    s = run("sumksrc2d", f"nt={nt}", f"dt={dt}", f"sdepth={sz}", f"sx={sx}", "fpeak=15")
    (d / "src.su").write_bytes(s.stdout)
    r = run("sumkrcev2d", f"sx={sx}", "gdel=2", f"offsetmax={nx}", f"nx={nx}")
    (d / "rec.su").write_bytes(r.stdout)
    return m


def base_args(nz=30, nx=40, nt=50, dt=1e-3):
    return [f"nz={nz}", f"nx={nx}", f"nt={nt}", "dx=10", "dz=10", f"dt={dt}", "fpeak=15", "vpfile=vp.bin",
            "vsfile=vs.bin", "rhofile=rho.bin", "epsilonfile=eps.bin", "deltafile=del.bin", "thetafile=th.bin",
            "sourcefile=src.su", "recposfile=rec.su", "seistype=3", "wsave=0", "ealThick=5"]


@pytest.mark.parametrize("prog", PROGS)
def test_selfdoc_without_arguments(prog):
    r = run(prog)
    assert r.returncode != 0 and b"nz" in r.stderr and b"sourcefile" in r.stderr


@pytest.mark.parametrize("missing", ["dx", "dt", "nz", "fpeak", "vpfile", "epsilonfile", "thetafile", "sourcefile"])
def test_missing_required_parameter(tmp_path, missing):
    write_case(tmp_path)
    args = [a for a in base_args() if not a.startswith(missing + "=")]
    r = run(PROGS[0], *args, cwd=tmp_path)
    assert r.returncode == 1 and b"<error>" in r.stderr


def test_unstable_dt_rejected_before_device(tmp_path):
    write_case(tmp_path, dt=1e-3)
    args = [a if not a.startswith("dt=") else "dt=0.01" for a in base_args()]
    r = run(PROGS[0], *args, cwd=tmp_path)
    assert r.returncode == 1 and b"unstable" in r.stderr


def test_source_outside_model(tmp_path):
    write_case(tmp_path, sx=38)
    s = run("sumksrc2d", "nt=50", "dt=0.001", "sdepth=2", "sx=41", "fpeak=15")
    (tmp_path / "src.su").write_bytes(s.stdout)
    r = run(PROGS[1], *base_args(), cwd=tmp_path)
    assert r.returncode == 1 and b"position of source in 2nd dimension (41)" in r.stderr


def test_receiver_outside_and_mismatch(tmp_path):
    write_case(tmp_path)
    sufile.write(tmp_path / "rec.su", [({"ns": 3, "ntr": 2}, np.array([1, 2, 99], "f4")),
                                       ({"ns": 3, "ntr": 2}, np.array([2, 2, 2], "f4"))])
    r = run(PROGS[2], *base_args(), cwd=tmp_path)
    assert r.returncode == 1 and b"position of receiver in 2nd dimension (99)" in r.stderr
    sufile.write(tmp_path / "rec.su", [({"ns": 3, "ntr": 2}, np.array([1, 2, 3], "f4")),
                                       ({"ns": 2, "ntr": 2}, np.array([2, 2], "f4"))])
    r = run(PROGS[2], *base_args(), cwd=tmp_path)
    assert r.returncode == 1 and b"not match" in r.stderr


def test_unreadable_model_file(tmp_path):
    write_case(tmp_path)
    os.remove(tmp_path / "eps.bin")
    r = run(PROGS[0], *base_args(), cwd=tmp_path)
    assert r.returncode == 1 and b"epsilon" in r.stderr


def test_parfile_and_last_value_wins(tmp_path):
    write_case(tmp_path)
    (tmp_path / "p.par").write_text("nz=30 nx=40 # comment\nnt=50 dx=10 dz=10\n")
    args = [a for a in base_args() if a.split("=")[0] not in ("nz", "nx", "nt", "dx", "dz", "dt")]
    r = run(PROGS[0], "par=p.par", "dt=0.01", "dt=0.5", *args, cwd=tmp_path)  # last dt (0.5) is unstable
    assert r.returncode == 1 and b"unstable" in r.stderr and b"0.0" in r.stderr


def test_default_vs_rho_files_written(tmp_path):
    m = write_case(tmp_path)
    args = [a for a in base_args() if not a.startswith(("vsfile", "rhofile"))]
    args = [a if not a.startswith("dt=") else "dt=0.5" for a in args]  # stop at the stability check
    r = run(PROGS[0], *args, cwd=tmp_path)
    assert r.returncode == 1
    vs = np.fromfile(tmp_path / "vs_00030_00040.bin", dtype=np.float32).reshape(40, 30)
    rho = np.fromfile(tmp_path / "rho_00030_00040.bin", dtype=np.float32).reshape(40, 30)
    vp = m["vp"].astype(np.float64)
    assert np.array_equal(vs, np.where(vp <= 1500, 0, 0.577350269189626 * vp).astype(np.float32))
    assert np.allclose(rho, np.where(vp <= 1500, 1000, 1741 * (vp / 1000) ** 0.25), rtol=1e-6)


@pytest.mark.parametrize("wtype", [1, 2, 3, 4])
def test_sumksrc2d_first_sample_zero_for_every_wavelet(tmp_path, wtype):
    # gauss_wlt, fdgauss_wlt, riker_wlt and driker_wlt all end with src[0] = 0
    # (src/lib/wavelet/wavelet.c:37, 67, 99, 130); a 20 ms delay gives every type a
    # clearly nonzero value there before that
    r = run("sumksrc2d", "nt=64", "dt=0.001", "sdepth=2", "sx=3", "fpeak=15", f"wtype={wtype}", "delay=0.02")
    assert r.returncode == 0, r.stderr
    p = tmp_path / "s.su"
    p.write_bytes(r.stdout)
    (_, data), = sufile.read(p)
    assert data[0] == 0.0 and np.abs(data[1:]).max() > 0
    if wtype == 4:  # Gaussian: 0.5/(pi f)^2 exp(-(pi f t)^2), A ignored (wavelet.c:31)
        t = np.arange(1, 4) * np.float32(0.001) - 0.02
        ref = 0.5 / (np.pi * 15.0) ** 2 * np.exp(-(np.pi * 15.0 * t) ** 2)
        assert np.allclose(data[1:4], ref, rtol=1e-5)


def test_sumksrc2d_reference_diagnostics():
    r = run("sumksrc2d", "nt=10", "dt=0.001", "sdepth=2", "sx=3")
    assert r.returncode == 0
    err = r.stderr.decode()
    for line in ("<warning>: didn't specify a wavelet file! The program will use one of wavelet in the document!",
                 "<warning>: didn't specify type of wavelet! The program will set it 1!",
                 "<warning>: didn't specify peak amplitude of wavelet! The program will set it as    1.00000.",
                 "<warning>: didn't specify fpeak frequency of wavelet! The program will set it as   10.00000 Hz.",
                 "<warning>: didn't specify delay time of wavelet! The program will set delay = 1.5 / fpeak =   "
                 "0.15000 Hz!"):
        assert line in err, line
    r = run("sumksrc2d", "nt=10", "sdepth=2", "sx=3")
    assert r.returncode == 1 and "<error>: must specify sampleing interval in time dimension!" in r.stderr.decode()


def test_program_reference_diagnostics(tmp_path):
    # the defaulted parameters are announced with the reference's lines (BT:146-294); the
    # run then stops at the stability check, before any device work
    write_case(tmp_path)
    keep = ("nz", "nx", "nt", "dx", "dz", "fpeak", "vpfile", "vsfile", "rhofile", "epsilonfile", "deltafile",
            "thetafile", "sourcefile")
    args = [a for a in base_args() if a.split("=")[0] in keep] + ["dt=0.5"]
    r = run(PROGS[2], *args, cwd=tmp_path)
    assert r.returncode == 1
    err = r.stderr.decode()
    for line in ("<warning>: didn't specify first sample in 2nd dimension. The program will set it as 0.000000.",
                 "<warning>: didn't specify first sample in 1st dimension. The program will set it as 0.000000.",
                 "<warning>: didn't specify the scalar to be apllied to dip angle. The program will set it as "
                 "1.000000.",
                 "<warning>: didn't specify the order of gaussian weighting function! The program will set it as "
                 "0.05 * ealThick + 1.5.",
                 "<warning>: didn't specify whether to use free surface boundary conditions. The program will not "
                 "use it.",
                 "<warning>: didn't specify the number of absorbing layers! The program will set it 30.",
                 "<warning>: didn't specify the mode of modelng! The program will it as 1 which means the program "
                 "will execute the forward modeling.",
                 "<warning>: didn't specify the type of source,  and will set stype as 1!",
                 "<warning>: didn't specify which wavefiled will be recorded. The program will set it as 0 which "
                 "means the program will not save the seismic record.",
                 "<warning>: didn't specify the sampling interval of snapshot. The program will set it as 500 / "
                 "(dt * 1000) = 1.",
                 "<warning>: didn't specify receiver_p position file!",
                 "<warning>: didn't specify prefix of snapshot file! The program will set it as snap.",
                 "<warning>: didn't specify the prefix of traces file. The program will set it as traces",
                 "the finite differenct scheme is unstable!"):
        assert line in err, line


def _resample(y, din, fin, t):
    import ctypes as ct

    lib = ct.CDLL(HOSTLIB)
    fp = ct.POINTER(ct.c_float)
    lib.resample8.argtypes = [ct.c_int, ct.c_float, ct.c_float, fp, ct.c_int, fp, fp]
    y = np.ascontiguousarray(y, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    out = np.zeros(len(t), np.float32)
    lib.resample8(len(y), din, fin, y.ctypes.data_as(fp), len(t), t.ctypes.data_as(fp), out.ctypes.data_as(fp))
    return out


@pytest.fixture(scope="module")
def hostlib():
    if not os.path.exists(HOSTLIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tti-fdmodeling_amd"), "_build/libfdmod_host.so"],
                       check=True)


@pytest.mark.parametrize("d", [0.125, 185 / 512, 0.5, 0.875])
def test_resample8_weights_are_the_least_squares_design(hostlib, d):
    """The impulse responses at offset d equal the dense solution of the normal equations
    of the 8-point least-squares interpolator (band fmax = 0.066 + 0.265 ln 8 of Nyquist):
    T w = c with T[j][k] = sinc(fmax (j - k)), c[j] = sinc(fmax (3 - j + d)), taps at
    samples i-3 .. i+4 for x = i + d."""
    from scipy.linalg import solve, toeplitz

    fmax = 0.066 + 0.265 * np.log(8.0)
    w = solve(toeplitz(np.sinc(fmax * np.arange(8))), np.sinc(fmax * (3 - np.arange(8) + d)))
    got = []
    for j in range(8):
        y = np.zeros(20, np.float32)
        y[10 - 3 + j] = 1.0
        got.append(_resample(y, 1.0, 0.0, [10 + d])[0])
    assert np.abs(np.array(got) - w).max() < 1e-6
    assert abs(w.sum() - 1.0) < 2e-3  # the design trades exact DC gain for the whole band


def test_resample8_band_limited_sinusoids(hostlib):
    """A source file sampled at 2 ms read into a 0.7 ms model (read2dsource, obsinfo.c:82):
    sinusoids up to 0.6 of the file's Nyquist frequency are interpolated at fractional
    offsets to within 4e-3 of their amplitude; equal intervals copy the trace; outside the
    trace the result is zero."""
    din, n = 0.002, 2000
    tin = np.arange(n) * din
    t = (np.arange(int(n * din / 0.0007)) * 0.0007).astype(np.float32)
    inside = (t > 10 * din) & (t < (n - 10) * din)
    for frac in (0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6):
        f = frac / (2 * din)
        out = _resample(np.sin(2 * np.pi * f * tin + 0.3), din, 0.0, t)
        exact = np.sin(2 * np.pi * f * t.astype(np.float64) + 0.3)
        assert np.abs(out[inside] - exact[inside]).max() < 4e-3, frac
    y = np.random.default_rng(0).standard_normal(500).astype(np.float32)
    assert np.array_equal(_resample(y, 0.001, 0.0, np.arange(500, dtype=np.float32) * np.float32(0.001)), y)
    assert np.all(_resample(y, 0.001, 0.0, [-0.02, 0.6, 1.0]) == 0.0)
