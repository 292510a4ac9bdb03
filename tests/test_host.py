"""CPU tests of the library boundary and host logic (no device is used).

Covers: every C-ABI function declared in include/dugks.h is exported; the
config.txt rules of the reference Config class; derived numerics; the tanh
mesh against the reference's mesh.txt fixture; initial states; the raw
restart-dump layout; failure behaviour without a GPU.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import C, GOLDEN, ROOT, RT, W, feq, moments
import dugks_amd
from dugks_amd import _lib, configs
from dugks_amd.engine import derive, initial_state
from pyoracle import Oracle

HEADER = os.path.join(ROOT, "include", "dugks.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s*(dugks_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_function():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's functions
    assert sorted(_lib.PROTOTYPES) == names


def test_tanh_mesh_matches_reference_fixture():
    ref = np.loadtxt(os.path.join(GOLDEN, "mesh_tanh_a085_n128_w61.72.txt"))
    assert ref.shape == (129,)
    z = dugks_amd.mesh_tanh(128, 0.85, 61.72)
    assert np.abs(z - ref).max() < 1e-12
    # symmetric about the centre and clustered at the walls
    assert abs(z[64] - 30.86) < 1e-12
    assert (np.diff(z)[:64] > 0).all() and np.diff(z)[0] < np.diff(z)[63] / 2


def test_mesh_file_axis(tmp_path):
    txt = configs.turbulent_channel(nx=4, ny=128, nz=4, mesh_y="mesh.txt")
    (tmp_path / "mesh.txt").write_text(open(os.path.join(GOLDEN, "mesh_tanh_a085_n128_w61.72.txt")).read())
    w = dugks_amd.config_spacing(txt, 1, base_dir=str(tmp_path))
    ref = np.diff(np.loadtxt(os.path.join(GOLDEN, "mesh_tanh_a085_n128_w61.72.txt"))) * (2.0 / 61.72)
    assert np.abs(w - ref).max() < 1e-14
    w2 = dugks_amd.config_spacing(configs.turbulent_channel(nx=4, ny=128, nz=4), 1)
    assert np.abs(w2 - ref).max() < 1e-14


def test_reference_laminar_config_derivation():
    text = open(os.path.join(GOLDEN, "laminar_config.txt")).read()
    dims, p = derive(text)
    assert dims == (64, 32, 32)
    f32 = np.float32
    # reals go through float like std::stof (quirk Q3)
    assert p["visc"] == float(f32(0.166)) and p["visc"] != 0.166
    dx = float(f32(64)) / 64
    assert p["dt"] == float(f32(0.7)) * dx / math.sqrt(6 * RT)  # src/numerics.cu:105
    assert abs(p["tau"] - p["visc"] / RT) < 1e-16
    # coefficients agree with the oracle's for the same dt and viscosity
    o = Oracle((4, 4, 4), (0, 1, 1), (4, 4, 4), dt=p["dt"], visc=p["visc"])
    oc = o.coefficients()
    for k in ("bw0", "bw1", "bw2", "fw0", "fw1", "fw2", "hdt"):
        assert abs(p[k] - oc[k]) <= 1e-16 * max(1, abs(oc[k])), k


def test_config_parsing_rules():
    base = configs.base()
    # first occurrence wins; blanks trimmed; text after a second '=' ignored;
    # lines without '=' ignored
    text = "# comment line\nvisc = 0.5 = junk\n" + base + "visc=0.25\nno equals sign here\n"
    _, p = derive(text)
    assert p["visc"] == 0.5
    # parse_real=double keeps values exact
    _, p = derive(configs.base(visc=0.166))
    assert p["visc"] == 0.166
    # a missing key is an error that names it (the reference crashed, Q10)
    bad = "\n".join(l for l in base.splitlines() if not l.startswith("nx="))
    with pytest.raises(dugks_amd.DugksError) as e:
        derive(bad)
    assert e.value.code == -2 and "nx" in str(e.value)
    with pytest.raises(dugks_amd.DugksError):
        derive(configs.base(case="nonsense"))
    with pytest.raises(dugks_amd.DugksError):
        derive(configs.base(precision="fp8"))
    with pytest.raises(dugks_amd.DugksError):  # stretched periodic axis
        derive(configs.base(mesh_x="tanh:0.8"))
    with pytest.raises(dugks_amd.DugksError):  # gradient scheme needs uniform axes
        derive(configs.base(periods_y=0, mesh_y="tanh:0.8", recon="grad"))
    with pytest.raises(dugks_amd.DugksError):  # general case needs a dump
        derive(configs.base(case="general"))


def test_dt_rules():
    txt = configs.turbulent_channel(nx=8, ny=32, nz=8)
    _, p = derive(txt)
    w = dugks_amd.config_spacing(txt, 1)
    assert p["dt"] == 0.5 * w.min() / math.sqrt(2)
    assert abs(p["dy"] - w.min()) < 1e-18
    _, p2 = derive(configs.base(CFL=0.5, xwidth=2.0, nx=8, ywidth=1.0, ny=16))
    assert p2["dt"] == 0.5 * (2.0 / 8) / math.sqrt(2)  # dx rule of the reference


def test_initial_states():
    # Taylor-Green (src/dugks.cu:222-240) with u0 = 0.05
    L = 2 * math.pi
    txt = configs.taylor_green(n=8, u0=0.05)
    f = initial_state(txt)
    rho, j = moments(f)
    x = (np.arange(8) + 0.5) * L / 8
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = 1 + 0.05 ** 2 / 16 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / RT
    ux = 0.05 * np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -0.05 * np.cos(X) * np.sin(Y) * np.cos(Z)
    assert np.abs(rho - r).max() < 1e-14
    assert np.abs(j[..., 0] / rho - ux).max() < 1e-14
    assert np.abs(j[..., 1] / rho - uy).max() < 1e-14
    assert np.abs(j[..., 2]).max() < 1e-15
    assert np.abs(f - feq(r, np.stack([ux, uy, 0 * ux], -1))).max() < 1e-15
    # plug flow of the reference's channel case (Q11)
    f = initial_state(configs.base(case="channel"))
    rho, j = moments(f)
    assert np.abs(j[..., 2] / rho - 0.1).max() < 1e-15
    # rest
    f = initial_state(configs.base(case="rest"))
    assert np.abs(f - W).max() < 1e-16


def test_turbulent_channel_state():
    txt = configs.turbulent_channel(nx=16, ny=32, nz=16)
    dims, p = derive(txt)
    f = initial_state(txt)
    rho, j = moments(f)
    # the Chapman-Enskog part carries no mass or momentum: rho = 1 exactly
    assert np.abs(rho - 1).max() < 1e-13
    u = j / rho[..., None]
    w = dugks_amd.config_spacing(txt, 1)
    yc = np.cumsum(w) - 0.5 * w
    eta = yc - 1.0
    um = u[..., 0].mean(axis=(0, 2))
    # mean profile: viscous sublayer / log law with u_tau = Re_tau nu / H
    utau = 180 * 3.0e-5
    yp = (1 - np.abs(eta)) * utau / 3.0e-5
    up = np.where(yp < 11.2, yp, np.log(yp) / 0.4 + 5.6)
    assert np.abs(um / utau - up).max() < 0.2
    # perturbation amplitude 10 % of the centreline velocity
    Uc = utau * (math.log(180) / 0.4 + 5.6)
    up3 = u - np.stack([np.broadcast_to(um[None, :, None], u.shape[:3]), 0 * um[None, :, None] + 0 * u[..., 1],
                        0 * u[..., 2]], -1)
    amp = np.sqrt((up3 ** 2).sum(-1)).max()
    assert 0.05 * Uc < amp < 0.15 * Uc
    # non-equilibrium part is present and traceless in the sense sum_i f_neq = 0
    fneq = f - feq(rho, u)
    assert np.abs(fneq).max() > 1e-9
    assert np.abs(fneq.sum(-1)).max() < 1e-13


def test_dump_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    nx, ny, nz = 3, 4, 2
    f = rng.random((nz, ny, nx, 19))
    p8 = str(tmp_path / "a.dat")
    dugks_amd.dump_write(p8, f, 8)
    raw = np.fromfile(p8, dtype=np.float64)
    # raw real values in [k][j][i][ip] order, no header (src/output.cu:22-41)
    assert raw.size == f.size and (raw == f.ravel()).all()
    k, j, i, q = 1, 2, 0, 7
    assert raw[((k * ny + j) * nx + i) * 19 + q] == f[k, j, i, q]
    assert (dugks_amd.dump_read(p8, f.size, 8) == f.ravel()).all()
    p4 = str(tmp_path / "b.dat")
    dugks_amd.dump_write(p4, f, 4)
    raw4 = np.fromfile(p4, dtype=np.float32)
    assert (raw4 == f.ravel().astype(np.float32)).all()
    assert (dugks_amd.dump_read(p4, f.size, 4) == f.ravel().astype(np.float32)).all()
    with pytest.raises(dugks_amd.DugksError) as e:
        dugks_amd.dump_read(p4, f.size, 8)  # too short for FP64
    assert e.value.code == -3


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_init_without_device_fails_cleanly():
    with pytest.raises(dugks_amd.DugksError) as e:
        dugks_amd.Engine(configs.base())
    assert e.value.code == -4


def test_driver_reports_bad_config(tmp_path):
    exe = os.path.join(ROOT, "dugks-gpu_amd", "lib", "dugks_main")
    r = subprocess.run([exe, str(tmp_path / "missing.txt")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "cannot open config file" in r.stderr


def test_velmean_matches_reference_script_output(tmp_path):
    # golden pair written by tests/golden/make_velmean_golden.py: the
    # reference's examples/channel/ProcessProfiles.py run on a synthetic
    # profiles-2.dat (8 rows, blocks at steps 1, 5001, ..., 20001, blocks
    # after step 5001 averaged, mean = column 2 (w), stress = column 3 (<u'w'>)).
    # The C++ restatement must write the same bytes.
    src = os.path.join(GOLDEN, "velmean_profiles-2.dat")
    out = tmp_path / "VELMEAN.dat"
    dugks_amd.velmean(src, 8, str(out), ignore_before=5001, mean_col=2, stress_col=3)
    with open(os.path.join(GOLDEN, "velmean_VELMEAN.dat"), "rb") as f:
        expect = f.read()
    assert out.read_bytes() == expect
    with pytest.raises(dugks_amd.DugksError):
        dugks_amd.velmean(src, 8, str(out), ignore_before=10 ** 6)


def test_velmean_writes_python_repr(tmp_path):
    # VELMEAN.dat rows are Python f-strings of floats: every value is written
    # as repr() writes it.  A one-block, two-row file with mean (x, x) and
    # stress (x, -x) makes the first output row hold x itself.
    rng = np.random.default_rng(11)
    vals = [0.0, -0.0, 1.0, -2.5, 0.1, 1e-4, 9.999999999999999e-05, 1e-5, 1.5e-7, 123.456,
            1e15, 9999999999999998.0, 1e16, 1.2345678901234567e17, 5e-324, 1.7976931348623157e308,
            2.0 ** 60, 1 / 3, 2 / 3, 4.8080371999999990e-05]
    vals += list(rng.standard_normal(40) * 10.0 ** rng.integers(-12, 18, 40))
    for k, x in enumerate(vals):
        src = tmp_path / f"p{k}.dat"
        row = [x, 0, 0, x, 0, 0, x, x, x]
        with open(src, "w") as f:
            f.write("00000001\n")
            f.write(" ".join(repr(float(v)) for v in row) + "\n")
            f.write(" ".join(repr(float(v)) for v in [x, 0, 0, -x, 0, 0, x, x, x]) + "\n")
        out = tmp_path / f"v{k}.dat"
        dugks_amd.velmean(str(src), 2, str(out), ignore_before=0, mean_col=0, stress_col=3)
        first = out.read_text().splitlines()[0].split()
        a = 0.0 + float(x)  # the block sums start from 0.0, as the script's zeros arrays do
        b = 0.0 - float(x)
        assert first == ["1.0", repr(0.5 * (a + a)), repr(0.5 * (a - b)), repr(0.5 * (a + a)),
                         repr(0.5 * (a + a)), repr(0.5 * (a + a))], (x, first)


def test_device_pointer_checks_reject_host_tensors():
    # a host tensor or a foreign object never reaches a C entry point as a
    # device pointer; None and raw integers pass through
    torch = pytest.importorskip("torch")
    from dugks_amd.engine import _dev_ptr
    with pytest.raises(ValueError):
        _dev_ptr(torch.zeros(19 * 8, dtype=torch.float64), 19 * 8, 0)
    with pytest.raises(ValueError):
        _dev_ptr(np.zeros(19 * 8), 19 * 8, 0)
    assert _dev_ptr(None, 19, 0) is None
    assert _dev_ptr(1234, 19, 0).value == 1234
