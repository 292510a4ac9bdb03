"""GPU parity of the HIP step (through the C ABI) against the FP64 oracle.

Bar: FP64 agrees with the oracle to <= 1e-13 relative per step on the same
inputs; FP32 and FP16-deviation storage within their stated tolerances.
Larger sizes are checked through size-independent properties (mass,
momentum input, restart continuation, analytic Poiseuille start-up).
"""
import math
import os

import numpy as np
import pytest

from conftest import C, W, feq, moments, random_state, reflect_state
import dugks_amd
from dugks_amd import KERNEL_FUSED, KERNEL_TWOPASS, PREC_FP16N, PREC_FP16S, PREC_FP32, PREC_FP64, configs
from pyoracle import Oracle

pytestmark = pytest.mark.gpu

KERNELS = [KERNEL_TWOPASS, KERNEL_FUSED]


def _engine(text, precision=PREC_FP64, kernel=KERNEL_TWOPASS):
    return dugks_amd.Engine(text, precision=precision, kernel=kernel)


def _oracle_for(eng, text, recon="tri"):
    """Oracle with the engine's geometry and coefficients."""
    p = eng.params
    nx, ny, nz = eng.shape
    kv = dict(l.split("=", 1) for l in text.strip().splitlines())
    periodic = tuple(int(kv.get(k, "1")) for k in ("periods_x", "periods_y", "periods_z"))
    width = tuple(float(kv[k]) for k in ("xwidth", "ywidth", "zwidth"))
    famb = tuple(float(kv.get(k, "0")) for k in ("ambient_fx", "ambient_fy", "ambient_fz"))
    spacing = [None if kv.get(f"mesh_{a}", "uniform") == "uniform" else eng.spacing(i)
               for i, a in enumerate("xyz")]
    return Oracle((nx, ny, nz), periodic, width, dt=p["dt"], visc=p["visc"], famb=famb,
                  spacing=spacing, recon=kv.get("recon", "tri"))


def _relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# cases both step kernels cover (the fused kernel needs nx % 32 == 0 and
# ny % 4 == 0, or nx % 16 == 0 and ny % 12 == 0, and a periodic uniform z axis)
CASES = {
    "periodic": dict(nx=32, ny=8, nz=4, xwidth=3.2, ywidth=0.8, zwidth=0.4),
    "periodic_wide": dict(nx=64, ny=8, nz=3, xwidth=6.4, ywidth=0.8, zwidth=0.3),
    "walls_x": dict(nx=32, ny=8, nz=4, periods_x=0, ambient_fz=2e-4, xwidth=3.2, ywidth=0.8, zwidth=0.4),
    "walls_y": dict(nx=32, ny=16, nz=4, periods_y=0, ambient_fx=2e-4, xwidth=3.2, ywidth=1.6, zwidth=0.4),
    "walls_xy": dict(nx=32, ny=16, nz=3, periods_x=0, periods_y=0, ambient_fz=1e-4, xwidth=3.2,
                     ywidth=1.6, zwidth=0.3),
    "stretched_y": dict(nx=32, ny=16, nz=4, periods_y=0, mesh_y="tanh:0.85", dt_rule="min",
                        ambient_fx=2e-4, xwidth=3.2, ywidth=2.0, zwidth=0.4),
    "stretched_x": dict(nx=32, ny=8, nz=4, periods_x=0, mesh_x="tanh:0.8", dt_rule="min",
                        ambient_fy=2e-4, xwidth=2.0, ywidth=0.8, zwidth=0.4),
    "force_all": dict(nx=32, ny=8, nz=4, ambient_fx=1e-4, ambient_fy=-2e-4, ambient_fz=3e-4,
                      xwidth=3.2, ywidth=0.8, zwidth=0.4),
    # a multiple of 8 tiles (the XCD-ordered tile mapping) with a partial
    # last z chunk (80 = 64 + 16)
    "tiles_xcd": dict(nx=64, ny=16, nz=80, periods_y=0, mesh_y="tanh:0.85", dt_rule="min",
                      ambient_fx=2e-4, xwidth=6.4, ywidth=2.0, zwidth=8.0),
    # the 16 x 12 tile (ny % 12 == 0): walls and stretching in y, walls in x,
    # several tiles in x and y with XCD ordering
    "tall_walls_y": dict(nx=32, ny=24, nz=6, periods_y=0, mesh_y="tanh:0.85", dt_rule="min",
                         ambient_fx=2e-4, xwidth=3.2, ywidth=2.0, zwidth=0.6),
    "tall_walls_x": dict(nx=16, ny=12, nz=5, periods_x=0, ambient_fz=2e-4, xwidth=1.6, ywidth=1.2,
                         zwidth=0.5),
    "tall_tiles": dict(nx=64, ny=24, nz=40, periods_y=0, ambient_fx=1e-4, ambient_fz=-1e-4, xwidth=6.4,
                       ywidth=2.4, zwidth=4.0),
}
# gradient reconstruction (calc_fx; uniform axes): both kernels, the fused
# one on 32 x 4 and 16 x 12 tiles, walls on x, y or both, several tiles
# with XCD ordering and z chunks (nz = 40 and 80 take 8-plane chunks), and
# nz = 3, where the ring's planes p - 1 and p + 1 are one plane apart
GRAD_CASES = {
    "grad_periodic": dict(nx=32, ny=8, nz=5, recon="grad", ambient_fx=1e-4, xwidth=3.2, ywidth=0.8, zwidth=0.5),
    "grad_walls_x": dict(nx=32, ny=8, nz=4, periods_x=0, recon="grad", ambient_fz=2e-4, xwidth=3.2, ywidth=0.8,
                         zwidth=0.4),
    "grad_walls_y": dict(nx=32, ny=16, nz=4, periods_y=0, recon="grad", ambient_fx=2e-4, xwidth=3.2, ywidth=1.6,
                         zwidth=0.4),
    "grad_walls_xy": dict(nx=32, ny=16, nz=3, periods_x=0, periods_y=0, recon="grad", ambient_fz=1e-4,
                          xwidth=3.2, ywidth=1.6, zwidth=0.3),
    "grad_tall_walls_y": dict(nx=32, ny=24, nz=6, periods_y=0, recon="grad", ambient_fx=2e-4, xwidth=3.2,
                              ywidth=2.4, zwidth=0.6),
    "grad_tall_walls_x": dict(nx=16, ny=12, nz=5, periods_x=0, recon="grad", ambient_fz=2e-4, xwidth=1.6,
                              ywidth=1.2, zwidth=0.5),
    "grad_tall_tiles": dict(nx=64, ny=24, nz=40, periods_y=0, recon="grad", ambient_fx=1e-4, ambient_fz=-1e-4,
                            xwidth=6.4, ywidth=2.4, zwidth=4.0),
    "grad_tiles_xcd": dict(nx=64, ny=16, nz=80, periods_y=0, recon="grad", ambient_fx=2e-4, xwidth=6.4,
                           ywidth=1.6, zwidth=8.0),
}
# configurations only the two-pass kernels cover
TWOPASS_CASES = {
    "ragged": dict(nx=13, ny=6, nz=5, xwidth=1.3, ywidth=0.6, zwidth=0.5),
    "walls_z": dict(nx=6, ny=4, nz=10, periods_z=0, ambient_fy=2e-4, xwidth=0.6, ywidth=0.4, zwidth=1.0),
    "grad_ragged": dict(nx=8, ny=6, nz=5, recon="grad", ambient_fx=1e-4, xwidth=0.8, ywidth=0.6, zwidth=0.5),
}


def _one_step_parity(rng, case, kernel):
    text = configs.base(visc=0.02, CFL=0.5, **case)
    eng = _engine(text, kernel=kernel)
    ora = _oracle_for(eng, text)
    f0 = random_state(eng.shape, rng)
    eng.set_f(f0)
    ora.set_f(f0)
    eng.step(1)
    ora.step(1)
    assert _relerr(eng.get_f(), ora.get_f()) <= 1e-13
    # a few more steps stay within round-off growth
    eng.step(4)
    ora.step(4)
    assert _relerr(eng.get_f(), ora.get_f()) <= 1e-12


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
@pytest.mark.parametrize("case", list(CASES))
def test_fp64_one_step_matches_oracle(rng, case, kernel):
    _one_step_parity(rng, CASES[case], kernel)


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
@pytest.mark.parametrize("case", list(GRAD_CASES))
def test_fp64_gradient_reconstruction_matches_oracle(rng, case, kernel):
    _one_step_parity(rng, GRAD_CASES[case], kernel)


def test_auto_selects_fused_kernel_for_gradient_reconstruction():
    eng = dugks_amd.Engine(configs.base(visc=0.02, **GRAD_CASES["grad_walls_y"]))
    assert eng.time_steps(2)["launches_per_step"] == 1


@pytest.mark.parametrize("case", list(TWOPASS_CASES))
def test_fp64_one_step_matches_oracle_twopass_only(rng, case):
    _one_step_parity(rng, TWOPASS_CASES[case], KERNEL_TWOPASS)
    # the automatic choice falls back to the two-pass kernels here
    eng = dugks_amd.Engine(configs.base(visc=0.02, **TWOPASS_CASES[case]))
    assert eng.time_steps(1)["launches_per_step"] == 5
    with pytest.raises(dugks_amd.DugksError):
        _engine(configs.base(visc=0.02, **TWOPASS_CASES[case]), kernel=KERNEL_FUSED).step(1)


def test_auto_selects_fused_kernel():
    eng = dugks_amd.Engine(configs.base(visc=0.02, **CASES["walls_y"]))
    assert eng.time_steps(2)["launches_per_step"] == 1


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_cell_force_matches_oracle(rng, kernel):
    text = configs.base(nx=32, ny=8, nz=4, periods_y=0, visc=0.02, ambient_fx=2e-4,
                        xwidth=3.2, ywidth=0.8, zwidth=0.4)
    eng = _engine(text, kernel=kernel)
    ora = _oracle_for(eng, text)
    f0 = random_state(eng.shape, rng)
    G = 1e-4 * rng.standard_normal((3,) + f0.shape[:3])
    eng.set_f(f0)
    ora.set_f(f0)
    eng.set_cell_force(*G)  # added to the ambient force at centres
    ora.set_cell_force(*(G + np.array([2e-4, 0, 0]).reshape(3, 1, 1, 1)))
    eng.step(2)
    ora.step(2)
    assert _relerr(eng.get_f(), ora.get_f()) <= 1e-13


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
# fused tiles: 32 x 4 (ny = 16), and for FP32 storage 32 x 12 (ny = 24)
@pytest.mark.parametrize("nx,ny", [(32, 16), (64, 16), (32, 24), (64, 24)])
# The FP16 modes are checked population by population against the rounding
# of their codec (tol = None), on top of the FP32 arithmetic of the decode,
# the step and the encode (twice the FP32 step's bar):
# - FP16 deviation storage rounds d = f / w - rho0 (about 3 c.u here) at
#   FP16: round to nearest moves f by at most w 2^-11 |d| (half an FP16 ulp,
#   relative), or w 2^-25 where d is subnormal;
# - FP16 non-equilibrium storage rounds only s = (f - feq(rho, u)) 1024 / w,
#   rho and u stay FP32: at most 2^-11 |f - feq|.
@pytest.mark.parametrize("precision,tol", [(PREC_FP32, 2e-6), (PREC_FP16S, None), (PREC_FP16N, None)],
                         ids=["fp32", "fp16", "fp16n"])
@pytest.mark.parametrize("recon", ["tri", "grad"])
def test_reduced_precision_one_step(rng, precision, tol, kernel, nx, ny, recon):
    text = configs.base(nx=nx, ny=ny, nz=4, periods_y=0, visc=0.02, ambient_fx=2e-4,
                        xwidth=nx / 10, ywidth=ny / 10, zwidth=0.4, recon=recon)
    eng = _engine(text, precision=precision, kernel=kernel)
    ora = _oracle_for(eng, text)
    f0 = random_state(eng.shape, rng)
    eng.set_f(f0)
    ora.set_f(eng.get_f())  # start from the stored (rounded) state
    eng.step(1)
    ora.step(1)
    got, want = eng.get_f(), ora.get_f()
    if tol is not None:
        assert _relerr(got, want) <= tol
        return
    if precision == PREC_FP16S:
        dev = want / W - eng.params["rho0"]
        codec = W * np.maximum(2.0 ** -11 * np.abs(dev), 2.0 ** -25)
    else:
        rho, j = moments(want)
        codec = 2.0 ** -11 * np.abs(want - feq(rho, j / rho[..., None]))
    bound = codec + 4e-6 * np.abs(want).max()
    assert np.all(np.abs(got - want) <= bound), float((np.abs(got - want) / bound).max())
    # the bound is not slack: the FP16 rounding is visible above the FP32 part
    assert _relerr(got, want) > 2e-6


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_macro_and_sums_match_oracle(rng, kernel):
    text = configs.base(nx=32, ny=8, nz=4, periods_x=0, visc=0.02, ambient_fy=3e-4, xwidth=0.8,
                        ywidth=0.8, zwidth=0.4)
    eng = _engine(text, kernel=kernel)
    ora = _oracle_for(eng, text)
    f0 = random_state(eng.shape, rng)
    eng.set_f(f0)
    ora.set_f(f0)
    for a, b in zip(eng.macro(), ora.macro()):
        assert np.abs(a - b).max() <= 1e-15 * max(1, np.abs(b).max())
    s = eng.global_sums()
    rho, j = moments(f0)
    assert abs(s[0] - rho.sum()) <= 1e-13 * rho.sum()
    assert np.allclose(s[1:4], j.sum((0, 1, 2)), rtol=1e-12, atol=1e-13)
    assert abs(s[7] - 0.8 * 0.8 * 0.4) < 1e-14
    eng.step(1)  # exercises the selected kernel


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_laminar_channel_startup_against_series(kernel):
    # 8 x 64 x 8 (32 x 64 x 8 for the fused kernel), walls at y = +-1,
    # Uc = 0.1, Re = 100, from rest, 20000 steps
    nx = 8 if kernel == KERNEL_TWOPASS else 32
    text = configs.laminar_channel(nsteps=20000, nx=nx, xwidth=nx * 2 / 64)
    eng = _engine(text, kernel=kernel)
    n = 20000
    eng.step(n)
    rho, ux, uy, uz = eng.macro()
    t = n * eng.params["dt"]
    nu = 1e-3
    eta = ((np.arange(64) + 0.5) * (2 / 64) - 1)
    s = 1 - eta ** 2
    for m in range(400):
        k = 2 * m + 1
        s = s - 32 * (-1) ** m / (k ** 3 * math.pi ** 3) * np.cos(k * math.pi * eta / 2) * \
            math.exp(-k * k * math.pi ** 2 * nu * t / 4)
    prof = ux[:, :, :].mean(axis=(0, 2)) / 0.1
    assert abs(t * nu - 0.221) < 0.01
    assert np.abs(prof - s).max() < 2e-3
    assert np.abs(uy).max() < 1e-12 and np.abs(uz).max() < 1e-12


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_laminar_channel_matches_oracle_over_2000_steps(kernel):
    nx = 8 if kernel == KERNEL_TWOPASS else 32
    text = configs.laminar_channel(nx=nx, xwidth=nx * 2 / 64)
    eng = _engine(text, kernel=kernel)
    ora = _oracle_for(eng, text)
    ora.set_f(eng.get_f())
    eng.step(2000)
    ora.step(2000)
    assert _relerr(eng.get_f(), ora.get_f()) <= 1e-12


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_mass_and_momentum_input_at_scale(kernel):
    # 64^3 Taylor-Green with a uniform force: mass conserved, momentum grows
    # by dt * M * F per step (closed form of the scheme)
    F = np.array([1e-5, -2e-5, 3e-5])
    text = configs.taylor_green(n=64, ambient_fx=F[0], ambient_fy=F[1], ambient_fz=F[2])
    eng = _engine(text, kernel=kernel)
    s0 = eng.global_sums()
    eng.step(10)
    s1 = eng.global_sums()
    assert abs(s1[0] - s0[0]) / s0[0] < 1e-13
    dj = s1[1:4] - s0[1:4]
    expect = 10 * eng.params["dt"] * s0[0] * F
    assert np.allclose(dj, expect, rtol=1e-8)


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_restart_continues_bitwise(tmp_path, kernel):
    text = configs.turbulent_channel(nx=32, ny=24, nz=16, flow_rate=0.0)
    a = _engine(text, kernel=kernel)
    a.step(5)
    a.write_dump(5, str(tmp_path))
    a.step(5, istep0=6)
    fa = a.get_f()
    b = _engine(text, kernel=kernel)
    b.read_dump(5, str(tmp_path))
    b.step(5, istep0=6)
    assert (b.get_f() == fa).all()
    # the dump is the raw reference layout
    raw = np.fromfile(os.path.join(tmp_path, "ddf00000005.dat"), dtype=np.float64)
    assert raw.size == 32 * 24 * 16 * 19


def test_restart_from_config(tmp_path):
    text = configs.base(nx=8, ny=8, nz=8, case="tg_eq", u0=0.05)
    a = _engine(text)
    a.step(3)
    a.write_dump(7, str(tmp_path))
    b = dugks_amd.Engine(configs.base(nx=8, ny=8, nz=8, case="general", restart=1, nrestart=7),
                         base_dir=str(tmp_path))
    assert (b.get_f() == a.get_f()).all()


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_flow_rate_control_holds_bulk_velocity(kernel):
    text = configs.turbulent_channel(nx=32, ny=32, nz=16)
    eng = _engine(text, kernel=kernel)
    h = eng.params["hdt"]
    s0 = eng.global_sums()
    F0 = eng.ambient_force()
    ub0 = s0[4] / s0[7] - h * F0[0]  # momentum bulk velocity <j/rho>
    eng.step(300)
    s1 = eng.global_sums()
    F = eng.ambient_force()
    assert abs(s1[4] / s1[7] - h * F[0] - ub0) < 1e-6 * ub0
    assert F[0] > 0 and F[1] == 0 and F[2] == 0


def test_outputs_written(tmp_path):
    eng = _engine(configs.base(nx=8, ny=6, nz=4, periods_x=0, case="channel", ambient_fz=1e-4))
    eng.step(2)
    eng.write_outputs(2, str(tmp_path))
    names = sorted(os.listdir(tmp_path))
    assert names == ["diag.dat", "profiles-2.dat", "s2X00000002.tec", "s2Y00000002.tec",
                     "s2Z00000002.tec"]
    rows = open(os.path.join(tmp_path, "profiles-2.dat")).read().splitlines()
    assert rows[0] == "00000002" and len(rows) == 1 + 8
    rho, ux, uy, uz = eng.macro()
    first = [float(v) for v in rows[1].split()]
    assert abs(first[2] - uz[:, :, 0].mean()) < 1e-7 * abs(uz).max()


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_perturb_forcing_matches_oracle(rng, kernel):
    # the reference's streak forcing (perturb = 1) recomputed every step on
    # the device against the oracle's field applied as the centre force
    from pyoracle import perturb_force
    text = configs.base(nx=32, ny=8, nz=6, periods_y=0, visc=0.02, ambient_fx=2e-4, perturb=1,
                        xwidth=3.2, ywidth=0.8, zwidth=0.6)
    eng = _engine(text, kernel=kernel)
    ora = _oracle_for(eng, text)
    f0 = random_state(eng.shape, rng)
    eng.set_f(f0)
    ora.set_f(f0)
    p = eng.params
    for step in (37, 38):
        force = perturb_force(eng.shape, (3.2, 0.8, 0.6), (p["dx"], p["dy"], p["dz"]), p["dt"], p["visc"],
                              step, (2e-4, 0.0, 0.0))
        assert np.abs(force[1]).max() > 0  # the forcing is on at these steps
        ora.set_cell_force(*force)
        eng.step(1, istep0=step)
        ora.step(1)
    assert _relerr(eng.get_f(), ora.get_f()) <= 1e-13


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32], ids=["fp64", "fp32"])
def test_device_statistics_match_fields(rng, precision):
    # statistics_serial (src/output.cu:223-334) on the device in FP64 against
    # the same sums taken in numpy from the macroscopic fields: profiles
    # along the wall axis (y here) and the global mean / min / max
    text = configs.turbulent_channel(nx=32, ny=24, nz=8, precision="fp64", flow_rate=0.0)
    eng = _engine(text, precision=precision, kernel=KERNEL_FUSED)
    eng.step(3)
    prof, diag = eng.statistics()
    rho, ux, uy, uz = eng.macro()  # [k][j][i]
    assert prof.shape == (24, 9)
    mean = lambda a: a.mean(axis=(0, 2))
    U, V, Wv = mean(ux), mean(uy), mean(uz)
    expect = np.stack([U, V, Wv, mean(ux * uz) - U * Wv, mean(ux * uy) - U * V, mean(uy * uz) - V * Wv,
                       mean(ux * ux) - U * U, mean(uy * uy) - V * V, mean(uz * uz) - Wv * Wv], axis=1)
    scale = np.abs(expect).max(axis=0)
    assert np.all(np.abs(prof - expect) <= 1e-10 * scale + 1e-18)
    fields = (rho, ux, uy, uz)
    assert np.allclose(diag[0:4], [a.mean() for a in fields], rtol=1e-12, atol=1e-15)
    assert np.array_equal(diag[4:8], [a.min() for a in fields])
    assert np.array_equal(diag[8:12], [a.max() for a in fields])


def test_driver_runs_reference_config(tmp_path):
    # the C++ driver (the reference's main.cu over the C ABI) on a short
    # channel config: steps, outputs on the noutput2 schedule (and at the
    # first step), a dump at ndump that holds the same state as the engine
    # stepping the same config
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "dugks-gpu_amd", "lib", "dugks_main")
    base = configs.base(nx=32, ny=12, nz=8, periods_y=0, visc=0.02, ambient_fx=2e-4, xwidth=3.2,
                        ywidth=1.2, zwidth=0.8, case="channel", nsteps=6, noutput2=3, ndump=6, nprint=2)
    (tmp_path / "config.txt").write_text(base)
    r = subprocess.run([exe, str(tmp_path / "config.txt"), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Main loop time" in r.stdout
    names = set(os.listdir(tmp_path))
    assert {"diag.dat", "profiles-2.dat", "ddf00000006.dat"} <= names
    diag = open(tmp_path / "diag.dat").read().split("\n")
    assert [l.split()[0] for l in diag if l.strip()] == ["00000001", "00000003", "00000006"]
    # the engine stepping the same config reaches the same dumped state
    eng = dugks_amd.Engine(base)
    eng.step(6, istep0=1)
    dumped = dugks_amd.dump_read(str(tmp_path / "ddf00000006.dat"), eng.ncell * 19).reshape(eng.fshape)
    assert np.array_equal(dumped, eng.get_f())


# mirror x_a -> L - x_a of one axis on mirror-symmetric meshes; the mirrored
# force component changes sign
MIRROR_CASES = {
    "walls_y_tall": (1, dict(nx=32, ny=24, nz=6, periods_y=0, mesh_y="tanh:0.85", dt_rule="min", ambient_fx=2e-4,
                             ambient_fz=1e-4, xwidth=3.2, ywidth=2.0, zwidth=0.6)),
    "walls_x": (0, dict(nx=32, ny=8, nz=4, periods_x=0, mesh_x="tanh:0.8", dt_rule="min", ambient_fy=1e-4,
                        ambient_fz=2e-4, xwidth=2.0, ywidth=0.8, zwidth=0.4)),
    "periodic_z": (2, dict(nx=32, ny=24, nz=5, ambient_fx=1e-4, ambient_fz=-2e-4, xwidth=3.2, ywidth=2.4,
                           zwidth=0.5)),
    # gradient reconstruction: the signs of its tangential differences
    "grad_walls_y_tall": (1, dict(nx=32, ny=24, nz=6, periods_y=0, recon="grad", ambient_fx=2e-4,
                                  ambient_fz=1e-4, xwidth=3.2, ywidth=2.4, zwidth=0.6)),
    "grad_walls_x": (0, dict(nx=32, ny=8, nz=4, periods_x=0, recon="grad", ambient_fy=1e-4, ambient_fz=2e-4,
                             xwidth=3.2, ywidth=0.8, zwidth=0.4)),
    "grad_periodic_z": (2, dict(nx=32, ny=24, nz=5, recon="grad", ambient_fx=1e-4, ambient_fz=-2e-4, xwidth=3.2,
                                ywidth=2.4, zwidth=0.5)),
}


@pytest.mark.parametrize("case", list(MIRROR_CASES))
@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_mirror_equivariance(rng, case, kernel):
    # step(R f) = R step(f): the tiles of the fused kernel map to mirrored
    # tiles, and every +/-c weight, wall ghost and bounce-back is exercised
    # on both sides
    axis, kw = MIRROR_CASES[case]
    key = "ambient_f" + "xyz"[axis]
    kwm = dict(kw)
    if key in kwm:
        kwm[key] = -kwm[key]
    a = _engine(configs.base(visc=0.02, **kw), kernel=kernel)
    b = _engine(configs.base(visc=0.02, **kwm), kernel=kernel)
    f0 = random_state(a.shape, rng)
    a.set_f(f0)
    a.step(3)
    b.set_f(reflect_state(f0, axis))
    b.step(3)
    assert _relerr(reflect_state(a.get_f(), axis), b.get_f()) <= 1e-13


@pytest.mark.parametrize("kernel", KERNELS, ids=["twopass", "fused"])
def test_fp16n_storage_round_trip(rng, kernel):
    # FP16 non-equilibrium storage keeps rho and u = j/rho in FP32 and rounds
    # s = (f - feq(rho, u)) 1024 / w to FP16: a state read back is within
    # half an FP16 ulp of s (2^-11 |f - feq|, relative) of the state written,
    # plus FP32 terms, population by population.  The rounding does not
    # preserve moments, so rho and j read back are within the sum of those
    # bounds, not FP32-exact; storing the read-back state again moves it by
    # at most the same rounding (the codec is not idempotent: rho and u of
    # the read-back state differ from the stored ones).
    text = configs.base(nx=32, ny=12, nz=4, periods_y=0, visc=0.02, ambient_fx=2e-4,
                        xwidth=3.2, ywidth=1.2, zwidth=0.4)
    eng = _engine(text, precision=PREC_FP16N, kernel=kernel)

    def bound(f):
        rho, j = moments(f)
        return np.abs(f - feq(rho, j / rho[..., None])) * 2.0 ** -11 + 1e-6 * W

    f0 = random_state(eng.shape, rng)
    eng.set_f(f0)
    f1 = eng.get_f()
    b0 = bound(f0)
    assert np.all(np.abs(f1 - f0) <= b0)
    r0, j0 = moments(f0)
    r1, j1 = moments(f1)
    assert np.all(np.abs(r1 - r0) <= b0.sum(-1))
    assert np.all(np.abs(j1 - j0) <= np.einsum("...q,qa->...a", b0, np.abs(C)))
    # the rounding is visible: FP32 storage would keep rho to ~1e-7
    assert np.abs(r1 - r0).max() > 1e-6
    eng.set_f(f1)
    assert np.all(np.abs(eng.get_f() - f1) <= bound(f1))
