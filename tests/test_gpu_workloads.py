"""The benchmark workloads at full size on the GPU.

- Taylor-Green vortex, 128^3 periodic, u0 = 0.05, Re = 1600, 1000 steps in
  FP64 and FP32: mass conserved, kinetic energy decays monotonically, and
  the FP32 kinetic energy, dissipation rate and dissipated energy agree
  with FP64.
- Re_tau = 180 channel, 128^3, tanh-stretched wall-normal mesh, constant
  flow rate, 10 000 steps: density, velocity, mean-velocity, rms and
  Reynolds-stress profile errors of each reduced-precision storage mode
  against FP64 (one test per statistic), and the bulk velocity held by the
  controller.
- Large channel, 256 x 192 x 256, 200 steps per mode: finite, mass conserved,
  bulk velocity held, FP32 against FP64.
- Adjoint channel 64 x 64 x 32, T = 400, at constant force and at constant
  flow rate: FP64 and FP32 gradients against central differences.

Field errors are relative to the FP64 maxima (each profile to its own
maximum), the definitions bench.py reports.
"""
import math

import numpy as np
import pytest

import dugks_amd
from dugks_amd import PREC_FP16N, PREC_FP16S, PREC_FP32, PREC_FP64, configs

pytestmark = pytest.mark.gpu

# field-error bar of the reduced-precision modes against FP64
FIELD_TOL = 1e-3


def _engine(text, precision):
    return dugks_amd.Engine(text, precision=precision, device=0)


def _velocity_stats(eng):
    """u, v, w fields and the wall-normal (y) profiles of mean u, the rms
    velocities and <u'v'>, FP64 on the host."""
    rho, ux, uy, uz = eng.macro()  # [k][j][i]
    mean = lambda a: a.mean(axis=(0, 2))
    U, V, Wm = mean(ux), mean(uy), mean(uz)
    prof = {"U": U,
            "urms": np.sqrt(np.maximum(mean(ux * ux) - U * U, 0)),
            "vrms": np.sqrt(np.maximum(mean(uy * uy) - V * V, 0)),
            "wrms": np.sqrt(np.maximum(mean(uz * uz) - Wm * Wm, 0)),
            "uv": mean(ux * uy) - U * V}
    return {"rho": rho, "u": (ux, uy, uz), "prof": prof}


def _field_errors(ref, got):
    du2 = sum((g - r) ** 2 for g, r in zip(got["u"], ref["u"]))
    umax = float(np.sqrt(sum(r * r for r in ref["u"]).max()))
    out = {"rho_max": float(np.abs(got["rho"] - ref["rho"]).max() / np.abs(ref["rho"]).max()),
           "u_max": float(np.sqrt(du2.max()) / umax)}
    for k, r in ref["prof"].items():
        out[k] = float(np.abs(got["prof"][k] - r).max() / np.abs(r).max())
    return out


def _bulk(eng):
    """Momentum bulk velocity <j/rho> along x (the controlled quantity)."""
    s = eng.global_sums()
    return s[4] / s[7] - eng.params["hdt"] * eng.ambient_force()[0]


def _dissipation(ux, uy, uz, L, nu):
    """eps = 2 nu <S:S> of a periodic cube field [k][j][i] (central differences)."""
    n = ux.shape[0]
    h = L / n
    d = lambda a, ax: (np.roll(a, -1, axis=ax) - np.roll(a, 1, axis=ax)) / (2 * h)
    g = [[d(c, 2 - b) for b in range(3)] for c in (ux, uy, uz)]  # g[a][b] = du_a/dx_b
    ss = sum((0.5 * (g[a][b] + g[b][a])) ** 2 for a in range(3) for b in range(3))
    return 2 * nu * float(ss.mean())


def test_taylor_green_128_fp32_against_fp64():
    steps, every = 1000, 20
    E, mass, eps = {}, {}, {}
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32)):
        text = configs.taylor_green(n=128, precision=name, nsteps=steps)
        eng = _engine(text, prec)
        m0 = eng.global_sums()[0]
        series = [eng.loss(dugks_amd.LOSS_KINETIC_ENERGY)]
        for k in range(steps // every):
            eng.step(every, istep0=1 + k * every)
            series.append(eng.loss(dugks_amd.LOSS_KINETIC_ENERGY))
        mass[name] = abs(eng.global_sums()[0] - m0) / m0
        E[name] = np.array(series)
        _, ux, uy, uz = eng.macro()
        eps[name] = _dissipation(ux, uy, uz, 2 * math.pi, eng.params["visc"])
        eng.close()
    assert mass["fp64"] <= 1e-12
    assert mass["fp32"] <= 1e-5
    for name in E:
        assert np.all(np.diff(E[name]) < 0), name  # monotone decay
    # the vortex has decayed measurably over the run (0.3 % at Re = 1600)
    assert E["fp64"][-1] < 0.998 * E["fp64"][0]
    assert np.abs(E["fp32"] - E["fp64"]).max() / np.abs(E["fp64"]).max() <= FIELD_TOL
    assert abs(eps["fp32"] - eps["fp64"]) / eps["fp64"] <= FIELD_TOL
    # the energy each run dissipated: FP32 fell 0.9 % short of FP64 while the
    # equilibrium's class products were rounded twice in FP32 (8e-5 now)
    d64, d32 = E["fp64"][0] - E["fp64"][-1], E["fp32"][0] - E["fp32"][-1]
    assert abs(d32 - d64) <= 1e-3 * d64, d32 / d64


@pytest.fixture(scope="module")
def channel_128():
    """FP64 reference run of the 128^3 Re_tau = 180 channel, 10 000 steps."""
    steps = 10000
    eng = _engine(configs.turbulent_channel(precision="fp64", nsteps=steps), PREC_FP64)
    ub0 = _bulk(eng)
    eng.step(steps, istep0=1)
    stats = _velocity_stats(eng)
    ub = _bulk(eng)
    eng.close()
    return steps, stats, ub0, ub


_CHANNEL_RUNS = {}


def _channel_run(channel_128, mode):
    """errors of a 10 000-step run in storage mode `mode` against FP64 (cached
    per module), and its bulk velocity before and after."""
    if mode not in _CHANNEL_RUNS:
        steps, ref, _, _ = channel_128
        prec = {"fp32": PREC_FP32, "fp16": PREC_FP16S, "fp16n": PREC_FP16N}[mode]
        eng = _engine(configs.turbulent_channel(precision=mode, nsteps=steps), prec)
        ub0 = _bulk(eng)
        eng.step(steps, istep0=1)
        got = _velocity_stats(eng)
        ub = _bulk(eng)
        eng.close()
        _CHANNEL_RUNS[mode] = (_field_errors(ref, got), ub0, ub)
    return _CHANNEL_RUNS[mode]


_METRICS = ["rho_max", "u_max", "U", "urms", "vrms", "wrms", "uv"]


@pytest.mark.parametrize("metric", _METRICS)
@pytest.mark.parametrize("mode", ["fp32", "fp16n"])
def test_channel_128_storage_modes_against_fp64(channel_128, mode, metric):
    _, _, ub0_ref, ub_ref = channel_128
    err, ub0, ub = _channel_run(channel_128, mode)
    # the controller holds the bulk velocity in both runs
    assert abs(ub_ref - ub0_ref) <= 1e-6 * ub0_ref
    assert abs(ub - ub0) <= 1e-5 * ub0
    assert err[metric] <= FIELD_TOL, err


# the deviation codec stores f/w - rho0 in FP16, which rounds the 3 c.u part
# of every population each step: its error grows far past the bar over
# 10 000 steps of this flow (u max 0.13, <u'v'> 0.83)
@pytest.mark.xfail(strict=True, reason="FP16 deviation storage rounds 3 c.u each step; "
                                       "known to exceed the 1e-3 field-error bar here")
def test_channel_128_fp16_deviation_storage_against_fp64(channel_128):
    err, _, _ = _channel_run(channel_128, "fp16")
    bad = {k: v for k, v in err.items() if not v <= FIELD_TOL}
    assert not bad, err


def _mass(eng):
    """sum rho dV (the conserved mass on a stretched mesh), FP64 on the host."""
    rho = eng.macro()[0]  # [k][j][i]
    dx, dy, dz = (eng.spacing(a) for a in range(3))
    return float(np.einsum("kji,k,j,i->", rho, dz, dy, dx))


def test_large_channel_256x192x256():
    steps = 200
    out = {}
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32), ("fp16n", PREC_FP16N), ("fp16", PREC_FP16S)):
        eng = _engine(configs.large_channel(precision=name, nsteps=steps), prec)
        m0 = _mass(eng)
        ub0 = _bulk(eng)
        eng.step(steps, istep0=1)
        s1 = eng.global_sums()
        assert all(math.isfinite(v) for v in s1), name
        # mass: round-off of the arithmetic type (FP16 storage rounds each state)
        tol = {"fp64": 1e-12, "fp32": 1e-5, "fp16n": 1e-5, "fp16": 1e-3}[name]
        assert abs(_mass(eng) - m0) / m0 <= tol, name
        # the controller holds the bulk velocity of the state before its
        # storage rounding; FP16 deviation storage moves it by ~1e-4
        assert abs(_bulk(eng) - ub0) <= (1e-3 if name == "fp16" else 1e-5) * ub0, name
        if name != "fp16":
            out[name] = _velocity_stats(eng)
        eng.close()
    for name in ("fp32", "fp16n"):
        err = _field_errors(out["fp64"], out[name])
        bad = {k: v for k, v in err.items() if not v <= FIELD_TOL}
        assert not bad, (name, err)


@pytest.mark.parametrize("flow", [False, True], ids=["constant_force", "constant_flow_rate"])
def test_adjoint_channel_64x64x32_gradient(flow):
    # the adjoint workload at full size: transitional channel 64 x 64 x 32
    # (seeded perturbed laminar state, Re = 2000), J = time-mean wall shear
    # stress over T = 400 steps, control G ~ N(0, 1e-6) per cell, sqrt(T)
    # checkpoints, driven by a constant force or at constant flow rate (the
    # sweep then differentiates the controller; every run restarts it from
    # the same force).  The FP64 gradient against central differences of J
    # on 32 seeded random directions (bar 1e-4); the FP32 gradient against
    # the FP64 one and against the same differences.
    torch = pytest.importorskip("torch")
    T, ndirs = 400, 32
    dev = torch.device("cuda", 0)
    grads, Js = {}, {}
    fds = []
    kw = dict(flow_rate="auto", flow_axis=0) if flow else {}
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32)):
        eng = _engine(configs.adjoint_channel(precision=name, **kw), prec)
        nx, ny, nz = eng.shape
        f0 = torch.empty((19, nz, ny, nx), dtype=torch.float64, device=dev)
        eng.get_f_device(f0)
        F0 = eng.ambient_force()
        G = torch.from_numpy(1e-6 * np.random.default_rng(2024).standard_normal((3, nz, ny, nx))).to(dev)
        gd = torch.empty_like(G)
        eng.set_ambient_force(F0)
        eng.set_f_device(f0)
        eng.set_cell_force_device(G)
        Js[name], _ = eng.forward_adjoint(T, gd)
        grads[name] = gd.cpu().numpy()
        if name == "fp64":
            rng = np.random.default_rng(7)

            def run_J(ctrl):
                eng.set_ambient_force(F0)
                eng.set_f_device(f0)
                eng.set_cell_force_device(ctrl)
                return eng.forward_record(T)

            for _ in range(ndirs):
                v = 1e-6 * rng.standard_normal(G.shape)
                vd = torch.from_numpy(v).to(dev)
                fd = (run_J(G + vd) - run_J(G - vd)) / 2.0
                fds.append((fd, v))
            eng.release_sweep()
        eng.close()
    g64, g32 = grads["fp64"], grads["fp32"]
    err64 = [abs(fd - float((g64 * v).sum())) / abs(fd) for fd, v in fds]
    err32 = [abs(fd - float((g32 * v).sum())) / abs(fd) for fd, v in fds]
    assert max(err64) <= 1e-4, max(err64)
    assert max(err32) <= 1e-4, max(err32)
    assert np.abs(g32 - g64).max() <= 1e-4 * np.abs(g64).max()
    # J sums the wall shear stress, a difference of near-wall velocities: its
    # FP32 value is about 1e-4 from FP64 (measured 1.1e-4)
    assert abs(Js["fp32"] - Js["fp64"]) <= 1e-3 * abs(Js["fp64"])
