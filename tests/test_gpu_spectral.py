"""DHIT spectral statistics on the device (stat_te of the reference,
src/fortran/dhit/src/stat.f90:5-205) against closed forms.

Velocity fields made of a few Fourier modes are loaded as f = feq(1, u), so
the engine's output velocity is u itself.  Their shell spectrum, dissipation
2 nu sum k^2 E(k), the derived scales and the derivative skewness and
flatness follow in closed form (the grid averages of trigonometric
polynomials below the Nyquist mode are exact).  The dissipation of an
evolved Taylor-Green state is checked against 2 nu <S:S> from the same
field's derivatives taken in numpy.
"""
import math
import os

import numpy as np
import pytest

from conftest import feq
import dugks_amd
from dugks_amd import PREC_FP64, configs

pytestmark = pytest.mark.gpu

L = 2 * math.pi


def _engine(n, visc=0.01, **kw):
    text = configs.base(nx=n, ny=n, nz=n, xwidth=L, ywidth=L, zwidth=L, visc=visc, **kw)
    return dugks_amd.Engine(text, precision=PREC_FP64, device=0)


def _grid(n):
    x = (np.arange(n) + 0.5) * (L / n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")  # [k][j][i]
    return X, Y, Z


def _load(eng, u):
    nz, ny, nx = u[0].shape
    eng.set_f(feq(np.ones((nz, ny, nx)), np.stack(u, -1)))


def test_shell_spectrum_and_scales_of_fourier_modes():
    n, nu = 32, 0.01
    A, B, C, D, E2 = 0.03, 0.02, 0.01, 0.015, 0.004
    X, Y, Z = _grid(n)
    # u_x: modes (0, 2, 0) and (1, 0, 0), (2, 0, 0); u_y: (0, 0, 3); u_z: (1, 1, 1)
    ux = A * np.cos(2 * Y) + D * np.sin(X) + E2 * np.sin(2 * X + 0.3)
    uy = B * np.sin(3 * Z)
    uz = C * np.cos(X + Y + Z)
    eng = _engine(n, visc=nu)
    _load(eng, (ux, uy, uz))
    st, spec = eng.spectral_statistics(nbins=n + 1)
    expect = np.zeros(n + 1)
    expect[2] += A * A / 4 + C * C / 4 + E2 * E2 / 4   # |kappa| = 2, sqrt(3) -> shell 2
    expect[3] += B * B / 4
    expect[1] += D * D / 4
    assert np.abs(spec - expect).max() <= 1e-15
    etot = expect.sum()
    eps = 2 * nu * (4 * A * A / 4 + 1 * D * D / 4 + 4 * E2 * E2 / 4 + 9 * B * B / 4 + 3 * C * C / 4)
    assert st["etotal"] == pytest.approx(etot, rel=1e-12)
    assert st["te"] == pytest.approx(etot, rel=1e-12)  # no mean flow
    assert st["eps"] == pytest.approx(eps, rel=1e-12)
    up = math.sqrt(2 / 3 * etot)
    assert st["uprime"] == pytest.approx(up, rel=1e-12)
    assert st["eta"] == pytest.approx((nu ** 3 / eps) ** 0.25, rel=1e-12)
    assert st["vk"] == pytest.approx((nu * eps) ** 0.25, rel=1e-12)
    assert st["tk"] == pytest.approx(math.sqrt(nu / eps), rel=1e-12)
    lam = math.sqrt(15 * nu * up * up / eps)
    assert st["taylor"] == pytest.approx(lam, rel=1e-12)
    assert st["re_lambda"] == pytest.approx(up * lam / nu, rel=1e-12)
    assert st["L"] == pytest.approx(up ** 3 / eps, rel=1e-12)
    assert st["kmax_eta"] == pytest.approx((L / 2 - 1.5 * L / n) * st["eta"], rel=1e-12)
    # longitudinal derivatives in closed form on the grid
    dudx = D * np.cos(X) + 2 * E2 * np.cos(2 * X + 0.3)
    dvdy = np.zeros_like(X)
    dwdz = -C * np.sin(X + Y + Z)
    m = [sum((d ** p).sum() for d in (dudx, dvdy, dwdz)) / (3 * n ** 3) for p in (2, 3, 4)]
    assert abs(m[1]) > 1e-9  # the (1, 0, 0) + (2, 0, 0) pair makes du/dx skewed
    assert st["skewness"] == pytest.approx(m[1] / m[0] ** 1.5, rel=1e-9)
    assert st["flatness"] == pytest.approx(m[2] / m[0] ** 2, rel=1e-11)


def test_mean_flow_is_excluded_from_the_spectrum():
    n = 16
    X, Y, Z = _grid(n)
    U0 = 0.05
    ux = U0 + 0.01 * np.sin(Y)
    eng = _engine(n)
    _load(eng, (ux, np.zeros_like(X), np.zeros_like(X)))
    st, spec = eng.spectral_statistics()
    assert st["te"] == pytest.approx(0.5 * (U0 ** 2 + 0.01 ** 2 / 2), rel=1e-12)
    assert st["etotal"] == pytest.approx(0.01 ** 2 / 4, rel=1e-12)
    assert spec[0] == 0 and spec[1] == pytest.approx(0.01 ** 2 / 4, rel=1e-12)


def test_dissipation_of_evolved_taylor_green_against_strain_rate():
    # 64^3 Taylor-Green vortex after 200 steps: eps from the shell sums
    # against 2 nu <S:S> of the same field (numpy FFT derivatives: equal up to
    # the small divergence of the weakly compressible flow) and against
    # central differences (second order)
    eng = dugks_amd.Engine(configs.taylor_green(n=64, u0=0.05, Re=400.0), precision=PREC_FP64, device=0)
    eng.step(200)
    st, _ = eng.spectral_statistics()
    _, ux, uy, uz = eng.macro()
    n = 64
    k = np.fft.fftfreq(n, d=1.0 / n)  # integer wavenumbers, L = 2 pi
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    kv = (kx, ky, kz)

    def grad(f):
        fh = np.fft.fftn(f)
        return [np.real(np.fft.ifftn(1j * kk * fh)) for kk in kv]

    g = [grad(c) for c in (ux, uy, uz)]  # g[a][b] = du_a / dx_b
    ss = sum((0.5 * (g[a][b] + g[b][a])) ** 2 for a in range(3) for b in range(3)).mean()
    div2 = ((g[0][0] + g[1][1] + g[2][2]) ** 2).mean()
    nu = eng.params["visc"]
    # 2 nu <S:S> = nu <|grad u|^2> + nu <(div u)^2> for a periodic field
    assert st["eps"] == pytest.approx(2 * nu * ss - nu * div2, rel=1e-10)
    assert st["eps"] == pytest.approx(2 * nu * ss, rel=1e-3)
    h = L / n
    cd = lambda a, ax: (np.roll(a, -1, axis=ax) - np.roll(a, 1, axis=ax)) / (2 * h)
    gc = [[cd(c, 2 - b) for b in range(3)] for c in (ux, uy, uz)]
    ssc = sum((0.5 * (gc[a][b] + gc[b][a])) ** 2 for a in range(3) for b in range(3)).mean()
    assert st["eps"] == pytest.approx(2 * nu * ssc, rel=2e-2)
    # skewness and flatness of du/dx, dv/dy, dw/dz from the same numpy
    # derivatives (the real part of the complex inverse transform, as the
    # reference takes it); the evolved vortex has a negative skewness
    d = [g[0][0], g[1][1], g[2][2]]
    m = [sum((x ** p).mean() for x in d) / 3 for p in (2, 3, 4)]
    assert st["skewness"] == pytest.approx(m[1] / m[0] ** 1.5, rel=1e-9)
    assert st["flatness"] == pytest.approx(m[2] / m[0] ** 2, rel=1e-9)
    assert st["skewness"] < -0.05


def test_spectral_output_files(tmp_path):
    n = 16
    X, Y, Z = _grid(n)
    eng = _engine(n)
    _load(eng, (0.01 * np.sin(Y), 0.02 * np.cos(Z), 0.01 * np.sin(X)))
    eng.write_spectral_outputs(5, str(tmp_path))
    eng.write_spectral_outputs(7, str(tmp_path))
    names = sorted(os.listdir(tmp_path))
    assert names == ["skew-flat.dat", "spe000000005.dat", "spe000000007.dat", "stat_te.dat",
                     "stat_te_spectral.dat"]
    st, spec = eng.spectral_statistics()
    lines = open(tmp_path / "stat_te.dat").read().splitlines()
    assert len(lines) == 2 and lines[0].split()[0] == "5"
    assert float(lines[0].split()[2]) == pytest.approx(st["te"], rel=1e-6)
    rows = [l.split() for l in open(tmp_path / "spe000000005.dat").read().splitlines()]
    assert [int(r[0]) for r in rows] == list(range(n + 1))
    assert np.allclose([float(r[1]) for r in rows], spec, rtol=1e-6, atol=1e-20)
    sp = open(tmp_path / "stat_te_spectral.dat").read().splitlines()[0].split()
    assert len(sp) == 13 and float(sp[4]) == 0.0
    assert float(sp[2]) == pytest.approx(st["etotal"], rel=1e-6)
    assert float(sp[3]) == pytest.approx(st["eps"], rel=1e-6)


def test_spectral_statistics_need_a_periodic_cube():
    eng = dugks_amd.Engine(configs.base(nx=16, ny=16, nz=16, periods_y=0), precision=PREC_FP64, device=0)
    with pytest.raises(dugks_amd.DugksError):
        eng.spectral_statistics()
    eng = dugks_amd.Engine(configs.base(nx=16, ny=16, nz=16, xwidth=1.0, ywidth=2.0), precision=PREC_FP64,
                           device=0)
    with pytest.raises(dugks_amd.DugksError):
        eng.spectral_statistics()
