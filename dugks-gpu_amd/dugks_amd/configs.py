"""config.txt texts of the workloads this engine is measured and tested on.

Keys are the reference's (src/numerics.cu:70-99, src/parallel.cu:95-97) plus
the engine's own (precision, recon, kernel, mesh_y, dt_rule, parse_real, u0,
flow_rate, retau, mean_profile, seed, ...).  Reals are parsed as double here
(parse_real=double) so the stated values are exact.
"""
import math


def _text(d):
    return "".join(f"{k}={v}\n" for k, v in d.items())


def base(**kw):
    """Small periodic test domain; keyword arguments override keys."""
    return _text(_base(**kw))


def _base(**kw):
    d = dict(case="rest", CFL=0.5, xwidth=1.0, ywidth=1.0, zwidth=1.0, nx=8, ny=8, nz=8,
             istep0=1, nsteps=0, noutput2=1000000, noutput3=1000000, ndump=1000000, restart=0,
             nrestart=0, perturb=0, visc=0.01, ambient_fx=0.0, ambient_fy=0.0, ambient_fz=0.0,
             periods_x=1, periods_y=1, periods_z=1, parse_real="double", precision="fp64")
    d.update(kw)
    return d


def laminar_channel(precision="fp64", nsteps=20000, **kw):
    """8 x 64 x 8 laminar channel, walls at y = +-1, centreline 0.1, Re = 100,
    CFL 0.5, from rest at rho = 1."""
    H, Uc, Re = 1.0, 0.1, 100.0
    nu = Uc * H / Re
    d = _base(case="rest", nx=8, ny=64, nz=8, xwidth=8 * 2 * H / 64, ywidth=2 * H,
             zwidth=8 * 2 * H / 64, visc=nu, ambient_fx=2 * nu * Uc / H ** 2, periods_y=0,
             nsteps=nsteps, precision=precision, CFL=0.5)
    d.update(kw)
    return _text(d)


def taylor_green(n=128, precision="fp64", nsteps=1000, u0=0.05, Re=1600.0, **kw):
    """Periodic Taylor-Green vortex on n^3, L = 2 pi, u0, Re = u0 (L/2pi) / nu."""
    L = 2 * math.pi
    d = _base(case="tg_eq", nx=n, ny=n, nz=n, xwidth=L, ywidth=L, zwidth=L, u0=u0,
             visc=u0 * (L / (2 * math.pi)) / Re, nsteps=nsteps, precision=precision, CFL=0.5)
    d.update(kw)
    return _text(d)


def turbulent_channel(nx=128, ny=128, nz=128, precision="fp64", nsteps=10000, retau=180.0,
                      visc=3.0e-5, alpha=0.85, CFL=0.5, flow_rate="auto", **kw):
    """Re_tau = 180 channel on a tanh-stretched wall-normal (y) mesh, domain
    4 pi x 2 x 4 pi/3, constant flow rate, synthetic log-law + Fourier state."""
    H = 1.0
    utau = retau * visc / H
    d = _base(case="turb_channel", nx=nx, ny=ny, nz=nz, xwidth=4 * math.pi, ywidth=2 * H,
             zwidth=4 * math.pi / 3, visc=visc, ambient_fx=utau ** 2 / H, periods_y=0,
             mesh_y=f"tanh:{alpha}", dt_rule="min", retau=retau, pert_amp=0.1, kmin=1, kmax=8,
             flow_rate=flow_rate, flow_axis=0, seed=20240611, nsteps=nsteps, precision=precision,
             CFL=CFL)
    d.update(kw)
    return _text(d)


def large_channel(precision="fp32", nsteps=100, **kw):
    """256 x 192 x 256 stretched Re_tau = 180 channel (same recipe)."""
    return turbulent_channel(nx=256, ny=192, nz=256, precision=precision, nsteps=nsteps, **kw)


def adjoint_channel(precision="fp64", Re=2000.0, nx=64, ny=64, nz=32, **kw):
    """Transitional channel 64 x 64 x 32, walls at y, seeded perturbed
    laminar state at Re = Uc H / nu = 2000, Uc = 0.1."""
    H, Uc = 1.0, 0.1
    nu = Uc * H / Re
    d = _base(case="turb_channel", mean_profile="laminar", u0=Uc, nx=nx, ny=ny, nz=nz,
             xwidth=4 * math.pi, ywidth=2 * H, zwidth=4 * math.pi / 3, visc=nu,
             ambient_fx=2 * nu * Uc / H ** 2, periods_y=0, dt_rule="min", pert_amp=0.1, kmin=1,
             kmax=4, seed=777, precision=precision, CFL=0.5)
    d.update(kw)
    return _text(d)
