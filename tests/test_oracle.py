"""Pins the FP64 CPU oracle (oracle/dugks_oracle.cpp) before it is trusted.

The reference ships no tests, golden vectors or known-answer tests for the
step (its only data fixture, the tanh mesh, is checked in test_host.py).  The
oracle is pinned here by closed forms and invariants each chosen so that a
plausible transcription mistake fails one of them, and by an independent
whole-array implementation (numpy_dugks.py).
"""
import math

import numpy as np
import pytest

from conftest import C, OPP, RT, W, feq, moments, random_state, reflect_state
from pyoracle import Oracle
import numpy_dugks


def test_lattice_closed_forms():
    # isotropy of D3Q19 up to fourth order; opposite velocities
    assert abs(W.sum() - 1) < 1e-15
    assert np.abs(W @ C).max() < 1e-15
    assert np.allclose(np.einsum("q,qa,qb->ab", W, C, C), RT * np.eye(3), atol=1e-15)
    d = np.eye(3)
    iso = RT ** 2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) +
                     np.einsum("ad,bc->abcd", d, d))
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", W, C, C, C, C), iso, atol=1e-15)
    assert (C[OPP] == -C).all() and (OPP[OPP] == np.arange(19)).all()


def test_coefficient_identities():
    o = Oracle((4, 4, 4), (1, 1, 1), (1, 1, 1), dt=0.013, visc=0.0071)
    c = o.coefficients()
    dt = 0.013
    assert abs(c["tau"] - 0.0071 / RT) < 1e-16
    assert abs(c["hdt"] - dt / 2) < 1e-18
    # with 2h = dt the centre and face weights are partitions of unity
    assert abs(c["bw1"] + c["bw2"] - 1) < 1e-15
    assert abs(c["fw1"] + c["fw2"] - 1) < 1e-15
    # force weight: bw0 + h bw2 = 3 dt/4 (total momentum input dt*rho*F per step)
    assert abs(c["bw0"] + c["hdt"] * c["bw2"] - 0.75 * dt) < 1e-16
    # face stage is the same relaxation over h/2: fw0 = tau h/(2tau+h)
    assert abs(c["fw0"] - c["tau"] * c["fw2"]) < 1e-16


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)])
def test_rest_and_uniform_equilibrium_are_fixed_points(periodic):
    shape = (6, 5, 7)
    o = Oracle(shape, periodic, (1.2, 1.0, 1.4), dt=0.05, visc=0.02)
    nx, ny, nz = shape
    u = np.zeros((nz, ny, nx, 3))
    if all(periodic):
        u[..., 0], u[..., 1], u[..., 2] = 0.04, -0.03, 0.02
    f0 = feq(np.full((nz, ny, nx), 1.01), u)
    o.set_f(f0)
    o.step(4)
    assert np.abs(o.get_f() - f0).max() < 1e-15


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1), (1, 1, 0)])
def test_mass_conservation(rng, periodic):
    # the force is tangential to the walls: a wall-normal force gives the
    # wall face a normal velocity (h/2) F and the scheme then leaks mass
    shape = (8, 6, 5)
    F = [0.0 if not periodic[a] else (1e-4, 0, 2e-4)[a] for a in range(3)]
    o = Oracle(shape, periodic, (1.0, 0.75, 0.625), dt=0.04, visc=0.01, famb=F)
    f0 = random_state(shape, rng)
    o.set_f(f0)
    o.step(6)
    m0, m1 = f0.sum(), o.get_f().sum()
    assert abs(m1 - m0) / m0 < 5e-15


def test_momentum_input_closed_form(rng):
    # periodic: total momentum grows by exactly dt * sum(rho) * F per step
    shape = (6, 6, 6)
    F = np.array([2e-4, -1e-4, 3e-4])
    dt = 0.06
    o = Oracle(shape, (1, 1, 1), (1, 1, 1), dt=dt, visc=0.03, famb=F)
    f0 = random_state(shape, rng)
    o.set_f(f0)
    o.step(1)
    r0, j0 = moments(f0)
    r1, j1 = moments(o.get_f())
    dj = j1.sum((0, 1, 2)) - j0.sum((0, 1, 2))
    assert np.allclose(dj, dt * r0.sum() * F, rtol=1e-10, atol=1e-16)
    # the same holds with a per-cell centre force
    Fc = 1e-4 * rng.standard_normal((3,) + f0.shape[:3])
    o2 = Oracle(shape, (1, 1, 1), (1, 1, 1), dt=dt, visc=0.03)
    o2.set_f(f0)
    o2.set_cell_force(*Fc)
    o2.step(1)
    r2, j2 = moments(o2.get_f())
    expect = dt * np.einsum("kji,akji->a", r0, Fc)
    assert np.allclose(j2.sum((0, 1, 2)) - j0.sum((0, 1, 2)), expect, rtol=1e-9, atol=1e-16)


def test_shear_wave_viscosity():
    # u_x = A sin(k y): amplitude decays as exp(-nu k^2 t)
    ny, L, nu = 32, 2 * math.pi, 0.01
    dx = L / ny
    dt = 0.5 * dx / math.sqrt(2)
    o = Oracle((2, ny, 2), (1, 1, 1), (2 * dx, L, 2 * dx), dt=dt, visc=nu)
    y = (np.arange(ny) + 0.5) * dx
    u = np.zeros((2, ny, 2, 3))
    u[..., 0] = 0.01 * np.sin(y)[None, :, None]
    o.set_f(feq(np.ones((2, ny, 2)), u))
    n = 400
    o.step(n)
    _, ux, _, _ = o.macro()
    A = (ux[0, :, 0] * np.sin(y)).sum() / (np.sin(y) ** 2).sum()
    rate = -math.log(A / 0.01) / (n * dt)
    assert abs(rate / nu - 1) < 0.01


def startup_series(eta, t, nu, H=1.0, nterms=400):
    s = 1 - eta ** 2
    for m in range(nterms):
        k = 2 * m + 1
        s = s - 32 * (-1) ** m / (k ** 3 * math.pi ** 3) * np.cos(k * math.pi * eta / 2) * \
            math.exp(-k * k * math.pi ** 2 * nu * t / (4 * H * H))
    return s


@pytest.mark.parametrize("wall_axis", [0, 1, 2])
def test_poiseuille_startup_series(wall_axis):
    # channel of half-width 1 started from rest under a constant force; the
    # walls may lie on any axis (the reference only allows x, quirk Q1)
    n, H, nu, Uc = 24, 1.0, 0.0125, 0.05
    d = 2 * H / n
    F = 2 * nu * Uc / H ** 2
    dt = 0.5 * d / math.sqrt(2)
    shape = [2, 2, 2]
    shape[wall_axis] = n
    per = [1, 1, 1]
    per[wall_axis] = 0
    width = [s * d for s in shape]
    sa = (wall_axis + 1) % 3
    fam = [0.0, 0.0, 0.0]
    fam[sa] = F
    o = Oracle(shape, per, width, dt=dt, visc=nu, famb=fam)
    nx, ny, nz = shape
    o.set_f(feq(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx, 3))))
    nst = int(round(0.3 * H * H / nu / dt))
    o.step(nst)
    _, ux, uy, uz = o.macro()
    u = (ux, uy, uz)[sa]
    prof = np.moveaxis(u, 2 - wall_axis, 0)[:, 0, 0]
    eta = ((np.arange(n) + 0.5) * d - H) / H
    err = np.abs(prof / Uc - startup_series(eta, nst * dt, nu)).max()
    assert err < 5e-3


def _rotate_state(f, perm):
    """Apply the lattice symmetry x_new[b] = x_old[perm[b]] (axis permutation)."""
    q_new = []
    for q in range(19):
        c = C[q]
        cn = np.array([c[perm[b]] for b in range(3)])
        q_new.append(int(np.where((C == cn).all(1))[0][0]))
    # f[z,y,x] axes 0:z,1:y,2:x ; permute spatial axes accordingly
    sp = f.transpose(2, 1, 0, 3)  # (x, y, z, q)
    sp = sp.transpose(perm[0], perm[1], perm[2], 3)
    out = np.empty_like(sp)
    out[..., q_new] = sp
    return out.transpose(2, 1, 0, 3)


@pytest.mark.parametrize("perm", [(1, 0, 2), (2, 1, 0), (1, 2, 0)])
def test_axis_permutation_equivariance(rng, perm):
    n = 6
    F = np.array([1e-4, 2e-4, -3e-4])
    f0 = random_state((n, n, n), rng)
    a = Oracle((n, n, n), (1, 1, 1), (1, 1, 1), dt=0.05, visc=0.02, famb=F)
    a.set_f(f0)
    a.step(2)
    b = Oracle((n, n, n), (1, 1, 1), (1, 1, 1), dt=0.05, visc=0.02, famb=F[list(perm)])
    b.set_f(_rotate_state(f0, perm))
    b.step(2)
    assert np.abs(_rotate_state(a.get_f(), perm) - b.get_f()).max() < 1e-15


def test_wall_axis_permutation_equivariance(rng):
    # an x-wall run permuted equals the y-wall run (the reference only has x)
    shape = (8, 4, 4)
    F = (0.0, 0.0, 2e-4)
    f0 = random_state(shape, rng)
    a = Oracle(shape, (0, 1, 1), (1.0, 0.5, 0.5), dt=0.03, visc=0.02, famb=F)
    a.set_f(f0)
    a.step(3)
    perm = (1, 0, 2)
    b = Oracle((4, 8, 4), (1, 0, 1), (0.5, 1.0, 0.5), dt=0.03, visc=0.02, famb=(0.0, 0.0, 2e-4))
    b.set_f(_rotate_state(f0, perm))
    b.step(3)
    assert np.abs(_rotate_state(a.get_f(), perm) - b.get_f()).max() < 1e-15


# (mirrored axis, periodic flags, spacing of the mirrored axis or None,
# force): mirror-symmetric meshes only; the mirrored force component changes
# sign
REFLECT_CASES = {
    "periodic_x": (0, (1, 1, 1), None, (1e-4, 2e-4, -3e-4)),
    "periodic_z": (2, (1, 1, 1), None, (1e-4, 2e-4, -3e-4)),
    "walls_y_stretched": (1, (1, 0, 1), np.array([0.05, 0.12, 0.33, 0.33, 0.12, 0.05]), (2e-4, 0.0, 1e-4)),
    "walls_x_stretched": (0, (0, 1, 1), np.array([0.1, 0.2, 0.3, 0.3, 0.2, 0.1]), (0.0, 1e-4, 2e-4)),
}


@pytest.mark.parametrize("case", list(REFLECT_CASES))
def test_mirror_equivariance(rng, case):
    # step(R f) = R step(f) for a mirror R of a lattice axis: pins the sign
    # and side of every +/-c weight (stretched tables included), the wall
    # ghosts and bounce-back at both walls, and the force terms
    axis, per, d, F = REFLECT_CASES[case]
    shape = (6, 6, 6)
    spacing = [None, None, None]
    if d is not None:
        spacing[axis] = d / d.sum()
    F = np.array(F)
    Fm = F.copy()
    Fm[axis] *= -1
    f0 = random_state(shape, rng)
    a = Oracle(shape, per, (1, 1, 1), dt=0.02, visc=0.02, famb=F, spacing=spacing)
    a.set_f(f0)
    a.step(3)
    b = Oracle(shape, per, (1, 1, 1), dt=0.02, visc=0.02, famb=Fm, spacing=spacing)
    b.set_f(reflect_state(f0, axis))
    b.step(3)
    assert np.abs(reflect_state(a.get_f(), axis) - b.get_f()).max() < 1e-14


@pytest.mark.parametrize("stretched,recon", [(False, "tri"), (True, "tri"), (False, "grad")])
def test_reconstruction_exact_for_linear_fields(rng, stretched, recon):
    # with walls on every axis the ghost extrapolation is linear, so a linear
    # f-bar field is reconstructed exactly at the foot point x_face - c h, by
    # the trilinear scheme (calc_fx_tri) and by the midpoint + gradient scheme
    # (calc_fx, uniform axes), for all 19 populations
    shape = (5, 6, 4)
    width = (1.0, 1.3, 0.8)
    dt = 0.02
    h = dt / 2
    spacing = [None, None, None]
    if stretched:
        w = np.array([0.05, 0.12, 0.3, 0.35, 0.31, 0.17])
        spacing[1] = w / w.sum() * width[1]
    o = Oracle(shape, (0, 0, 0), width, dt=dt, visc=0.01, spacing=spacing, recon=recon)
    faces_x = []
    for a in range(3):
        d = spacing[a] if spacing[a] is not None else np.full(shape[a], width[a] / shape[a])
        faces_x.append(np.concatenate([[0.0], np.cumsum(d)]))
    centres = [0.5 * (fx[1:] + fx[:-1]) for fx in faces_x]
    A = rng.standard_normal(19)
    B = rng.standard_normal((19, 3))
    X, Y, Z = np.meshgrid(centres[0], centres[1], centres[2], indexing="ij")
    pos = np.stack([X, Y, Z], -1).transpose(2, 1, 0, 3)  # (z,y,x,3)
    fbar = A + pos @ B.T
    faces = o.reconstruct(fbar)
    nx, ny, nz = shape
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                p = pos[k, j, i]
                idx = (i, j, k)
                for s in range(6):
                    a, side = s // 2, (1 if s % 2 else -1)
                    xf = p.copy()
                    xf[a] = faces_x[a][idx[a] + (1 if side > 0 else 0)]
                    for q in range(1, 19):
                        expect = A[q] + (xf - C[q] * h) @ B[q]
                        assert abs(faces[k, j, i, s, q] - expect) < 1e-12
                    if not stretched:
                        expect0 = A[0] + xf @ B[0]
                        assert abs(faces[k, j, i, s, 0] - expect0) < 1e-12


def test_gradient_minus_trilinear_on_quadratic_fields(rng):
    # f-bar = sum_b K_b x_b^2 (per population).  Along the face normal both
    # schemes make the same error (the normal lerp / midpoint-slope pair is
    # exact to the same quadratic term); along a tangential axis b the
    # trilinear lerp between the cell row and the upwind row has error
    # h d_b |c_b| - h^2 c_b^2 at the foot point and the central-difference
    # gradient has error -h^2 c_b^2, so tri - grad = h d_b |c_b| summed over
    # the tangential axes.  A wrong sign, a one-sided difference or d instead
    # of 2 d in a tangential gradient of calc_fx breaks this.
    shape = (6, 7, 5)
    width = (1.2, 1.4, 0.5)
    d = [w / n for w, n in zip(width, shape)]
    dt = 0.03
    h = dt / 2
    centres = [(np.arange(n) + 0.5) * dd for n, dd in zip(shape, d)]
    X, Y, Z = np.meshgrid(*centres, indexing="ij")
    pos = np.stack([X, Y, Z], -1).transpose(2, 1, 0, 3)  # (z,y,x,3)
    K = rng.standard_normal((19, 3))
    fbar = 2.0 + (pos ** 2) @ K.T
    faces = {}
    for recon in ("tri", "grad"):
        o = Oracle(shape, (0, 0, 0), width, dt=dt, visc=0.01, recon=recon)
        faces[recon] = o.reconstruct(fbar)
    diff = faces["tri"] - faces["grad"]
    nx, ny, nz = shape
    # cells whose stencils (one row beyond each tangential neighbour) stay
    # away from the wall ghosts, whose extrapolation is not exact here
    sl = (slice(1, nz - 1), slice(1, ny - 1), slice(1, nx - 1))
    for s in range(6):
        a = s // 2
        expect = sum(K[:, b] * h * d[b] * np.abs(C[:, b]) for b in range(3) if b != a)
        got = diff[sl][..., s, :]
        assert np.abs(got - expect).max() < 1e-13, s
    assert np.abs(diff).max() > 1e-4


def test_gradient_and_trilinear_agree_on_normal_populations(rng):
    shape = (5, 5, 5)
    f0 = random_state(shape, rng)
    tri = Oracle(shape, (1, 1, 1), (1, 1, 1), dt=0.05, visc=0.02, recon="tri")
    grd = Oracle(shape, (1, 1, 1), (1, 1, 1), dt=0.05, visc=0.02, recon="grad")
    ft, fg = tri.reconstruct(f0), grd.reconstruct(f0)
    for s in range(6):
        a = s // 2
        normal = [q for q in range(19) if C[q, (a + 1) % 3] == 0 and C[q, (a + 2) % 3] == 0]
        assert np.abs(ft[..., s, normal] - fg[..., s, normal]).max() < 1e-15
    # and they differ for oblique populations (the schemes are distinct)
    assert np.abs(ft - fg).max() > 1e-6


@pytest.mark.parametrize("periodic,F,cell_force", [
    ((1, 1, 1), (0, 0, 0), False),
    ((1, 1, 1), (1e-4, -2e-4, 3e-4), True),
    ((0, 1, 1), (0, 0, 2e-4), False),
    ((1, 0, 1), (2e-4, 0, 0), True),
    ((1, 1, 0), (0, 1e-4, 0), False),
    ((0, 0, 1), (0, 0, 1e-4), False),
])
def test_oracle_matches_independent_implementation(rng, periodic, F, cell_force):
    shape = (7, 6, 5)
    width = (1.4, 1.2, 1.0)
    dx = [w / n for w, n in zip(width, shape)]
    dt, nu = 0.03, 0.015
    f0 = random_state(shape, rng)
    o = Oracle(shape, periodic, width, dt=dt, visc=nu, famb=F)
    o.set_f(f0)
    Fcell = None
    if cell_force:
        Fcell = np.asarray(F).reshape(3, 1, 1, 1) + 1e-4 * rng.standard_normal((3,) + f0.shape[:3])
        o.set_cell_force(*Fcell)
    o.step(1)
    ref = numpy_dugks.step(f0, periodic, dx, dt, nu, F, Fcell)
    assert np.abs(o.get_f() - ref).max() / np.abs(ref).max() < 1e-13


def _oracle_case(rng, periodic=(1, 0, 1), recon="tri", spacing=(None, None, None)):
    shape = (5, 6, 4)
    width = (1.0, 1.2, 0.8)
    o = Oracle(shape, periodic, width, dt=0.02, visc=0.01, famb=(2e-4, 0, 1e-4),
               spacing=spacing, recon=recon)
    f0 = random_state(shape, rng)
    Fc = np.array([2e-4, 0, 1e-4]).reshape(3, 1, 1, 1) + 1e-4 * rng.standard_normal((3,) + f0.shape[:3])
    return o, f0, Fc


@pytest.mark.parametrize("variant", ["walls", "periodic", "grad", "stretched"])
def test_oracle_adjoint_matches_central_differences(rng, variant):
    kw = {}
    if variant == "periodic":
        kw = dict(periodic=(1, 1, 1))
    elif variant == "grad":
        kw = dict(periodic=(1, 1, 1), recon="grad")
    elif variant == "stretched":
        w = np.array([0.1, 0.2, 0.3, 0.3, 0.2, 0.1])
        kw = dict(spacing=(None, w / w.sum() * 1.2, None))
    o, f0, Fc = _oracle_case(rng, **kw)
    lam = rng.standard_normal(f0.shape)
    o.set_f(f0)
    o.set_cell_force(*Fc)
    lam_prev, gF = o.adjoint_step(lam)

    def J(f, Fcell):
        o.set_f(f)
        o.set_cell_force(*Fcell)
        o.step(1)
        return float((o.get_f() * lam).sum())

    # central differences in random directions; J is O(1) so its round-off
    # gives a difference-quotient noise floor of about 1e-9 at eps = 1e-6
    eps = 1e-6
    for trial in range(3):
        v = rng.standard_normal(f0.shape) * 1e-2
        fd = (J(f0 + eps * v, Fc) - J(f0 - eps * v, Fc)) / (2 * eps)
        an = float((lam_prev * v).sum())
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-8
        w = rng.standard_normal(Fc.shape)
        fd = (J(f0, Fc + eps * w) - J(f0, Fc - eps * w)) / (2 * eps)
        an = float((gF.transpose(3, 0, 1, 2) * w).sum())
        assert abs(an) > 1e-4
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-8


def test_perturb_force_structure():
    # the streak forcing (calc_perturb_force): ambient outside the band
    # 3 < i < nx/4 + 2, the pure driving force FF = 2 u*^2 / Lx at step 0
    # (amplitude A0 ~ sin(2 pi step dt / 2000) vanishes), and a perturbation
    # that scales with that sine
    from pyoracle import perturb_force
    shape, width, d = (32, 8, 6), (3.2, 0.8, 0.6), (0.1, 0.1, 0.1)
    dt, visc, famb = 0.5, 0.02, (1e-4, 2e-4, -3e-4)
    band = np.zeros(32, bool)
    band[3:9] = True  # 1-based i = 4..9
    fx0, fy0, fz0 = perturb_force(shape, width, d, dt, visc, 0, famb)
    ustar = 2 * 180 * visc / width[0]
    FF = 2 * ustar ** 2 / width[0]
    for f, a in zip((fx0, fy0, fz0), famb):
        assert np.all(f[:, :, ~band] == a)
    assert np.all(fx0[:, :, band] == 0) and np.all(fy0[:, :, band] == 0)
    assert np.allclose(fz0[:, :, band], FF, rtol=1e-15)
    s1, s2 = 100, 333
    r = math.sin(2 * math.pi * s2 * dt / 2000) / math.sin(2 * math.pi * s1 * dt / 2000)
    f1 = perturb_force(shape, width, d, dt, visc, s1, famb)
    f2 = perturb_force(shape, width, d, dt, visc, s2, famb)
    for a in range(2):
        assert np.abs(f1[a][:, :, band]).max() > 0
        assert np.allclose(f2[a][:, :, band], r * f1[a][:, :, band], rtol=1e-12, atol=1e-20)
    assert np.allclose(f2[2][:, :, band] - FF, r * (f1[2][:, :, band] - FF), rtol=1e-12, atol=1e-12 * FF)


def test_perturb_force_spatial_pattern():
    # calc_perturb_force (src/dugks.cu:1086-1126) in the band: with
    # X = 2 pi (i - 3) dx / (Lx / 4), Y = 2 pi (j - 1/2) dy / Ly,
    # Z = 2 pi (k - 1/2) dz / Lz (1-based cell indices),
    #   F'x ~ (1 - cos X) cos(3 Y) cos(2 Z),
    #   F'y ~ sin X cos(3 Y) sin(2 Z),
    #   F'z - FF ~ sin X sin(3 Y) cos(2 Z).
    # Projected on the Fourier modes of the cell-centred y and z grids each
    # component holds one mode, beta = 3 in y and gamma = 2 in z, with the
    # stated cos / sin phase (a real or an imaginary coefficient); along x it
    # follows its band profile, even or odd about X = pi.
    from pyoracle import perturb_force
    nx, ny, nz = 64, 24, 16
    width = (6.4, 2.4, 1.6)
    d = tuple(w / n for w, n in zip(width, (nx, ny, nz)))
    dt, visc, step = 0.5, 0.02, 250
    f = perturb_force((nx, ny, nz), width, d, dt, visc, step, (1e-4, 0.0, 0.0))
    ustar = 2 * 180 * visc / width[0]
    FF = 2 * ustar ** 2 / width[0]
    band = np.arange(3, nx // 4 + 1)  # 0-based i of 1-based 4 .. nx/4 + 1
    comps = [f[0][:, :, band], f[1][:, :, band], f[2][:, :, band] - FF]  # [k][j][m]
    jj, kk = np.arange(ny) + 0.5, np.arange(nz) + 0.5
    ey = np.exp(-2j * np.pi * np.outer(np.arange(ny), jj) / ny)  # [mode][j]
    ez = np.exp(-2j * np.pi * np.outer(np.arange(nz), kk) / nz)
    # (y phase, z phase) of each component: 0 = cos (real), 1 = sin (imaginary)
    phases = [(0, 0), (0, 1), (1, 0)]
    X = 2 * np.pi * (band + 1 - 3) / (nx // 4)
    shapes = [1 - np.cos(X), np.sin(X), np.sin(X)]
    for comp, (py, pz), sx in zip(comps, phases, shapes):
        c = np.einsum("my,nz,zyx->nmx", ey, ez, comp)  # [z mode][y mode][x]
        scale = np.abs(c).max()
        assert scale > 0
        modes = np.zeros(c.shape[:2], bool)
        modes[np.ix_([2, nz - 2], [3, ny - 3])] = True
        assert np.abs(c[~modes]).max() <= 1e-12 * scale
        c32 = c[2, 3]
        real, imag = np.abs(c32.real).max(), np.abs(c32.imag).max()
        if py + pz == 1:   # one sine factor: imaginary coefficient
            assert real <= 1e-12 * scale and imag > 0.1 * scale
        else:              # cos cos or sin sin: real coefficient
            assert imag <= 1e-12 * scale and real > 0.1 * scale
        # x profile: the mode's coefficient is proportional to the band shape
        prof = c32.real if py + pz != 1 else c32.imag
        nzero = np.abs(sx) > 1e-9
        ratio = prof[nzero] / sx[nzero]
        assert np.abs(ratio - ratio[0]).max() <= 1e-12 * np.abs(ratio[0])
    # parity about the band centre X = pi (m <-> nx/4 - m, 1-based m = i - 3)
    m = band + 1 - 3
    q = nx // 4
    pairs = [(a, int(np.where(m == q - m[a])[0][0])) for a in range(len(m)) if 2 <= m[a] <= q - 2]
    tol = [1e-12 * np.abs(c).max() for c in comps]
    for a, b in pairs:
        assert np.abs(comps[0][:, :, a] - comps[0][:, :, b]).max() <= tol[0]
        assert np.abs(comps[1][:, :, a] + comps[1][:, :, b]).max() <= tol[1]
        assert np.abs(comps[2][:, :, a] + comps[2][:, :, b]).max() <= tol[2]
