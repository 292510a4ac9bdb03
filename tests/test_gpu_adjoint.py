"""GPU adjoint (through the C ABI) against the FP64 CPU oracle adjoint and
central finite differences.

The oracle's adjoint is itself pinned by central differences in
test_oracle.py.  Here: one adjoint step vs the oracle on the same state and
cotangent; the loss functionals vs their definitions; the sqrt(T)-checkpointed
forward+adjoint vs central differences of J in random control directions and
vs the chained oracle adjoint.
"""
import numpy as np
import pytest

from conftest import C, random_state
import dugks_amd
from dugks_amd import KERNEL_FUSED, KERNEL_TWOPASS, PREC_FP16S, PREC_FP32, PREC_FP64, configs
from pyoracle import Oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _engine(text, precision=PREC_FP64, kernel=-1):
    return dugks_amd.Engine(text, precision=precision, kernel=kernel,
                            stream=torch.cuda.current_stream().cuda_stream)


def _oracle(eng, text):
    kv = dict(l.split("=", 1) for l in text.strip().splitlines())
    per = tuple(int(kv.get(k, "1")) for k in ("periods_x", "periods_y", "periods_z"))
    width = tuple(float(kv[k]) for k in ("xwidth", "ywidth", "zwidth"))
    famb = np.array([float(kv.get(k, "0")) for k in ("ambient_fx", "ambient_fy", "ambient_fz")])
    sp = [None if kv.get(f"mesh_{a}", "uniform") == "uniform" else eng.spacing(i)
          for i, a in enumerate("xyz")]
    p = eng.params
    return Oracle(eng.shape, per, width, dt=p["dt"], visc=p["visc"], famb=famb, spacing=sp), famb


def _soa(aos):  # (nz,ny,nx,q) -> (q,nz,ny,nx)
    return np.ascontiguousarray(np.moveaxis(aos, -1, 0))


def _aos(soa):
    return np.ascontiguousarray(np.moveaxis(soa, 0, -1))


CASES = {
    "periodic": dict(nx=32, ny=8, nz=4, xwidth=3.2, ywidth=0.8, zwidth=0.4, ambient_fx=1e-4),
    "walls_y": dict(nx=32, ny=16, nz=4, periods_y=0, ambient_fx=2e-4, xwidth=3.2, ywidth=1.6, zwidth=0.4),
    "walls_x_stretched": dict(nx=16, ny=8, nz=4, periods_x=0, mesh_x="tanh:0.8", dt_rule="min",
                              ambient_fz=2e-4, xwidth=2.0, ywidth=0.8, zwidth=0.4),
    "walls_z": dict(nx=8, ny=6, nz=10, periods_z=0, ambient_fy=2e-4, xwidth=0.8, ywidth=0.6, zwidth=1.0),
    # cases the fused adjoint-face pass covers (nx % 32 == 0, z periodic)
    "walls_x": dict(nx=32, ny=8, nz=4, periods_x=0, ambient_fz=2e-4, xwidth=3.2, ywidth=0.8, zwidth=0.4),
    # x walls across two tiles: the inner tile edges take lam' from the halo column
    "walls_x_tiles": dict(nx=64, ny=8, nz=6, periods_x=0, mesh_x="tanh:0.8", dt_rule="min", ambient_fz=2e-4,
                          xwidth=6.4, ywidth=0.8, zwidth=0.6),
    "tiles": dict(nx=64, ny=16, nz=80, periods_y=0, mesh_y="tanh:0.85", dt_rule="min", ambient_fx=2e-4,
                  xwidth=6.4, ywidth=2.0, zwidth=8.0),
    # 16 x 12 tiles
    "tall": dict(nx=32, ny=24, nz=10, periods_y=0, mesh_y="tanh:0.85", dt_rule="min", ambient_fx=2e-4,
                 xwidth=3.2, ywidth=2.0, zwidth=1.0),
    "tall_walls_x": dict(nx=16, ny=12, nz=6, periods_x=0, ambient_fz=2e-4, xwidth=1.6, ywidth=1.2, zwidth=0.6),
    "tall_periodic": dict(nx=32, ny=24, nz=6, xwidth=3.2, ywidth=2.4, zwidth=0.6, ambient_fx=1e-4),
}
FUSED_CASES = ("periodic", "walls_y", "walls_x", "walls_x_tiles", "tiles", "tall", "tall_walls_x", "tall_periodic")


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kernel", [-1, KERNEL_TWOPASS], ids=["auto", "twopass"])
@pytest.mark.parametrize("precision,tol", [(PREC_FP64, 1e-12), (PREC_FP32, 2e-5), (PREC_FP16S, 2e-5)],
                         ids=["fp64", "fp32", "fp16"])
def test_adjoint_step_matches_oracle(rng, case, kernel, precision, tol):
    # FP16 storage: the oracle steps from the decoded stored state
    if kernel == KERNEL_TWOPASS and case not in FUSED_CASES:
        pytest.skip("the automatic choice is the two-pass path here already")
    text = configs.base(visc=0.02, **CASES[case])
    eng = _engine(text, precision=precision, kernel=kernel)
    ora, famb = _oracle(eng, text)
    f0 = random_state(eng.shape, rng)
    eng.set_f(f0)
    f0 = eng.get_f()  # the stored (possibly rounded) state
    nx, ny, nz = eng.shape
    G = 1e-4 * rng.standard_normal((3, nz, ny, nx))
    eng.set_cell_force(*G)
    ora.set_f(f0)
    ora.set_cell_force(*(G + famb.reshape(3, 1, 1, 1)))
    lam = rng.standard_normal(f0.shape)
    lp_ref, g_ref = ora.adjoint_step(lam)
    dev = torch.device("cuda")
    lam_d = torch.from_numpy(_soa(lam)).to(dev)
    lp_d = torch.empty_like(lam_d)
    g_d = torch.zeros((3, nz, ny, nx), dtype=torch.float64, device=dev)
    eng.adjoint_step(lam_d, lp_d, g_d)
    torch.cuda.synchronize()
    lp = _aos(lp_d.cpu().numpy())
    g = np.moveaxis(g_d.cpu().numpy(), 0, -1)
    assert np.abs(lp - lp_ref).max() <= tol * np.abs(lp_ref).max()
    assert np.abs(g - g_ref).max() <= tol * np.abs(g_ref).max()


# configurations of the z-marching gather (nx % 32 == 0, ny % 8 == 0)
LAM32_CASES = {
    "channel": dict(nx=64, ny=24, nz=40, periods_y=0, mesh_y="tanh:0.85", dt_rule="min", ambient_fx=2e-4,
                    xwidth=6.4, ywidth=2.0, zwidth=4.0),
    "walls_x": dict(nx=64, ny=16, nz=20, periods_x=0, mesh_x="tanh:0.8", dt_rule="min", ambient_fz=2e-4,
                    xwidth=6.4, ywidth=1.6, zwidth=2.0),
    "periodic": dict(nx=32, ny=16, nz=8, xwidth=3.2, ywidth=1.6, zwidth=0.8, ambient_fx=1e-4),
}


@pytest.mark.parametrize("case", list(LAM32_CASES))
@pytest.mark.parametrize("precision", [PREC_FP32, PREC_FP16S], ids=["fp32", "fp16"])
def test_adjoint_step_f32_lambda_matches_f64_lambda(rng, case, precision):
    # the FP32 modes compute lam in FP32 arithmetic: with lam' exactly
    # representable in FP32, FP32 storage of lam' and lam changes no bit
    eng = _engine(configs.base(visc=0.02, **LAM32_CASES[case]), precision=precision)
    eng.set_f(random_state(eng.shape, rng))
    nx, ny, nz = eng.shape
    eng.set_cell_force(*(1e-4 * rng.standard_normal((3, nz, ny, nx))))
    dev = torch.device("cuda")
    lam32 = torch.from_numpy(rng.standard_normal((19, nz, ny, nx)).astype(np.float32)).to(dev)
    lam64 = lam32.double()
    out = []
    for lam in (lam64, lam32):
        lp = torch.empty_like(lam)
        g = torch.zeros((3, nz, ny, nx), dtype=torch.float64, device=dev)
        eng.adjoint_step(lam, lp, g)
        torch.cuda.synchronize()
        out.append((lp.double().cpu().numpy(), g.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.abs(out[0][0]).max() > 0


def test_adjoint_step_f32_lambda_unsupported_configurations():
    dev = torch.device("cuda")
    # FP64 arithmetic, and a grid the z-marching gather does not cover (nx % 32 != 0)
    wide = dict(nx=48, ny=36, nz=8, periods_y=0, ambient_fx=2e-4, xwidth=4.8, ywidth=2.0, zwidth=0.8)
    for text, prec in ((configs.base(visc=0.02, **LAM32_CASES["periodic"]), PREC_FP64),
                       (configs.base(visc=0.02, **wide), PREC_FP32)):
        eng = _engine(text, precision=prec)
        nx, ny, nz = eng.shape
        lam = torch.zeros((19, nz, ny, nx), dtype=torch.float32, device=dev)
        with pytest.raises(dugks_amd.DugksError):
            eng.adjoint_step(lam, torch.empty_like(lam))


def test_adjoint_step_rejects_overlapping_buffers(rng):
    eng = _engine(configs.base(visc=0.02, **LAM32_CASES["periodic"]), precision=PREC_FP32)
    nx, ny, nz = eng.shape
    lam = torch.zeros((19, nz, ny, nx), dtype=torch.float64, device=torch.device("cuda"))
    with pytest.raises(dugks_amd.DugksError):
        eng.adjoint_step(lam, lam)


def _wall_shear_grad(eng, famb, nu):
    """dJ/df of the mean wall shear stress (definition in include/dugks.h)."""
    nx, ny, nz = eng.shape
    h = eng.params["hdt"]
    w = eng.spacing(1)
    g = np.zeros((nz, ny, nx, 19))
    for jw in (0, ny - 1):
        coef = nu / (0.5 * w[jw] * 2.0 * nx * nz)
        g[:, jw, :, :] = coef * (C[:, 0] + h * famb[0])
    return g


def test_loss_functionals_match_definitions(rng):
    text = configs.adjoint_channel(nx=32, ny=16, nz=4)
    eng = _engine(text)
    f = eng.get_f()
    famb = eng.ambient_force()
    nu = eng.params["visc"]
    g = _wall_shear_grad(eng, famb, nu)
    dev = torch.device("cuda")
    gd = torch.empty((19,) + f.shape[:3], dtype=torch.float64, device=dev)
    J = eng.loss(dugks_amd.LOSS_WALL_SHEAR, gd)
    torch.cuda.synchronize()
    assert abs(J - (g * f).sum()) <= 1e-13 * abs(J)
    assert np.abs(_aos(gd.cpu().numpy()) - g).max() <= 1e-15 * np.abs(g).max()
    # kinetic energy 0.5 <rho |u|^2>, u = j/rho + h F
    K = eng.loss(dugks_amd.LOSS_KINETIC_ENERGY, gd)
    torch.cuda.synchronize()
    rho = f.sum(-1)
    u = np.einsum("kjiq,qa->kjia", f, C) / rho[..., None] + eng.params["hdt"] * famb
    w = eng.spacing(1)
    dV = (4 * np.pi / 32) * w[None, :, None] * (4 * np.pi / 3 / 4)
    Kref = (0.5 * rho * (u ** 2).sum(-1) * dV).sum() / (4 * np.pi * 2 * 4 * np.pi / 3)
    assert abs(K - Kref) <= 1e-12 * Kref


def _run_J(eng, f0, G, T, kind=dugks_amd.LOSS_WALL_SHEAR):
    eng.set_f(f0)
    eng.set_cell_force(*G)
    total = 0.0
    for n in range(T):
        eng.step(1, istep0=1 + n)
        total += eng.loss(kind)
    return total / T


def _adjoint_problem(rng, precision, nx=32, ny=16, nz=8, T=24):
    # transitional channel (seeded perturbed laminar state, Re = 2000),
    # wall-shear loss averaged over T steps, control G ~ N(0, 1e-6)
    text = configs.adjoint_channel(nx=nx, ny=ny, nz=nz)
    eng = _engine(text, precision=precision)
    f0 = eng.get_f()
    G = 1e-6 * np.random.default_rng(99).standard_normal((3, nz, ny, nx))
    return eng, f0, G, T


def _grad(eng, f0, G, T, interval=0):
    nx, ny, nz = eng.shape
    eng.set_f(f0)
    eng.set_cell_force(*G)
    gd = torch.empty((3, nz, ny, nx), dtype=torch.float64, device=torch.device("cuda"))
    J, peak = eng.forward_adjoint(T, gd, interval=interval)
    torch.cuda.synchronize()
    return J, gd.cpu().numpy(), peak


def test_forward_adjoint_matches_central_differences(rng):
    eng, f0, G, T = _adjoint_problem(rng, PREC_FP64)
    J, grad, _ = _grad(eng, f0, G, T)
    # recomputation from checkpoints is exact: every interval gives the same gradient
    for interval in (1, 5, T):
        Jk, gk, _ = _grad(eng, f0, G, T, interval)
        assert Jk == J and (gk == grad).all()
    assert abs(J - _run_J(eng, f0, G, T)) <= 1e-12 * abs(J)
    # perturbations of the control's own size (1e-6): the response is linear
    # to O(1e-12) there, and J's round-off stays far below the difference
    eps = 1.0
    for trial in range(4):
        v = rng.standard_normal(G.shape) * 1e-6
        jp = _run_J(eng, f0, G + eps * v, T)
        jm = _run_J(eng, f0, G - eps * v, T)
        fd = (jp - jm) / (2 * eps)
        an = float((grad * v).sum())
        assert abs(fd - an) <= 1e-6 * abs(an), (trial, fd, an)


@pytest.mark.parametrize("ny", [16, 24])
def test_fp32_gradient_against_fp64(rng, ny):
    # ny = 16: the z-marching gather with FP32 lam; ny = 24: 16 x 12 step tiles
    e64, f64, G, T = _adjoint_problem(rng, PREC_FP64, ny=ny)
    J64, g64, _ = _grad(e64, f64, G, T)
    e32, f32, _, _ = _adjoint_problem(rng, PREC_FP32, ny=ny)
    J32, g32, _ = _grad(e32, f32, G, T)
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    assert np.abs(g32 - g64).max() <= 1e-3 * np.abs(g64).max()


def test_forward_adjoint_matches_chained_oracle_adjoint(rng):
    text = configs.adjoint_channel(nx=16, ny=16, nz=4)
    eng = _engine(text)
    ora, famb = _oracle(eng, text)
    f0 = eng.get_f()
    nx, ny, nz = eng.shape
    G = 1e-6 * rng.standard_normal((3, nz, ny, nx))
    T = 10
    dev = torch.device("cuda")
    gd = torch.empty((3, nz, ny, nx), dtype=torch.float64, device=dev)
    g0 = torch.empty((19, nz, ny, nx), dtype=torch.float64, device=dev)
    eng.set_cell_force(*G)
    J, _ = eng.forward_adjoint(T, gd, interval=3, grad_f0=g0)
    torch.cuda.synchronize()
    # oracle: forward with stored states, then the chained one-step adjoints
    ora.set_cell_force(*(G + famb.reshape(3, 1, 1, 1)))
    states = [f0]
    ora.set_f(f0)
    for n in range(T):
        ora.step(1)
        states.append(ora.get_f())
    dJ = _wall_shear_grad(eng, famb, eng.params["visc"]) / T
    Jref = sum((dJ * s).sum() for s in states[1:])
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    lam = np.zeros_like(f0)
    gF = np.zeros((nz, ny, nx, 3))
    for n in range(T - 1, -1, -1):
        lam = lam + dJ
        ora.set_f(states[n])
        lam, gF = ora.adjoint_step(lam, gF)
    grad = np.moveaxis(gd.cpu().numpy(), 0, -1)
    assert np.abs(grad - gF).max() <= 1e-11 * np.abs(gF).max()
    assert np.abs(_aos(g0.cpu().numpy()) - lam).max() <= 1e-11 * np.abs(lam).max()


# ---- the adjoint through the flow-rate controller ----
#
# The same transitional channel driven at constant flow rate: after every step
# the controller sets the ambient force from the bulk velocity it reads, so
# dJ/dG and dJ/dF^0 include the controller's response.  Each run restarts the
# controller (set_ambient_force, then set_f) from the same force.


def _flow_problem(precision, nx=32, ny=16, nz=8, T=24, **kw):
    text = configs.adjoint_channel(nx=nx, ny=ny, nz=nz, flow_rate="auto", flow_axis=0, **kw)
    eng = _engine(text, precision=precision)
    G = 1e-6 * np.random.default_rng(99).standard_normal((3, nz, ny, nx))
    return eng, eng.get_f(), eng.ambient_force(), G, T


def _flow_start(eng, f0, F0, G):
    eng.set_ambient_force(F0)
    eng.set_f(f0)
    eng.set_cell_force(*G)


def _flow_grad(eng, f0, F0, G, T, kind=dugks_amd.LOSS_WALL_SHEAR, interval=0):
    nx, ny, nz = eng.shape
    _flow_start(eng, f0, F0, G)
    gd = torch.empty((3, nz, ny, nx), dtype=torch.float64, device=torch.device("cuda"))
    J = eng.forward_record(T, loss_kind=kind, interval=interval)
    F_end = eng.ambient_force()
    eng.adjoint_replay(gd)
    torch.cuda.synchronize()
    # the replay leaves the controller's force where the forward ended
    assert (eng.ambient_force() == F_end).all()
    return J, gd.cpu().numpy(), eng.sweep_force_gradient()


def _flow_J(eng, f0, F0, G, T, kind=dugks_amd.LOSS_WALL_SHEAR):
    _flow_start(eng, f0, F0, G)
    return eng.forward_record(T, loss_kind=kind)


@pytest.mark.parametrize("kind", [dugks_amd.LOSS_WALL_SHEAR, dugks_amd.LOSS_KINETIC_ENERGY],
                         ids=["wall_shear", "kinetic_energy"])
def test_flow_controlled_adjoint_matches_central_differences(rng, kind):
    eng, f0, F0, G, T = _flow_problem(PREC_FP64)
    J, grad, dF0 = _flow_grad(eng, f0, F0, G, T, kind)
    # the forward moved the force (the controller is active)
    assert eng.ambient_force()[0] != F0[0]
    # recomputation from checkpoints with the recorded forces is exact
    for interval in (1, 5, T):
        Jk, gk, dk = _flow_grad(eng, f0, F0, G, T, kind, interval)
        assert Jk == J and (gk == grad).all() and dk == dF0
    for trial in range(4):
        v = rng.standard_normal(G.shape) * 1e-6
        fd = (_flow_J(eng, f0, F0, G + v, T, kind) - _flow_J(eng, f0, F0, G - v, T, kind)) / 2
        an = float((grad * v).sum())
        # measured 8e-10 (wall shear), 3e-9 (kinetic energy)
        assert abs(fd - an) <= 1e-7 * abs(an), (trial, fd, an)
    # the initial ambient force along the flow axis (measured 3e-8 / 5e-8,
    # the O(d^2) truncation of the difference)
    d = 1e-3 * abs(F0[0])
    Fp, Fm = F0.copy(), F0.copy()
    Fp[0] += d
    Fm[0] -= d
    fd = (_flow_J(eng, f0, Fp, G, T, kind) - _flow_J(eng, f0, Fm, G, T, kind)) / (2 * d)
    assert abs(fd - dF0) <= 1e-6 * abs(dF0), (fd, dF0)


def test_flow_controlled_gradient_differs_from_constant_force(rng):
    # the controller's response is a large part of the gradient: the same
    # problem at constant force gives a gradient far outside the 1e-6 bar
    eng, f0, F0, G, T = _flow_problem(PREC_FP64)
    _, g_flow, _ = _flow_grad(eng, f0, F0, G, T)
    e2, f2, _, _ = _adjoint_problem(rng, PREC_FP64)
    assert (f2 == f0).all()
    _, g_const, _ = _grad(e2, f2, G, T)
    # (measured 0.13)
    assert np.abs(g_flow - g_const).max() > 1e-2 * np.abs(g_const).max()


@pytest.mark.parametrize("precision", [PREC_FP32, PREC_FP16S], ids=["fp32", "fp16"])
def test_flow_controlled_reduced_precision_gradient_against_fp64(rng, precision):
    # FP32 arithmetic with lam in FP32 (the z-marching gather's compact
    # records): against the FP64 sweep
    e64, f64, F64, G, T = _flow_problem(PREC_FP64)
    J64, g64, d64 = _flow_grad(e64, f64, F64, G, T)
    e32, f32, F32, _, _ = _flow_problem(precision)
    J32, g32, d32 = _flow_grad(e32, f32, F32, G, T)
    # measured: FP32 1.1e-6, FP16 deviation storage 1.6e-5 (gradient)
    tol = 1e-5 if precision == PREC_FP32 else 2e-4
    assert abs(J32 - J64) <= tol * abs(J64)
    assert np.abs(g32 - g64).max() <= tol * np.abs(g64).max()
    assert abs(d32 - d64) <= tol * abs(d64)


def test_flow_controlled_adjoint_unsupported_configurations():
    # FP32 without the FP32-lam form (ny % 8 != 0), and streak forcing
    eng, *_ = _flow_problem(PREC_FP32, ny=12)
    with pytest.raises(dugks_amd.DugksError):
        eng.forward_record(4)
    eng, *_ = _flow_problem(PREC_FP64, perturb=1)
    with pytest.raises(dugks_amd.DugksError):
        eng.forward_record(4)
    # no sweep with the controller replayed yet
    eng, *_ = _flow_problem(PREC_FP64)
    with pytest.raises(dugks_amd.DugksError):
        eng.sweep_force_gradient()


def test_autograd_function(rng):
    from dugks_amd.autograd import control_loss
    text = configs.adjoint_channel(nx=32, ny=16, nz=4)
    eng = _engine(text)
    nx, ny, nz = eng.shape
    dev = torch.device("cuda")
    f0 = torch.empty((19, nz, ny, nx), dtype=torch.float64, device=dev)
    eng.get_f_device(f0)
    G = (1e-6 * torch.randn((3, nz, ny, nx), dtype=torch.float64, device=dev)).requires_grad_(True)
    f0g = f0.clone().requires_grad_(True)
    eng.release_sweep()
    J = control_loss(G, eng, f0g, 8)
    # forward() ran only the forward sweep: checkpoints, no adjoint buffers yet
    fwd_bytes = eng.sweep_bytes()
    assert fwd_bytes > 0
    J.backward()
    assert eng.sweep_bytes() > fwd_bytes
    g1, gf1 = G.grad.clone(), f0g.grad.clone()
    gd = torch.empty_like(g1)
    g0 = torch.empty_like(f0)
    eng.set_f_device(f0)
    eng.set_cell_force_device(G.detach())
    J2, _ = eng.forward_adjoint(8, gd, grad_f0=g0)
    torch.cuda.synchronize()
    assert abs(J.item() - J2) <= 1e-14 * abs(J2)
    assert torch.equal(g1, gd)
    assert torch.equal(gf1, g0) and float(g0.abs().max()) > 0
    with torch.no_grad():
        J3 = control_loss(G.detach(), eng, f0, 8)
    assert abs(J3.item() - J2) <= 1e-13 * abs(J2)


def test_device_pointers_are_checked():
    eng = _engine(configs.base(visc=0.02, **LAM32_CASES["periodic"]), precision=PREC_FP32)
    nx, ny, nz = eng.shape
    dev = torch.device("cuda", eng.device)
    ok = torch.zeros((19, nz, ny, nx), dtype=torch.float64, device=dev)
    eng.set_f_device(ok)
    bad = [torch.zeros((19, nz, ny, nx - 1), dtype=torch.float64, device=dev),  # size
           torch.zeros((19, nz, ny, nx), dtype=torch.float32, device=dev),      # dtype
           torch.zeros((19, nx, ny, nz), dtype=torch.float64, device=dev).transpose(1, 3),  # layout
           torch.zeros((19, nz, ny, nx), dtype=torch.float64)]                  # host memory
    for t in bad:
        with pytest.raises(ValueError):
            eng.set_f_device(t)
    with pytest.raises(ValueError):
        eng.set_cell_force_device(ok)  # needs 3 values per cell
