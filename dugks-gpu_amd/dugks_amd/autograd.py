"""torch.autograd.Function over the engine's hand-written adjoint.

J(G, f0) = (1/T) sum_{n=1..T} loss(f^n) where f^{n+1} = step(f^n; F_amb + G),
f^0 = f0, and G is a per-cell centre body-force field (the FORCEX/Y/Z arrays
of the reference's calc_fb1, src/dugks.cu:1131-1196).  forward() runs the T
steps once, keeping sqrt(T) checkpoints on the engine (dugks_forward_record);
backward() runs the discrete adjoint over them (dugks_adjoint_replay) and
returns dJ/dG and dJ/df0.  Without a backward the checkpoints stay on the
engine until its next sweep or release_sweep().
"""
import torch

from . import _lib


class ControlLoss(torch.autograd.Function):
    """J = ControlLoss.apply(G, engine, f0, T, loss_kind, interval).

    G: float64 CUDA tensor (3, nz, ny, nx) on the engine's device.
    f0: float64 CUDA tensor (19, nz, ny, nx), the initial state (SoA).
    Returns a 0-d float64 tensor.  The engine is left at f^T.
    """

    @staticmethod
    def forward(ctx, G, engine, f0, T, loss_kind=_lib.LOSS_WALL_SHEAR, interval=0):
        G = G.detach().contiguous()
        engine.set_f_device(f0.detach().contiguous())
        engine.set_cell_force_device(G)
        J = engine.forward_record(int(T), loss_kind=loss_kind, interval=int(interval))
        ctx.engine = engine
        ctx.save_for_backward(G)
        return torch.tensor(J, dtype=torch.float64, device=G.device)

    @staticmethod
    def backward(ctx, gJ):
        (G,) = ctx.saved_tensors
        eng = ctx.engine
        # the adjoint linearises about the recorded forward, whose control was G
        eng.set_cell_force_device(G)
        grad_G = torch.empty_like(G)
        want_f0 = ctx.needs_input_grad[2]
        nx, ny, nz = eng.shape
        grad_f0 = torch.empty((19, nz, ny, nx), dtype=torch.float64, device=G.device) if want_f0 else None
        eng.adjoint_replay(grad_G, grad_f0)
        return gJ * grad_G, None, (gJ * grad_f0 if want_f0 else None), None, None, None


def control_loss(G, engine, f0, T, loss_kind=_lib.LOSS_WALL_SHEAR, interval=0):
    return ControlLoss.apply(G, engine, f0, T, loss_kind, interval)
