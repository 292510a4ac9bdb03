"""torch.autograd.Function over the engine's hand-written adjoint.

J(G) = (1/T) sum_{n=1..T} loss(f^n) where f^{n+1} = step(f^n; F_amb + G) and
G is a per-cell centre body-force field (the FORCEX/Y/Z arrays of the
reference's calc_fb1, src/dugks.cu:1131-1196).  The backward pass is the
discrete adjoint with sqrt(T) checkpointing, run by dugks_forward_adjoint in
the same call as the forward sweep (the checkpoints live in device memory
only for the duration of that call), so forward() returns J and keeps dJ/dG
for backward().
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
        if ctx.needs_input_grad[0]:
            grad = torch.empty_like(G)
            J, _ = engine.forward_adjoint(int(T), grad, loss_kind=loss_kind, interval=int(interval))
            ctx.save_for_backward(grad)
        else:
            total = 0.0
            for n in range(int(T)):
                engine.step(1, istep0=1 + n)
                total += engine.loss(loss_kind)
            J = total / int(T)
            ctx.save_for_backward(torch.zeros_like(G))
        return torch.tensor(J, dtype=torch.float64, device=G.device)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None, None


def control_loss(G, engine, f0, T, loss_kind=_lib.LOSS_WALL_SHEAR, interval=0):
    return ControlLoss.apply(G, engine, f0, T, loss_kind, interval)
