"""torch.autograd.Function over the device forward/adjoint pair, for the
verification harness only: the objective's gradient with respect to the
initial state comes from sd2d_vjp, so autograd-based checks can drive it."""
import numpy as np
import torch


class DeviceDragObjective(torch.autograd.Function):
    """J(q0) = time-averaged wall drag of `nsteps` RK_S5P3 steps from q0
    (FP64, CPU tensors in the Fortran-order shape (9, N, N, NCELL))."""

    @staticmethod
    def forward(ctx, q0, solver, nsteps, ckpt):
        solver.set_state(np.asfortranarray(q0.detach().numpy()), [0.0, 0.0, 0.0, 0.0])
        J = solver.fwd(nsteps, ckpt)
        ctx.solver = solver
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, jbar):
        q0b, _, _ = ctx.solver.vjp(None, float(jbar))
        return torch.from_numpy(np.ascontiguousarray(q0b)), None, None, None
