"""Differentiable Mdot: a torch.autograd.Function over the solver's C ABI.

    from doublyperiodicstokes_amd.autograd import mdot
    U, W = mdot(solver, positions, forces, torques)   # device tensors
    loss(U, W).backward()   # gradients for positions (forces-only), forces, torques

The backward pass is dps_mdot_adjoint: force and torque gradients are one
Mdot of the output gradients (the solve taken as self-adjoint), and position
gradients come from derivative-kernel interpolation of the velocity grids of
the inputs and of the output gradients (second derivatives of the torque
kernel for the torque terms).  Position gradients with torques need the
derivative torque kernels (not useRegKernel).
"""
import torch


class _Mdot(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, positions, forces, torques):
        solver.setPositions(positions)
        has_t = torques is not None and torques.numel() > 0
        U, W = solver.mdot_device(forces, torques if has_t else None)
        ctx.solver = solver
        ctx.has_t = has_t
        ctx.save_for_backward(positions, forces, torques if has_t else forces.new_empty(0))
        return U, (W if W is not None else U.new_empty(0))

    @staticmethod
    def backward(ctx, gU, gW):
        solver = ctx.solver
        positions, forces, torques = ctx.saved_tensors
        if gU is None:
            gU = torch.zeros(forces.numel(), dtype=forces.dtype, device=forces.device)
        gU = gU.contiguous()
        # the solver may have been moved since the forward call
        solver.setPositions(positions)
        need_x, need_f, need_t = ctx.needs_input_grad[1], ctx.needs_input_grad[2], ctx.needs_input_grad[3]
        gF = gT = gX = None
        if ctx.has_t:
            if gW is None or gW.numel() == 0:
                gW = torch.zeros_like(gU)
            gF, gT, gX = solver.mdot_adjoint_device(forces, gU, torques=torques, grad_angular=gW.contiguous(),
                                                    positions=need_x)
        else:
            gF, _, gX = solver.mdot_adjoint_device(forces, gU, positions=need_x)
        return (None,
                gX.to(positions.dtype) if need_x else None,
                gF.to(forces.dtype) if need_f else None,
                gT.to(torques.dtype) if (ctx.has_t and need_t) else None)


def mdot(solver, positions, forces, torques=None):
    """Differentiable (U, W) = Mdot(forces, torques) at `positions` (W is None
    without torques).  `solver` is an initialised DPStokes; all tensors are
    device tensors of 3N values."""
    U, W = _Mdot.apply(solver, positions, forces, torques)
    return U, (W if W.numel() else None)
