"""torch.autograd binding of the engine's adjoint gradient.

A thin shim over the C-ABI session (include/gsvit_session.h): the forward
pass writes the design permittivities into the engine and runs the whole
simulation (gsv_objective), the backward pass runs the engine's adjoint
(gsv_objective_gradient).  Tensors cross the boundary through host memory;
the simulation itself stays on the GPU inside the engine.

    from gsvit_copy_amd import Session
    from gsvit_copy_amd.autograd import DesignObjective

    s = Session("inverse_design.par")        # DESIGN_BOX + OBJECTIVE_DFT
    s.adjoint_setup()
    eps = torch.tensor(s.design_eps(), requires_grad=True)
    J = DesignObjective.apply(eps, s)
    J.backward()                              # eps.grad = dJ / d eps_r
"""
import numpy as np
import torch


class DesignObjective(torch.autograd.Function):
    """J(eps_r) for the session's DESIGN_BOX / OBJECTIVE_DFT problem.

    eps_r: tensor of the design box's shape ([i][j][k]); returns a 0-d float64
    tensor.  The gradient is computed in FP32 by the engine's adjoint and
    returned in eps_r's dtype and device.
    """

    @staticmethod
    def forward(ctx, eps_r, session):
        session.set_design_eps(eps_r.detach().to("cpu", torch.float32).numpy())
        J = session.objective()
        ctx.session = session
        ctx.like = (eps_r.dtype, eps_r.device)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.session.objective_gradient()
        dtype, device = ctx.like
        grad = torch.from_numpy(np.ascontiguousarray(g)).to(device=device, dtype=dtype)
        return grad * grad_out.to(dtype), None


def objective(session, eps_r):
    """J(eps_r) as a differentiable torch scalar."""
    return DesignObjective.apply(eps_r, session)
