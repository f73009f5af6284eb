"""torch.autograd.Function shim over the adjoint kernel (constant fitting).

    fit = ConstantFit(ctx, pop, const_slot)        # compiles parametric programs
    consts = fit.initial().requires_grad_()        # [E, 8] on the GPU
    opt = torch.optim.Adam([consts], lr=1e-2)
    for _ in range(20):
        opt.zero_grad()
        mse = fit.mse(consts)                      # [E]; one forward+adjoint launch
        mse.nansum().backward()
        opt.step()

forward() runs sr_param_grad (MSE and dMSE/dconstants in one pass over the
data) on torch's current stream and keeps the gradient for backward(); when the
constants do not require grad it runs the forward-only kernel.  Expressions
with no finite prediction get NaN MSE and zero gradient.
"""
from __future__ import annotations

import torch

from .api import Context, Population


class _ExprMse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, consts, fit):
        consts_c = consts.detach().contiguous()
        E = consts_c.shape[0]
        mse = torch.empty(E, device=consts.device, dtype=torch.float32)
        fit.ctx.set_stream(torch.cuda.current_stream(consts.device).cuda_stream)
        if consts.requires_grad:
            grad = torch.empty(E, 8, device=consts.device, dtype=torch.float32)
            fit.ctx.param_grad(consts_c.data_ptr(), mse.data_ptr(), grad.data_ptr())
            ctx.save_for_backward(grad)
        else:
            fit.ctx.param_eval(consts_c.data_ptr(), mse.data_ptr())
        return mse

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        g = torch.nan_to_num(grad, nan=0.0, posinf=0.0, neginf=0.0)
        return gout.unsqueeze(1) * g, None


class ConstantFit:
    """Parametric population bound to an evaluator context."""

    def __init__(self, ctx: Context, pop: Population, const_slot, device="cuda"):
        self.ctx = ctx
        self.pop = pop
        self._init = ctx.param_prepare(pop, const_slot)
        self.nslots, self.tokens = ctx.param_nslots()
        self.device = torch.device(device)

    def initial(self) -> torch.Tensor:
        return torch.tensor(self._init, device=self.device, dtype=torch.float32)

    def mse(self, consts: torch.Tensor) -> torch.Tensor:
        if consts.dtype != torch.float32 or consts.shape != (self.pop.num_exprs, 8):
            raise ValueError("consts must be float32 [num_exprs, 8]")
        return _ExprMse.apply(consts, self)

    def fd_grad(self, consts: torch.Tensor, h: float = 1e-3) -> torch.Tensor:
        """Central differences on the GPU (2 forward passes per free constant)."""
        c = consts.detach().contiguous()
        g = torch.empty_like(c)
        self.ctx.set_stream(torch.cuda.current_stream(c.device).cuda_stream)
        self.ctx.param_fd_grad(c.data_ptr(), h, g.data_ptr())
        return g

    def adjoint_grad(self, consts: torch.Tensor):
        c = consts.detach().contiguous()
        mse = torch.empty(c.shape[0], device=c.device, dtype=torch.float32)
        g = torch.empty_like(c)
        self.ctx.set_stream(torch.cuda.current_stream(c.device).cuda_stream)
        self.ctx.param_grad(c.data_ptr(), mse.data_ptr(), g.data_ptr())
        return mse, g
