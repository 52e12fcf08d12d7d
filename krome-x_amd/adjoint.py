"""Differentiable batched solves: torch.autograd over the discrete adjoint
of the per-cell BDF integration (the library's krome_adjoint_cells_device,
a hand-written HIP kernel, csrc/kx_adjoint.h).

    import krome, adjoint
    net = krome.load("primordial3")
    opts = net.options(evolve_T=True, rtol=1e-8)
    solve = adjoint.Solve(net, opts)
    n1, T1, flags = solve(lnk, n0, T0, dt, par)
    loss = (torch.log(n1[:, net.idx("H2")])).sum()
    loss.backward()          # lnk.grad, n0.grad, T0.grad

Arguments are float64 CUDA tensors, cell-major as in Network.solve:
lnk [ncell, nrea] (or [nrea], shared by every cell) offsets of the
logarithms of the rate coefficients (each rate k_r exp(lnk_r); zeros: the
network's own rates), n0 [ncell, nspec], T0, dt [ncell], par
[ncell, npar].  The forward integrates with the adjoint kernel's own FP64
integrator; the backward integrates again from the same inputs, keeping a
checkpoint every opts' checkpoint_steps accepted steps (32 by default),
and runs the steps backwards segment by segment from the checkpoints, so
nothing but the inputs is kept between the two.  The step sizes, orders
and accepted attempts of the integration are those of the forward run
(the gradient is that of the discrete scheme they define).  A cell that
fails (flags != 0) keeps its inputs, and its gradient is that of the
identity (no rate gradient).  With Tgas fixed, T0 gets no gradient.
"""
from __future__ import annotations

import torch

import krome

__all__ = ["Solve"]


class _SolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lnk, n0, T0, dt, par, solver):
        net, opts = solver.net, solver.opts
        n = n0.detach().t().contiguous()
        T = T0.detach().clone()
        fl = torch.zeros(T.shape[0], dtype=torch.int32, device=T.device)
        r, _, mem = net.adjoint_device(n, T, dt, par, opts,
                                       lnk=lnk.detach().contiguous(),
                                       flags=fl)
        torch.cuda.synchronize()
        if r < -100 and r not in (-110, -111):
            raise krome.KromeError("adjoint forward: %s (%d)"
                                   % (net.flag_name(r), r))
        solver.last_flag = r
        ctx.solver = solver
        ctx.save_for_backward(lnk, n0, T0, dt, par)
        ctx.mark_non_differentiable(fl)
        return n.t().contiguous(), T, fl

    @staticmethod
    def backward(ctx, gn, gT, _gfl):
        lnk, n0, T0, dt, par = ctx.saved_tensors
        solver = ctx.solver
        net, opts = solver.net, solver.opts
        n = n0.detach().t().contiguous()
        T = T0.detach().clone()
        gns = (torch.zeros_like(n) if gn is None
               else gn.detach().t().contiguous())
        gTs = (torch.zeros_like(T) if gT is None
               else gT.detach().contiguous())
        fl = torch.zeros(T.shape[0], dtype=torch.int32, device=T.device)
        r, grads, mem = net.adjoint_device(n, T, dt, par, opts,
                                           lnk=lnk.detach().contiguous(),
                                           flags=fl, gn=gns, gT=gTs)
        torch.cuda.synchronize()
        if r < -100 and r not in (-110, -111):
            raise krome.KromeError("adjoint backward: %s (%d)"
                                   % (net.flag_name(r), r))
        solver.last_flag = r
        solver.last_memory = mem
        solver.last_flags = fl
        g_lnk, g_n0, g_T0 = grads
        return g_lnk, g_n0.t(), g_T0, None, None, None


class Solve:
    """the differentiable solve of one network's cells with options opts
    (krome.Options); last_memory: the peak device bytes of the last
    backward pass (workspace and checkpoints), last_flags its flags"""

    def __init__(self, net, opts):
        self.net = net
        self.opts = opts
        self.last_flag = 0
        self.last_memory = 0.0
        self.last_flags = None

    def __call__(self, lnk, n0, T0, dt, par):
        ncell = n0.shape[0]
        if lnk.dim() == 1:
            lnk = lnk.expand(ncell, -1)
        if not torch.is_tensor(par):
            par = torch.as_tensor(par, dtype=torch.float64, device=n0.device)
        return _SolveFn.apply(lnk.contiguous(), n0, T0, dt.contiguous(),
                              par.contiguous(), self)
