"""Differentiable BP objectives: torch.autograd.Functions over the device adjoint.

`bp_lognorm(sites, topo)` and `bp_tfim_energy(sites, topo, J, h)` run BP to convergence on
the GPU in forward; backward returns the gradient with respect to every site tensor from
the fixed-point adjoint (implicit differentiation of m* = F(m*, A); for the log-norm the
message term vanishes by Bethe stationarity).  The Julia counterpart is the rrule in
julia/MI355XBP.jl.  Site tensors travel as a flat FP64 tensor in the topology's site
layout.
"""
from __future__ import annotations

import torch

from .engine import Engine


def _solve(topo, sites, dtype, tol, maxiter):
    e = Engine(topo, sites.detach().to("cpu", torch.float64).numpy(), dtype=dtype)
    e.init_messages("identity")
    # max-abs change certifies FP64 messages to `tol`; the fidelity criterion only to sqrt(tol)
    r = e.run(maxiter=maxiter, tol=tol, criterion="maxabs" if dtype == "f64" else "fidelity")
    if not bool(r.converged[0]):
        e.close()
        raise RuntimeError(f"BP did not converge in {maxiter} sweeps (delta {r.delta[0]:.3e})")
    return e


class _LogNorm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sites, topo, dtype, tol, maxiter):
        e = _solve(topo, sites, dtype, tol, maxiter)
        val = e.lognorm()
        ctx.engine = e
        ctx.dev = sites.device
        ctx.dt = sites.dtype
        return torch.tensor(float(val.real if isinstance(val, complex) else val),
                            dtype=sites.dtype, device=sites.device)

    @staticmethod
    def backward(ctx, gout):
        e = ctx.engine
        try:
            g = torch.from_numpy(e.grad_lognorm()).to(ctx.dev, ctx.dt)
        finally:
            e.close()
        return gout * g, None, None, None, None


class _TFIM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sites, topo, J, h, dtype, tol, maxiter, adj_tol, adj_maxiter):
        e = _solve(topo, sites, dtype, tol, maxiter)
        E, g, it, res = e.tfim_energy(J, h, grad=True, adj_tol=adj_tol, adj_maxiter=adj_maxiter)
        e.close()
        ctx.save_for_backward(torch.from_numpy(g).to(sites.device, sites.dtype))
        ctx.adjoint = (it, res)
        return torch.tensor(E, dtype=sites.dtype, device=sites.device)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return gout * g, None, None, None, None, None, None, None, None


def bp_lognorm(sites: torch.Tensor, topo, dtype: str = "f64", tol: float = 1e-12,
               maxiter: int = 2000) -> torch.Tensor:
    return _LogNorm.apply(sites, topo, dtype, tol, maxiter)


def bp_tfim_energy(sites: torch.Tensor, topo, J: float = 1.0, h: float = 3.0,
                   dtype: str = "f64", tol: float = 1e-12, maxiter: int = 2000,
                   adj_tol: float = 1e-10, adj_maxiter: int = 2000) -> torch.Tensor:
    return _TFIM.apply(sites, topo, J, h, dtype, tol, maxiter, adj_tol, adj_maxiter)
