"""torch.autograd.Function over the hand-written adjoint (include/hc.h,
hc_adj_*), for the check harness: the loss is computed and differentiated by
libhc.so; torch only carries the tensors."""
from __future__ import annotations

import numpy as np
import torch


class HydrographMisfit(torch.autograd.Function):
    """loss = HydrographMisfit.apply(n, mult, adj): n is the per-cell Manning
    n on the reference's padded layout, mult the rain-block multipliers, adj
    a hail_caesar_amd.hc.Adjoint whose context receives n."""

    @staticmethod
    def forward(ctx, n: torch.Tensor, mult: torch.Tensor, adj) -> torch.Tensor:
        adj.ctx.upload("mannings", n.detach().cpu().double().numpy())
        loss, _ = adj.forward(mult.detach().cpu().double().numpy())
        ctx.adj = adj
        ctx.n_meta = (n.dtype, n.device)
        ctx.m_meta = (mult.dtype, mult.device)
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        gn, gm = ctx.adj.backward()
        g = float(grad_out)
        dn = torch.from_numpy(np.ascontiguousarray(gn * g)).to(dtype=ctx.n_meta[0], device=ctx.n_meta[1])
        dm = torch.from_numpy(gm * g).to(dtype=ctx.m_meta[0], device=ctx.m_meta[1])
        return dn, dm, None
