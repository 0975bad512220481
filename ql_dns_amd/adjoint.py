"""Gradients through the GPU adjoint as a torch.autograd.Function.

`final_mf_energy(bx0, by0, q, Rm, ...)` is the final mean-field magnetic energy
0.5 <Bx^2 + By^2>_z after `nsteps` steps of the QL main loop (main.cpp:108-123;
FP64 GPU forward, the same kernels as ``ql_steps``) as a differentiable
function of the initial real profiles Bx(z), By(z) on the NZ grid, the shear q
and the magnetic Reynolds number Rm (eta = 1/Rm).  The fluctuations start from
a fixed state; Ux = Uy = 0 at the start.  The backward pass is the library's
hand-written adjoint (``ql_adjoint_forward`` / ``ql_adjoint_backward``) with
checkpointed recomputation.

The mean fields are stored as NZ x their Fourier coefficients (the dealiased
band of the unnormalised forward transform of the profile), so by Parseval
0.5 <Bx^2 + By^2>_z = 0.5 / NZ^2 * sum over the stored Bx, By coefficients of |M|^2.
"""
from __future__ import annotations

import numpy as np
import torch

from .capi import QLContext
from .params import QLParams


def kept_index(p: QLParams) -> np.ndarray:
    """Positions of the nz stored kz coefficients in the length-NZ spectrum
    (kz order 0..nzl, -nzl..-1; Auxiliary/Kdata.cpp:74-78)."""
    nzl = (p.nz - 1) // 2
    return np.concatenate([np.arange(nzl + 1), np.arange(p.NZ - nzl, p.NZ)])


def mf_energy(p: QLParams, mf: np.ndarray) -> float:
    """0.5 <Bx^2 + By^2>_z from the stored coefficients."""
    return 0.5 * float((np.abs(mf[:2]) ** 2).sum()) / p.NZ ** 2


class _FinalMFEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mf_ri, q, Rm, p, lin0, nsteps, every):
        pp = p.replace(q=float(q), eta=1.0 / float(Rm))
        c = QLContext(pp)
        mf = torch.view_as_complex(mf_ri.detach().contiguous()).numpy()
        c.set_state(lin0, mf)
        c.linear_ops_init(0.0)
        c.adjoint_forward(nsteps, every, 0.0)
        _, mfT, _ = c.get_state()
        ctx.c, ctx.p, ctx.mfT, ctx.Rm = c, pp, mfT, float(Rm)
        return torch.tensor(mf_energy(pp, mfT), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        p = ctx.p
        gmf = np.zeros_like(ctx.mfT)
        gmf[:2] = ctx.mfT[:2] / p.NZ ** 2  # d(0.5 |M|^2 / NZ^2)/dRe + i d/dIm
        _, gm0, gpar = ctx.c.adjoint_backward(None, gmf)
        ctx.c.close()
        g = float(g)
        g_mf = torch.view_as_real(torch.from_numpy(gm0 * g))
        g_q = torch.tensor(gpar[0] * g, dtype=torch.float64)
        g_Rm = torch.tensor(-gpar[3] / ctx.Rm ** 2 * g, dtype=torch.float64)  # eta = 1/Rm
        return g_mf, g_q, g_Rm, None, None, None, None


def final_mf_energy(bx0: torch.Tensor, by0: torch.Tensor, q: torch.Tensor, Rm: torch.Tensor,
                    p: QLParams, lin0: np.ndarray, nsteps: int, every: int) -> torch.Tensor:
    """Differentiable final mean-field magnetic energy.  bx0, by0: float64 [NZ]
    profiles; q, Rm: 0-dim float64 tensors; lin0: initial fluctuations
    [nxy][4][nz] complex (held fixed)."""
    kept = torch.from_numpy(kept_index(p))
    spec = torch.fft.fft(torch.stack([bx0, by0]).to(torch.complex128), dim=-1)[:, kept]
    mf = torch.cat([spec, torch.zeros(p.nMF - 2, p.nz, dtype=torch.complex128)])
    return _FinalMFEnergy.apply(torch.view_as_real(mf), q, Rm, p, lin0, nsteps, every)
