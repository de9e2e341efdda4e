"""torch.autograd bridge for the two-electron energy and its adjoint.

``TwoElectronEnergy.apply(xyz, builder, d_packed, hfscale, coulscale)``
returns E2 = 0.5 Tr D G[D] at nuclear positions ``xyz`` (natm x 3, bohr)
for a fixed AO density; its backward returns dE2/dR from the device adjoint
kernel (the derivative-ERI contraction of OpenQP's grd2 path,
source/integrals/grd2.F90:145-391), scaled by the incoming gradient.  With
hfscale = coulscale = 1 and the converged SCF density this is the
two-electron part of the RHF nuclear gradient that hf_2e_grad adds
(source/modules/hf_gradient.F90:229-337).
"""
from __future__ import annotations

import numpy as np
import torch

from .lib import FockBuilder


def packed_trace(dp: np.ndarray, fp: np.ndarray, nbf: int) -> float:
    """0.5 Tr(D G) on packed symmetric matrices (traceprod_sym_packed, mathlib.F90:69-85)."""
    idx = np.arange(1, nbf + 1)
    diag = idx * (idx + 1) // 2 - 1
    return 0.5 * (2.0 * float(np.dot(dp, fp)) - float(np.dot(dp[diag], fp[diag])))


class TwoElectronEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz: torch.Tensor, builder: FockBuilder, d_packed: np.ndarray, hfscale: float = 1.0,
                coulscale: float = 1.0):
        x = xyz.detach().to("cpu", torch.float64).numpy().reshape(-1, 3)
        builder.set_geometry(x)
        f = builder.fock_jk(d_packed, se=hfscale, sc=coulscale)
        e2 = packed_trace(d_packed, f, builder.bas.nbf)
        ctx.builder, ctx.d, ctx.hs, ctx.cs = builder, d_packed, hfscale, coulscale
        ctx.dev, ctx.shape = xyz.device, xyz.shape
        return torch.tensor(e2, dtype=torch.float64, device=xyz.device)

    @staticmethod
    def backward(ctx, grad_out):
        de = ctx.builder.grad2(ctx.d, hfscale=ctx.hs, coulscale=ctx.cs)
        g = torch.tensor(de, dtype=torch.float64, device=ctx.dev).reshape(ctx.shape)
        return grad_out * g, None, None, None, None
