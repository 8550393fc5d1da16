"""torch.autograd harness over the hand-written adjoint.

EwaldEnergy.apply(x, y, z, mc) returns the total energy U_total (kT) of the
configuration (x, y, z) (box units, float64 tensors of N atoms) as computed by
the library's full recompute (mc_calc_total_sums + mc_total_energy,
MC.f90:477-698: real space, k space Coulomb + LJ-Ewald, tails, external
term).  Its backward calls the adjoint kernel (yamc_forces), which recomputes
every atom's sin/cos(k.r) tables on the device instead of keeping them from
the forward pass, and returns dU_total/dr = -g for all atoms.  The graph
holds no device tensors of the library: the positions are staged through the
C ABI, and the gradient comes back as host arrays converted to the input
tensors' device and dtype.
"""
from __future__ import annotations

import numpy as np
import torch

from .ewald import EwaldMC


class EwaldEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, y: torch.Tensor, z: torch.Tensor, mc: EwaldMC):
        host = [t.detach().to("cpu", torch.float64).numpy() for t in (x, y, z)]
        mc.set_positions(*host)
        e, overlap = mc.mc_init()
        ctx.mc = mc
        ctx.like = x
        return torch.tensor(e["U_total"], dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g, _ = ctx.mc.forces()  # -dU/dr, positions as staged by forward
        scale = float(grad_out.detach().cpu())
        dudr = torch.from_numpy(np.ascontiguousarray(-g * scale)).to(ctx.like.device, ctx.like.dtype)
        return dudr[:, 0], dudr[:, 1], dudr[:, 2], None


def ewald_energy(mc: EwaldMC, x: torch.Tensor, y: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
    """U_total(x, y, z) with the adjoint as its gradient."""
    return EwaldEnergy.apply(x, y, z, mc)
