"""total_energy (hamiltonian.jl:70-132) as a differentiable torch function
backed by the library's adjoint kernels (csmc_energy_grad_spins,
csmc_energy_grad_couplings): gradients with respect to the spins of every
system of a batch and to the couplings of the lattice the batch shares.

The energy and both gradients are evaluated on the GPU in FP64 from the
batch's device state; torch only carries the values across the autograd
graph (host tensors in, host tensors out).
"""
import numpy as np
import torch

from ._lib import FP64


class _TotalEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spins, couplings, batch):
        if couplings is not None:
            lat = batch.lattices[0]
            lat.update_couplings(lat.unit_cell.with_params(couplings.detach().cpu().numpy()))
        s = spins.detach().cpu().numpy().astype(np.float64).reshape(-1, 3)
        for i in range(batch.n):
            batch.set_spins(i, s[batch.offsets[i]:batch.offsets[i + 1]])
        E = batch.total_energy()
        gS = np.concatenate(batch.grad_spins()) if ctx.needs_input_grad[0] else None
        gJ = batch.grad_couplings() if couplings is not None and ctx.needs_input_grad[1] else None
        ctx.shape = spins.shape
        ctx.offsets = batch.offsets
        ctx.save_for_backward(torch.from_numpy(gS) if gS is not None else torch.zeros(0, dtype=torch.float64),
                              torch.from_numpy(gJ) if gJ is not None else torch.zeros(0, dtype=torch.float64))
        return torch.from_numpy(E).to(spins.dtype)

    @staticmethod
    def backward(ctx, gE):
        gS, gJ = ctx.saved_tensors
        gE = gE.detach().to(torch.float64).cpu()
        dS = dJ = None
        if ctx.needs_input_grad[0]:
            per_site = torch.repeat_interleave(gE, torch.from_numpy(np.diff(ctx.offsets)))
            dS = (gS * per_site[:, None]).reshape(ctx.shape)
        if ctx.needs_input_grad[1]:
            dJ = (gE[:, None] * gJ).sum(dim=0)
        return dS, dJ, None


def total_energy(batch, spins, couplings=None):
    """Energies E[n_sys] of `batch` at `spins` (all systems' sites
    concatenated, shape (total_sites, 3), or (n_sys, N, 3) for equal sizes).

    `couplings` (optional): a vector in the layout of UnitCell.params() for
    the lattice every system of the batch shares; the lattice is updated to
    these values before the evaluation, and the gradient flows to them."""
    if batch.precision != FP64:
        raise ValueError("the adjoint runs on an FP64 batch")
    if couplings is not None:
        first = batch.lattices[0]
        if any(lat is not first for lat in batch.lattices):
            raise ValueError("coupling gradients need every system to share one lattice")
    return _TotalEnergy.apply(spins, couplings, batch)
