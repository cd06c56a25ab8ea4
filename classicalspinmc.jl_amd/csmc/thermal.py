"""Thermal derivatives of the energy with respect to couplings.

For a Boltzmann ensemble at temperature T, the derivative of the mean energy
with respect to a coupling J is

    d<E>/dJ = <dE/dJ> - (<E dE/dJ> - <E><dE/dJ>) / T,

so it follows from samples of E and of the adjoint dE/dJ
(csmc_energy_grad_couplings) along Metropolis chains.  The batch's systems
are independent chains on one shared lattice; every sample evaluates the
energy and the coupling gradient of all of them on the GPU in FP64.
"""
import numpy as np

from . import _lib


def thermal_coupling_gradient(batch, T, n_samples, sweeps_between=10, overrelaxation=0,
                              thermalize=200, alg=_lib.METROPOLIS, sigma=60.0, seed=0, directions=None):
    """d<E>/dJ for every coupling entry of the batch's lattice (layout of
    UnitCell.params()) at temperature T.

    Each system runs `thermalize` Metropolis sweeps, then `n_samples` samples
    `sweeps_between` Metropolis sweeps apart (each preceded by
    `overrelaxation` overrelaxation sweeps).  With `directions` (a (P,) or
    (k, P) array), the derivatives along those directions of coupling space
    are returned instead: a coupling scaled into several entries (J2 times
    the identity in six bilinear terms, say) is one direction.  Returns
    (grad, err, means) with err the jackknife standard error over the
    independent chains and means a dict of <E>, <dE/dJ>, the per-system mean
    energies and the sample count."""
    first = batch.lattices[0]
    if any(lat is not first for lat in batch.lattices):
        raise ValueError("the systems must share one lattice")
    if batch.precision != _lib.FP64:
        raise ValueError("the adjoint runs on an FP64 batch")
    n = batch.n
    Tv = np.full(n, float(T))
    sg = np.full(n, float(sigma))
    P = first.unit_cell.n_params()
    W = None
    if directions is not None:
        W = np.atleast_2d(np.asarray(directions, dtype=np.float64))
        if W.shape[1] != P:
            raise ValueError(f"directions need {P} entries each, got {W.shape[1]}")

    def advance(k):
        for _ in range(k):
            if overrelaxation:
                batch.sweep(_lib.OVERRELAX, n_sweeps=overrelaxation)
            batch.sweep(alg, T=Tv, sigma=sg, n_sweeps=1, seed=seed)

    advance(thermalize)
    Q = P if W is None else W.shape[0]
    sE = np.zeros(n)
    sG = np.zeros((n, Q))
    sEG = np.zeros((n, Q))
    for _ in range(n_samples):
        advance(sweeps_between)
        E = batch.total_energy()
        G = batch.grad_couplings()
        if W is not None:
            G = G @ W.T
        sE += E
        sG += G
        sEG += E[:, None] * G

    def estimate(mask):
        m = n_samples * mask.sum()
        mE = sE[mask].sum() / m
        mG = sG[mask].sum(axis=0) / m
        mEG = sEG[mask].sum(axis=0) / m
        return mG - (mEG - mE * mG) / T, mE, mG

    full, mE, mG = estimate(np.ones(n, bool))
    if n > 1:
        jk = np.array([estimate(np.arange(n) != k)[0] for k in range(n)])
        err = np.sqrt((n - 1) / n * ((jk - jk.mean(axis=0)) ** 2).sum(axis=0))
    else:
        err = np.full(Q, np.nan)
    if W is not None and np.ndim(directions) == 1:
        full, err, mG = full[0], err[0], mG[0]
    return full, err, {"energy": mE, "dE_dJ": mG, "energy_per_system": sE / n_samples,
                       "samples": n * n_samples}
