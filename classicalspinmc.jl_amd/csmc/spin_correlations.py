"""The structure-factor runners of spin_correlations.jl over saved
configurations (host side; the correlations themselves are computed on the
GPU, every configuration of a directory in one batched call).

compute_equal_time_correlations   spin_correlations.jl:6-43 (csmc.lattice)
runEqualTimeStructureFactor       spin_correlations.jl:48-110
compute_equal_time_structure_factor  spin_correlations.jl:116-145

The reference splits the configurations over MPI ranks; here one process
takes them all and computes their correlations in batches of `batch`
systems on one device.
"""
import datetime
import os

import numpy as np

from . import hdf5io
from .batch import SpinBatch


def runEqualTimeStructureFactor(path, lat, ks, override=False, batch=256, ctx=None):
    """runEqualTimeStructureFactor!(path, lat, ks, override): for every
    configuration path * "IC_<i>.h5", i = 0 .. (number of files in path) - 1,
    compute the equal-time correlations of its spins on `lat` at the momenta
    ks (D x n_k) and write them to the file's spin_correlations group (SSF,
    SSF_momentum); files that already have the group are skipped unless
    `override`."""
    ks = np.asarray(ks, dtype=np.float64)
    n_ic = len(os.listdir(path))
    todo = []
    for ic in range(n_ic):
        fname = f"{path}IC_{ic}.h5"
        with hdf5io.h5open(fname, "r") as f:
            exists = f.has("spin_correlations")
        if exists and not override:
            print(f"Skipping IC_{ic}")
            continue
        todo.append((ic, fname))
    for b0 in range(0, len(todo), batch):
        part = todo[b0:b0 + batch]
        b = SpinBatch([lat] * len(part), precision=64, ctx=ctx)
        for r, (ic, fname) in enumerate(part):
            hdf5io.read_spin_configuration(lat, fname)
            b.set_spins(r, lat._spins)
        print(f"Computing SSF {b0 + 1}-{b0 + len(part)}/{len(todo)} on rank 0")
        S = b.spin_correlations(ks)
        b.close()
        for r, (ic, fname) in enumerate(part):
            print(f"Writing IC {ic} to file on rank 0")
            with hdf5io.h5open(fname, "r+") as f:
                with f.group("spin_correlations", create=True) as g:
                    hdf5io.overwrite_keys(g, {"SSF": np.ascontiguousarray(S[r].T),
                                              "SSF_momentum": np.ascontiguousarray(ks.T)})
    print(f"Calculation completed on rank 0 on {datetime.datetime.now():%d %b %Y %H:%M:%S}")


def compute_equal_time_structure_factor(path, dest):
    """compute_equal_time_structure_factor(path, dest): the mean of the SSF
    datasets of every file in `path`, written with the momenta of the first
    file to dest's spin_correlations group.  Returns (SSF 9 x n_k,
    momenta D x n_k)."""
    print(f"Initializing LogBinner in {path}")
    files = sorted(os.listdir(path))
    with hdf5io.h5open(os.path.join(path, files[0]), "r") as f0:
        ks = f0.read("spin_correlations/SSF_momentum")
    print(f"Collecting correlations from {len(files)} files")
    total = None
    for name in files:
        with hdf5io.h5open(os.path.join(path, name), "r") as f:
            s = f.read("spin_correlations/SSF")
        total = s.copy() if total is None else total + s
    mean = total / len(files)
    print(f"Writing to {dest}")
    with hdf5io.h5open(dest, "r+") as d:
        with d.group("spin_correlations", create=True) as g:
            hdf5io.overwrite_keys(g, {"SSF": mean, "SSF_momentum": ks})
    print("Done")
    return mean.T.copy(), np.asarray(ks).T.copy()
