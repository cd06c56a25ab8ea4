"""Writes tests/golden/golden.npz: small input/output vectors of the sweep
hot path, produced by the FP64 oracle (oracle/csmc_oracle.c) on seeded
inputs, and tests/golden/reference_kats.json, the known answers the
reference's own tests hold.

The reference is a Julia package and cannot run here, so the vectors come
from the oracle; the oracle itself is pinned by the reference's known
answers and by closed forms (tests/test_oracle.py).  The colour order of
each case is the one the library's lattice builder produces; a change of
colouring invalidates the sweep vectors and needs a regeneration:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "classicalspinmc.jl_amd"), os.path.dirname(HERE)]

import csmc  # noqa: E402  (host-only table builder, no device needed)
import oracle.oracle as orc  # noqa: E402
from golden_cases import CASES, colour_order  # noqa: E402

SEED = 0x5EED0C0FFEE
TEMP = 0.7
SIGMA = 0.4
CTR0 = 5


def main():
    out = {}
    for name, mk, shape, S, bc in CASES:
        uc = mk()
        lat = csmc.Lattice(shape, uc, S, bc=bc, seed=11)
        T = orc.Tables(uc, shape, bc == "periodic")
        s = lat._spins.copy()
        order = colour_order(lat)
        E, parts = orc.total_energy(T, s, parts=True)
        p = name + "/"
        out[p + "spins"] = s
        out[p + "order"] = order
        out[p + "bsite"] = T.bsite
        out[p + "local_field"] = orc.local_fields(T, s)
        out[p + "site_energy"] = orc.site_energies(T, s)
        out[p + "total_energy"] = np.array([E])
        out[p + "energy_parts"] = parts
        so = s.copy()
        for _ in range(2):
            orc.overrelax(T, so, order)
        out[p + "overrelax2"] = so
        sd = s.copy()
        orc.deterministic(T, sd, order, S)
        out[p + "deterministic1"] = sd
        for kind, tag in ((0, "metropolis"), (1, "cone")):
            sm = s.copy()
            acc, decisions = [], []
            for sw in range(2):
                a, dec, _ = orc.metropolis_coloured(T, sm, order, S, TEMP, kind, SIGMA, SEED, 0, CTR0 + sw)
                acc.append(a)
                decisions.append(dec)
            out[p + tag + "_accepted"] = np.array(acc)
            out[p + tag + "_decisions"] = np.stack(decisions).astype(np.int8)
            out[p + tag + "_spins"] = sm
    out["params"] = np.array([SEED & 0xFFFFFFFF, SEED >> 32, CTR0], dtype=np.int64)
    out["params_f"] = np.array([TEMP, SIGMA])
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)

    kats = {
        "source": "known answers held by the reference's own tests",
        "zeeman_single_site": {"ref": "test/latticetests.jl:9-18", "E": -1.0, "H": [-1.0, 0.0, 0.0]},
        "square_fm_L2": {"ref": "test/latticetests.jl:20-30", "E_per_site": -2.0},
        "spin_norm": {"ref": "test/latticetests.jl:6", "norm": 1.0},
        "kgg_honeycomb_L4_ground_state": {"ref": "test/mctests.jl:43-58", "E_per_site_rounded4": -0.6444},
    }
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
