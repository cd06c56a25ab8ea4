"""Sparse contraction of cubic and quartic couplings.  The host labels each
(basis, slot) coupling tensor with the pattern of its non-zero entries
(p = q and r = t for (Si.Sj)(Sk.Sl), p = r and q = t, p = t and q = r,
pairwise distinct indices for S.(S x S), all zero for a padding slot), and
the field kernels add only the products with those entries, in the order of
the full contraction.

Checked here: the labels the library reports for two cells, and that the
same sweeps, exchanges and energies in a context with CSMC_DENSE_TENSORS=1
(every tensor contracted in full) give the same bits.  Parity with the
oracle for the same cell is in test_gpu_parity (case tensor_patterns)."""
import os
import re

import numpy as np
import pytest

import csmc
from csmc import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(dense, verbose=False):
    env = {"CSMC_DENSE_TENSORS": "1" if dense else "0", "CSMC_VERBOSE": "1" if verbose else "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return csmc.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _labels(text):
    m = re.findall(r"cubic slots dense (\d+) alt (\d+) zero (\d+), quartic slots dense (\d+) pq (\d+) "
                   r"pr (\d+) pt (\d+) zero (\d+)", text)
    assert len(m) == 1, text
    return tuple(int(v) for v in m[0])


# (dense, alt, zero) cubic, (dense, pq, pr, pt, zero) quartic
LABELS = [
    # basis 1: alt + padding, basis 2: alt + dense; quartic basis 1: pr, pq
    # (role 1 of the second term), dense; basis 2: pt (role 1), pq, padding
    ("tensor_pattern_cell", W.tensor_pattern_cell, (4, 4), (1, 2, 1, 1, 2, 1, 1, 1)),
    # 4 bases x 6 chirality faces, 4 bases x 6 biquadratic bonds
    ("pyrochlore_mixed", W.pyrochlore_mixed, (2, 2, 2), (0, 24, 0, 0, 24, 0, 0, 0)),
]


@pytest.mark.parametrize("name,mk,shape,want", LABELS, ids=[c[0] for c in LABELS])
def test_pattern_labels(name, mk, shape, want, capfd):
    print(f"[tensor_patterns] labels {name}", flush=True)
    lat = csmc.Lattice(shape, mk(), 1.0, seed=1)
    for dense in (False, True):
        capfd.readouterr()
        ctx = _ctx(dense, verbose=True)
        b = csmc.SpinBatch([lat], precision=64, ctx=ctx)
        got = _labels(capfd.readouterr().err)
        if dense:
            n3 = sum(want[:3])
            n4 = sum(want[3:])
            assert got == (n3 - want[2], 0, want[2], n4 - want[7], 0, 0, 0, want[7])
        else:
            assert got == want
        b.close()
        ctx.close()


CASES = [("tensor_pattern_cell", W.tensor_pattern_cell, (6, 6)),
         ("pyrochlore_mixed", W.pyrochlore_mixed, (2, 2, 2))]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("name,mk,shape", CASES, ids=[c[0] for c in CASES])
def test_sparse_equals_dense_bit_for_bit(name, mk, shape, prec):
    print(f"[tensor_patterns] sparse vs dense {name} fp{prec}", flush=True)
    lat = csmc.Lattice(shape, mk(), 1.0, seed=2)
    R = 3
    T = np.array([0.3, 0.7, 1.5])
    ctxs = [_ctx(False), _ctx(True)]
    bs = [csmc.SpinBatch([lat] * R, precision=prec, ctx=c) for c in ctxs]
    for b in bs:
        b.randomize(9)
    H = [b.local_field(1) for b in bs]
    assert np.array_equal(H[0], H[1])
    for plain in (False, True):
        for kind in (csmc.OVERRELAX, csmc.METROPOLIS, csmc.CONE, csmc.DETERMINISTIC):
            acc = [b.sweep(kind, T=T, sigma=0.6, n_sweeps=2, seed=31, plain=plain) for b in bs]
            for r in range(R):
                assert np.array_equal(bs[0].get_spins(r), bs[1].get_spins(r)), (kind, plain, r)
            if kind in (csmc.METROPOLIS, csmc.CONE):
                assert np.array_equal(acc[0], acc[1]) and acc[0].sum() > 0
    E = [b.total_energy() for b in bs]
    assert np.array_equal(E[0], E[1])
    # parallel tempering: exchanges decided from the energies the sweeps keep
    params = dict(t_thermalization=10, t_measurement=30, probe_rate=3, swap_rate=5, overrelaxation_rate=2,
                  report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    for b in bs:
        b.pt_init(T, params, csmc.METROPOLIS, seed=4)
        b.pt_advance(-1)
    st = [b.pt_state() for b in bs]
    assert np.array_equal(st[0]["system_of_slot"], st[1]["system_of_slot"])
    assert np.array_equal(st[0]["accepted_exchange"], st[1]["accepted_exchange"])
    for r in range(R):
        assert np.array_equal(bs[0].get_spins(r), bs[1].get_spins(r)), ("pt", r)
    for b in bs:
        b.close()
    for c in ctxs:
        c.close()
