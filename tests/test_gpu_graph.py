"""Plain-form sweeps captured as one hipGraph (csmc_graph_sweeps) against the
same sweeps made eagerly (csmc_sweep_plain): spins, Philox counters, accepted
counts and adaptive cone widths bit for bit, for every replay.  The graph is
replayed several times, with eager work on the batch (energies, a spin
upload, counter reads) between the replays, and each replay is checked
against its eager twin."""
import numpy as np
import pytest

import csmc
from csmc import workloads as W

pytestmark = pytest.mark.gpu

KINDS = [("overrelax", csmc.OVERRELAX), ("metropolis", csmc.METROPOLIS), ("cone", csmc.CONE),
         ("cone_adaptive", csmc.CONE_ADAPTIVE), ("deterministic", csmc.DETERMINISTIC)]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("kind", KINDS, ids=[k[0] for k in KINDS])
def test_graph_replays_equal_eager_sweeps(kind, prec):
    name, k = kind
    print(f"[graph] {name} fp{prec}", flush=True)
    lats = [csmc.Lattice((8, 8), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=3),
            csmc.Lattice((6, 6), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=4)]
    T = np.array([0.4, 0.9])
    sigma = np.array([0.7, 1.3])
    eager = csmc.SpinBatch(lats, precision=prec)
    graph = csmc.SpinBatch(lats, precision=prec)
    for b in (eager, graph):
        b.randomize(11)
    g = graph.graph_sweeps(k, T=T, sigma=sigma, n_sweeps=3, seed=29)
    sg_eager = sigma.copy()
    for rep in range(4):
        sg = sigma.copy() if k != csmc.CONE_ADAPTIVE else sg_eager.copy()
        # (adaptive cone: the graph continues from the widths the device
        # updated, the eager calls from the widths they returned)
        acc = eager.sweep(k, T=T, sigma=sg, n_sweeps=3, seed=29, plain=True)
        g.launch()
        gacc, gsg = g.accepted()
        for r in range(2):
            assert np.array_equal(graph.get_spins(r), eager.get_spins(r)), (name, rep, r)
        assert np.array_equal(graph.counters(), eager.counters()), (name, rep)
        if k in (csmc.METROPOLIS, csmc.CONE, csmc.CONE_ADAPTIVE):
            assert np.array_equal(gacc, acc), (name, rep)
            assert acc.sum() > 0
        if k == csmc.CONE_ADAPTIVE:
            assert np.array_equal(gsg, sg), (name, rep)
            sg_eager = sg
        # eager work on the graph's batch between replays
        graph.total_energy()
        eager.total_energy()
        if rep == 1:
            s0 = eager.get_spins(0)
            graph.set_spins(0, s0)
            eager.set_spins(0, s0)
    g.close()
    eager.close()
    graph.close()


def test_graph_errors():
    lat = csmc.Lattice((4, 4), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=1)
    b = csmc.SpinBatch([lat], precision=64)
    with pytest.raises(csmc.CsmcError, match="temperatures"):
        b.graph_sweeps(csmc.METROPOLIS, n_sweeps=1)
    with pytest.raises(csmc.CsmcError, match="n_sweeps"):
        b.graph_sweeps(csmc.OVERRELAX, n_sweeps=0)
    b.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_batch_wide_passes_equal_plain_sweeps(prec):
    """csmc_sweep_passes (one launch per colour pass for every system) and
    csmc_sweep_plain (one system at a time) make the same updates with the
    same draws: equal bit for bit on an uneven batch whose lattices have
    different colour counts"""
    cells = W.uneven_batch(n=5, sizes=(5, 6, 8, 9, 12), seed=7)
    lats = [csmc.Lattice(shape, uc, 1.0, seed=20 + k) for k, (uc, shape) in enumerate(cells)]
    assert len({lat.info.n_colours for lat in lats}) > 1
    T = np.linspace(0.3, 1.2, len(lats))
    a = csmc.SpinBatch(lats, precision=prec)
    b = csmc.SpinBatch(lats, precision=prec)
    for x in (a, b):
        x.randomize(5)
    for kind in (csmc.OVERRELAX, csmc.METROPOLIS, csmc.CONE, csmc.DETERMINISTIC):
        acc_a = a.sweep(kind, T=T, sigma=0.8, n_sweeps=2, seed=13, plain=True)
        acc_b = b.sweep_passes(kind, T=T, sigma=0.8, n_sweeps=2, seed=13)
        for r in range(len(lats)):
            assert np.array_equal(a.get_spins(r), b.get_spins(r)), (kind, r)
        assert np.array_equal(a.counters(), b.counters())
        if kind in (csmc.METROPOLIS, csmc.CONE):
            assert np.array_equal(acc_a, acc_b) and acc_a.sum() > 0
    with pytest.raises(csmc.CsmcError, match="adaptive"):
        b.sweep_passes(csmc.CONE_ADAPTIVE, T=T)
    a.close()
    b.close()
