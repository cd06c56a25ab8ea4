"""The tiled batch-wide sweeps (tile.hip) against the colour-pass form
(CSMC_TILE=0) on the device: every spin, every accepted count and the
counter words must be equal bit for bit, for overrelaxation, deterministic,
Metropolis and cone sweeps in FP32 and FP64, on lattices coloured by basis
(honeycomb, pyrochlore) and by a cell pattern (square, triangular), and on a
batch of two lattices whose tile counts differ.  The colour-pass form is
itself pinned to the oracle's colour-class loops (test_gpu_parity.py); the
CPU replay of the tile index arithmetic is in test_tiles.py."""
import numpy as np
import pytest

import csmc
from csmc import workloads as W
from csmc.unitcell import Pyrochlore, UnitCell, addBasisSite, addBilinear, addOnSite, addZeemanCoupling

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx_pair():
    import os
    tiled = csmc.Context(0)
    os.environ["CSMC_TILE"] = "0"
    try:
        plain = csmc.Context(0)
    finally:
        del os.environ["CSMC_TILE"]
    yield tiled, plain
    tiled.close()
    plain.close()


def pyrochlore_heisenberg():
    uc = Pyrochlore()
    for b1, b2, o in W.PYRO_BONDS:
        addBilinear(uc, b1, b2, np.eye(3), o)
    return uc


def square_xxz():
    """diagonal couplings whose products round (the tiles' diagonal form
    must round them as the colour passes do)"""
    uc = W.square_heisenberg(J=1.0)
    for d in ((1, 0), (0, 1)):
        addBilinear(uc, 1, 1, np.diag([0.7, 0.7, -1.3]), d)
    return uc


def xxz_chain():
    """a one-dimensional chain: two raw colours (x mod 2), diagonal couplings
    whose products round, a field"""
    uc = UnitCell([1.0])
    addBasisSite(uc, np.zeros(1))
    for d in ((1,), (-1,)):
        addBilinear(uc, 1, 1, np.diag([0.7, 0.7, -1.3]), d)
    addZeemanCoupling(uc, 1, np.array([0.0, 0.1, 0.2]))
    return uc


def square_onsite():
    uc = W.square_heisenberg()
    addOnSite(uc, 1, np.diag([0.1, -0.05, 0.2]))
    return uc


CASES = [
    ("honeycomb_64", lambda: [(W.honeycomb_kitaev(1.0, -2.0), (64, 64))] * 4),
    ("square_onsite_64", lambda: [(square_onsite(), (64, 64))] * 3),
    ("square_xxz_64", lambda: [(square_xxz(), (64, 64))] * 2),
    ("xxz_chain_4096", lambda: [(xxz_chain(), (4096,))] * 3),
    ("triangular_dm_48", lambda: [(W.triangular_dm(-1.0, 0.1, (0.0, 0.0, 0.2)), (48, 48))] * 3),
    ("pyrochlore_8", lambda: [(pyrochlore_heisenberg(), (8, 8, 8))] * 2),
    ("pyrochlore_12", lambda: [(pyrochlore_heisenberg(), (12, 12, 12))] * 2),
    ("mixed_64_32", lambda: [(W.honeycomb_kitaev(1.0, -2.0), (64, 64)), (W.honeycomb_kitaev(1.0, -2.0), (32, 32))]),
]


def _batches(ctx_pair, spec, prec):
    tiled, plain = ctx_pair
    out = []
    for ctx in (tiled, plain):
        lats, cache = [], {}
        for uc, shape in spec:
            key = (id(uc), shape)
            if key not in cache:
                cache[key] = csmc.Lattice(shape, uc, 0.5, seed=1)
            lats.append(cache[key])
        b = csmc.SpinBatch(lats, precision=prec, ctx=ctx)
        b.randomize(11)
        b.set_counters(np.arange(b.n) * 7 + 3)
        out.append(b)
    return out


def _equal(bt, bp):
    for i in range(bt.n):
        assert np.array_equal(bt.get_spins(i), bp.get_spins(i)), f"system {i}"
    assert np.array_equal(bt.counters(), bp.counters())


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_tiled_sweeps_equal_colour_passes(ctx_pair, name, spec, prec):
    print(f"[tiles] {name} FP{prec}", flush=True)
    bt, bp = _batches(ctx_pair, spec(), prec)
    nt, why = bt.tiles()
    assert nt > 0, why
    assert bp.tiles()[0] == 0
    # two sweeps a launch where the box of two sweeps fits the lattice and
    # the 48 KiB LDS budget (pyrochlore 12^3 in FP64: no tile shape does)
    one = name == "pyrochlore_8" or (name == "pyrochlore_12" and prec == 64)
    assert bt.tile_sweeps() == (1 if one else 2)
    n = bt.n
    T = np.linspace(0.2, 1.5, n)
    sg = np.full(n, 0.4)
    for kind, kw in [(csmc.OVERRELAX, {}), (csmc.METROPOLIS, dict(T=T)), (csmc.CONE, dict(T=T, sigma=sg)),
                     (csmc.OVERRELAX, {}), (csmc.DETERMINISTIC, {})]:
        for sweeps in (1, 2, 3):
            at = bt.sweep_passes(kind, n_sweeps=sweeps, seed=5, **kw)
            ap = bp.sweep_passes(kind, n_sweeps=sweeps, seed=5, **kw)
            assert np.array_equal(at, ap), (kind, sweeps)
            _equal(bt, bp)
    # energies evaluate the state the tiled sweeps left (in either buffer)
    assert np.array_equal(bt.total_energy(), bp.total_energy())
    bt.close()
    bp.close()


def test_graph_after_tiled_sweeps(ctx_pair):
    """a captured plain-form graph still sweeps the batch's state after
    tiled sweeps moved it to the other buffer"""
    spec = [(W.honeycomb_kitaev(1.0, -2.0), (64, 64))] * 2
    bt, bp = _batches(ctx_pair, spec, 32)
    gt = bt.graph_sweeps(csmc.OVERRELAX, n_sweeps=1)
    gp = bp.graph_sweeps(csmc.OVERRELAX, n_sweeps=1)
    for _ in range(2):
        bt.sweep_passes(csmc.OVERRELAX, n_sweeps=1)  # odd: the state moves to the other buffer
        bp.sweep_passes(csmc.OVERRELAX, n_sweeps=1)
        gt.launch()
        gp.launch()
        _equal(bt, bp)
    gt.close()
    gp.close()
    bt.close()
    bp.close()


@pytest.mark.parametrize("sweeps_per_launch", [1, 3])
def test_sweeps_per_launch(sweeps_per_launch):
    """CSMC_TILE_SWEEPS: one and three sweeps a tiled launch give the
    colour-pass results for sweep counts around the launch size"""
    import os
    os.environ["CSMC_TILE_SWEEPS"] = str(sweeps_per_launch)
    try:
        tiled = csmc.Context(0)
    finally:
        del os.environ["CSMC_TILE_SWEEPS"]
    os.environ["CSMC_TILE"] = "0"
    try:
        plain = csmc.Context(0)
    finally:
        del os.environ["CSMC_TILE"]
    for name, spec in [("honeycomb_64", [(W.honeycomb_kitaev(1.0, -2.0), (64, 64))] * 3),
                       ("triangular_dm_48", [(W.triangular_dm(-1.0, 0.1, (0.0, 0.0, 0.2)), (48, 48))] * 2)]:
        print(f"[tiles] {name} {sweeps_per_launch} sweeps a launch", flush=True)
        bt, bp = _batches((tiled, plain), spec, 32)
        assert bt.tile_sweeps() == sweeps_per_launch
        T = np.linspace(0.3, 1.2, bt.n)
        for kind, kw in [(csmc.OVERRELAX, {}), (csmc.METROPOLIS, dict(T=T))]:
            for sweeps in (2, 3, 4, 7):
                at = bt.sweep_passes(kind, n_sweeps=sweeps, seed=9, **kw)
                ap = bp.sweep_passes(kind, n_sweeps=sweeps, seed=9, **kw)
                assert np.array_equal(at, ap), (kind, sweeps)
                _equal(bt, bp)
        bt.close()
        bp.close()
    tiled.close()
    plain.close()
