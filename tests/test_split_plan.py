"""Host planning of the split parallel-tempering kernel (csmc_split_plan,
no device needed): strips of cells along the first lattice dimension, each
with the halo one window of sweeps needs.  The library replays every window
phase by phase with a validity bit per local site (a site updated from a
stale or missing neighbour, or one whose colour moves on without it, turns
invalid) from every sweep boundary of the window, and counts owned sites
that end a phase invalid and window-end energy neighbours that are invalid.
These tests hold it to zero for the workload lattices and schedules, and
check the sizes against what the geometry fixes."""
import ctypes as C

import numpy as np
import pytest

from csmc import _lib, workloads as W
from csmc.unitcell import Pyrochlore, addBasisSite, addBilinear


def plan(uc, shape, G, nor, periodic=1):
    if len(uc.basis) == 0:  # the Lattice constructor's default site (lattice.jl:68-70)
        addBasisSite(uc, np.zeros(uc.D))
    s, keep = uc._pack()
    sh = np.ascontiguousarray(np.array(shape, np.int32))
    st = np.zeros(7 + 4 * G, np.int64)
    _lib.check(_lib.lib().csmc_split_plan(C.byref(s), _lib.ptr(sh), periodic, G, nor, _lib.ptr(st)))
    parts = st[7:].reshape(G, 4)
    return dict(G=st[0], P=st[1], maxloc=st[2], maxslots=st[3], work_max=st[4], work_full=st[5],
                bad=st[6], n_loc=parts[:, 0], n_own=parts[:, 1], n_halo=parts[:, 2], n_bnd=parts[:, 3])


def pyrochlore_heisenberg():
    uc = Pyrochlore()
    for b1, b2, o in W.PYRO_BONDS:
        addBilinear(uc, b1, b2, np.eye(3), o)
    return uc


LATTICES = [
    ("honeycomb_64x64", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64), 8192),
    ("honeycomb_40x8", lambda: W.honeycomb_kitaev(1.0, -2.0), (40, 8), 640),
    ("square_30x6", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (30, 6), 180),
    ("triangular_24x5", lambda: W.triangular_dm(-0.9, -0.05, [0.3, 0.1, 0.0]), (24, 5), 120),
    ("pyrochlore_8x3x3", pyrochlore_heisenberg, (8, 3, 3), 288),
]


@pytest.mark.parametrize("lat", LATTICES, ids=[x[0] for x in LATTICES])
@pytest.mark.parametrize("G", [2, 3, 4])
@pytest.mark.parametrize("nor", [0, 1, 3, 10])
def test_split_plan_is_sound(lat, G, nor):
    name, mk, shape, N = lat
    p = plan(mk(), shape, G, nor)
    assert p["bad"] == 0
    assert p["G"] == G
    assert p["n_own"].sum() == N
    # strips of floor/ceil(L0 / G) cell layers
    L0 = shape[0]
    per_layer = N // L0
    sizes = np.diff([(g * L0) // G for g in range(G + 1)])
    assert sorted(p["n_own"]) == sorted(sizes * per_layer)
    assert np.all(p["n_loc"] <= N) and np.all(p["n_loc"] >= p["n_own"])
    assert np.array_equal(p["n_halo"], p["n_loc"] - p["n_own"])
    assert np.all(p["n_bnd"] <= p["n_own"])
    assert p["work_max"] <= p["work_full"]
    assert p["maxloc"] == p["n_loc"].max()


def test_honeycomb_halo_depth():
    """honeycomb, strips along a1: every colour phase reaches one A-B bond
    further, and the window-end energy one more, so the halo is nor + 1 cell
    layers on each side (2 sites a layer, 64 layers across)"""
    for nor in (0, 1, 3, 10):
        p = plan(W.honeycomb_kitaev(1.0, -2.0), (64, 64), 4, nor)
        assert p["bad"] == 0
        # 16 owned layers; each side needs nor + 1 layers for the sweeps
        assert np.all(p["n_own"] == 16 * 128)
        assert np.all(p["n_loc"] <= (16 + 2 * (nor + 1) + 2) * 128)
        assert np.all(p["n_loc"] >= (16 + 2 * (nor + 1) - 2) * 128)


def test_split_cost_of_the_headline_workload():
    """honeycomb L=64, 10 overrelaxation sweeps per Metropolis sweep: four
    parts do well under half of the whole system's work each"""
    p = plan(W.honeycomb_kitaev(1.0, -2.0), (64, 64), 4, 10)
    assert p["work_max"] / p["work_full"] < 0.5
    assert p["maxslots"] <= 8


def test_split_plan_rejects_bad_requests():
    uc = W.honeycomb_kitaev(1.0, -2.0)
    with pytest.raises(_lib.CsmcError):
        plan(uc, (3, 8), 4, 1)  # fewer cell layers than parts
    with pytest.raises(_lib.CsmcError):
        plan(uc, (16, 16), 9, 1)  # more parts than the kernel holds
