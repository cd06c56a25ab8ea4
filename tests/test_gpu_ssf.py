"""Equal-time spin correlations on the GPU (csmc_spin_correlations) against
the oracle's restatement of compute_equal_time_correlations
(spin_correlations.jl:6-43), itself pinned to closed forms and an FFT on the
host.  Both device paths: the fused FP64 MFMA kernel (default) and the
rocBLAS path (CSMC_SSF=blas).  FP64 and FP32 spin storage (the FP32 spins
widen exactly, so both compare at FP64 accuracy), 2D and 3D lattices, site
and momentum counts off the tile sizes, 1 to 70 systems in one call (every
column-tile width of the fused kernel), more momenta than one rocBLAS block
holds, and the ferromagnet's S^zz(0) = N S^2."""
import os

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import _lib, workloads as W
from csmc.reciprocal import get_allowed_wavevectors

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-12):
    # relative to the largest entry: S(k) spans many orders of magnitude
    return np.max(np.abs(a - b)) <= rtol * max(1.0, float(np.max(np.abs(b))))


def _ctx(path):
    old = os.environ.get("CSMC_SSF")
    os.environ["CSMC_SSF"] = path
    try:
        return csmc.Context(0)
    finally:
        if old is None:
            os.environ.pop("CSMC_SSF", None)
        else:
            os.environ["CSMC_SSF"] = old


CASES = [  # name, cell, shape, S, precision, systems, momenta
    ("honeycomb_fp64_R3", lambda: W.honeycomb_kitaev(1.0, -2.0), (8, 8), 0.5, 64, 3, "allowed"),
    ("honeycomb_fp32_R4_N120", lambda: W.honeycomb_kitaev(1.0, -2.0), (6, 10), 0.5, 32, 4, "allowed"),
    ("honeycomb_fp32_R6_N70", lambda: W.honeycomb_kitaev(1.0, -2.0), (5, 7), 0.5, 32, 6, "allowed"),
    ("pyrochlore_fp32_R2", W.pyrochlore_mixed, (3, 3, 3), 1.0, 32, 2, "allowed"),
    # rows of 32 sites: every chunk of 16 is regular, the phases rotate
    ("honeycomb_fp32_R5_rotating", lambda: W.honeycomb_kitaev(1.0, -2.0), (16, 32), 0.5, 32, 5, "allowed"),
    ("square_fp64_R1", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (9, 9), 1.0, 64, 1, "allowed"),
    ("square_fp64_R2_2500k", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (16, 16), 1.0, 64, 2, "random"),
    ("square_fp64_R2_bigk", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (16, 16), 1.0, 64, 2, "big"),
    ("square_fp32_R16", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (12, 12), 1.0, 32, 16, "random"),
    ("square_fp32_R22", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (10, 10), 1.0, 32, 22, "random"),
    ("square_fp32_R40", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (10, 10), 1.0, 32, 40, "random"),
    ("square_fp32_R70", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (10, 10), 1.0, 32, 70, "random"),
]


@pytest.mark.parametrize("path", ["mfma", "blas"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_batch_correlations_match_oracle(case, path):
    name, mk, shape, S, prec, R, kind = case
    print(f"[ssf] {name} {path}", flush=True)
    uc = mk()
    lat = csmc.Lattice(shape, uc, S, seed=5)
    if kind == "allowed":
        ks = get_allowed_wavevectors(uc, shape, min=-1, max=1)
    elif kind == "big":
        # |k.r| up to ~3e5: beyond the fused kernel's own sincos (1e5), so the
        # library sincos runs; the argument k.r itself differs between two
        # FP64 evaluation orders by ~|k.r| 1e-16, hence the looser bound
        ks = np.random.default_rng(9).uniform(-1e4, 1e4, size=(len(shape), 300))
    else:
        # 2500: more momenta than one block of the rocBLAS path (2048) holds
        n_k = 2500 if "2500k" in name else 333
        ks = np.random.default_rng(9).uniform(-7.0, 7.0, size=(len(shape), n_k))
    rtol = 1e-9 if kind == "big" else 1e-12
    ctx = _ctx(path)
    b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
    b.randomize(11)
    got = b.spin_correlations(ks)
    assert got.shape == (R, 9, ks.shape[1])
    for r in sorted({0, R // 2, R - 1}):
        ref = orc.equal_time_correlations(lat.site_positions, b.get_spins(r), ks)
        assert _close(got[r], ref, rtol), (name, r)
    # S^ab = S^ba holds for every system
    sw = got.reshape(R, 3, 3, -1)
    assert _close(sw, np.transpose(sw, (0, 2, 1, 3)), 1e-13)
    # a second call on the same buffers gives the same result
    assert np.array_equal(b.spin_correlations(ks), got)
    b.close()
    ctx.close()


def test_lattice_form_and_ferromagnet():
    uc = W.honeycomb_kitaev(1.0, -2.0)
    shape, S = (12, 12), 1.5
    lat = csmc.Lattice(shape, uc, S, seed=2)
    ks = get_allowed_wavevectors(uc, shape)
    got = csmc.compute_equal_time_correlations(lat, ks)
    ref = orc.equal_time_correlations(lat.site_positions, lat.spins.T, ks)
    assert got.shape == (9, ks.shape[1])
    assert _close(got, ref)
    fm = np.zeros((lat.size, 3))
    fm[:, 2] = S
    lat.spins = fm
    got = csmc.compute_equal_time_correlations(lat, np.zeros((2, 1)))
    assert got[8, 0] == pytest.approx(lat.size * S * S, rel=1e-13)
    assert np.max(np.abs(got[:8, 0])) < 1e-12


def test_full_size_batch_against_oracle_on_a_subset():
    """the bench lattice (honeycomb 64 x 64, 8192 sites) with 64 FP32
    replicas in one call; two systems checked against the oracle"""
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((64, 64), uc, 0.5, seed=4)
    R = 64
    ks = get_allowed_wavevectors(uc, (16, 16), min=-1, max=1)
    print(f"[ssf] full size N={lat.size} R={R} n_k={ks.shape[1]}", flush=True)
    b = csmc.SpinBatch([lat] * R, precision=32)
    b.randomize(21)
    got = b.spin_correlations(ks)
    for r in (0, R - 1):
        ref = orc.equal_time_correlations(lat.site_positions, b.get_spins(r), ks)
        assert _close(got[r], ref), r
    # fewer momenta on the same buffers
    few = b.spin_correlations(ks[:, :5])
    assert _close(few, got[:, :, :5])
    b.close()


def test_rejects_mixed_lattices_and_bad_momenta():
    a = csmc.Lattice((4, 4), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=1)
    c = csmc.Lattice((6, 6), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=1)
    b = csmc.SpinBatch([a, c], precision=64)
    with pytest.raises(_lib.CsmcError, match="share one lattice"):
        b.spin_correlations(np.zeros((2, 3)))
    with pytest.raises(ValueError):
        b.spin_correlations(np.zeros(3))
    b.close()
