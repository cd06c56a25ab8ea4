"""csmc_energy_grad_couplings_noalloc (no counterpart in the reference: the
coupling gradient for callers that cannot allocate device memory) against
csmc_energy_grad_couplings, through the C-ABI: two calls in a row, the
second on a smaller input, each equal to the allocating entry point's
output to 1e-12 of the largest entry (both sum in LDS with atomics, so the
order of the additions may differ); and the error paths the host checks
before any launch."""
import numpy as np
import pytest

import csmc
from csmc import _lib
from csmc import workloads as W

pytestmark = pytest.mark.gpu


def _grads(b, fn):
    n = b.lattices[0].unit_cell.n_params()
    out = np.zeros((b.n, n))
    rc = getattr(_lib.lib(), fn)(b._h, b.id, _lib.ptr(out))
    return rc, out


def _check(b):
    rc0, ref = _grads(b, "csmc_energy_grad_couplings")
    assert rc0 == _lib.OK
    rc1, got = _grads(b, "csmc_energy_grad_couplings_noalloc")
    assert rc1 == _lib.OK, _lib.lib().csmc_last_error()
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, float(np.max(np.abs(ref))))
    return got


def test_noalloc_coupling_gradients_match_allocating_entry_point():
    ctx = csmc.Context(0)
    rng = np.random.default_rng(9)
    # first call: 24 triangular J1-J2 systems, L=24, FP64
    cells = [W.triangular_j1j2(1.0, float(j)) for j in rng.uniform(0.0, 0.2, size=24)]
    lats = [csmc.Lattice((24, 24), uc, 1.0, seed=k) for k, uc in enumerate(cells)]
    b1 = csmc.SpinBatch(lats, precision=64, ctx=ctx)
    b1.randomize(3)
    g1 = _check(b1)
    assert g1.shape[0] == 24
    # second call, smaller input: 3 pyrochlore systems (cubic and quartic
    # terms, a longer gradient), FP32
    p = csmc.Lattice((2, 2, 2), W.pyrochlore_mixed(), 1.0, seed=4)
    b2 = csmc.SpinBatch([p] * 3, precision=32, ctx=ctx)
    b2.randomize(5)
    g2 = _check(b2)
    assert g2.shape[0] == 3
    # the first batch again after the smaller call: the shared array holds
    # the last call's contents, every entry is rewritten
    _check(b1)
    b1.close()
    b2.close()
    ctx.close()


def test_noalloc_refuses_what_it_cannot_do_without_allocating():
    ctx = csmc.Context(0)
    lib = _lib.lib()
    # lattice tables not uploaded yet: refused (an upload would allocate)
    lat = csmc.Lattice((2, 2, 2), W.pyrochlore_mixed(), 1.0, seed=1)
    b = csmc.SpinBatch([lat] * 2, precision=64, ctx=ctx)
    out = np.zeros((2, lat.unit_cell.n_params()))
    assert lib.csmc_energy_grad_couplings_noalloc(b._h, b.id, _lib.ptr(out)) == _lib.ESTATE
    assert b"not on the device" in lib.csmc_last_error()
    b.randomize(1)
    assert lib.csmc_energy_grad_couplings_noalloc(b._h, b.id, _lib.ptr(out)) == _lib.OK
    # more systems than the scratch holds: refused before any launch
    n = lat.unit_cell.n_params()
    big = (1 << 20) // (8 * n) + 1
    bb = csmc.SpinBatch([lat] * big, precision=64, ctx=ctx)
    bb.randomize(2)
    out = np.zeros((big, n))
    assert lib.csmc_energy_grad_couplings_noalloc(bb._h, bb.id, _lib.ptr(out)) == _lib.EINVAL
    assert b"scratch" in lib.csmc_last_error()
    assert lib.csmc_energy_grad_couplings_noalloc(bb._h, bb.id, None) == _lib.EINVAL
    bb.close()
    b.close()
    ctx.close()
