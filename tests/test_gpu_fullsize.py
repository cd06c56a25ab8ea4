"""The workloads at their full sizes, checked through properties that do not
depend on size (and against the oracle where it finishes in seconds):

* pyrochlore L=16 (16384 spins, cubic + quartic + onsite terms): local
  fields and energies of two FP64 replicas equal the oracle's to 1e-12;
  an FP32 parallel-tempering run keeps |s| = S and finite energies;
* triangular J1-J2 L=96 (9216 spins), a batch of 128 configurations:
  dE/dS equals the local field (symmetric couplings, reverse bonds present),
  dE/dJ1 summed over the J1 terms equals the J1 energy, and central finite
  differences of the oracle energy at random sites;
* the uneven batch of 256 lattices (L 16..128): every system's FP64 energy
  equals the oracle's."""
import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from conftest import random_spins

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b))))


def test_pyrochlore_full_size():
    uc = W.pyrochlore_mixed()
    lat = csmc.Lattice((16, 16, 16), uc, 1.0, seed=4)
    T = orc.Tables(uc, (16, 16, 16))
    rng = np.random.default_rng(1)
    confs = [random_spins(lat.size, 1.0, rng) for _ in range(2)]
    b = csmc.SpinBatch([lat] * 2, precision=64)
    for r, s in enumerate(confs):
        b.set_spins(r, s)
    E = b.total_energy()
    for r, s in enumerate(confs):
        assert abs(E[r] - orc.total_energy(T, s)) <= 1e-12 * lat.size
        idx = rng.choice(lat.size, 64, replace=False)
        H = b.local_field(r)
        Ho = np.array([orc.local_field(T, s, int(i) + 1) for i in idx])
        assert _rel(H[idx], Ho) < 1e-12
    b.close()
    R = 32
    pb = csmc.SpinBatch([lat] * R, precision=32)
    pb.randomize(4)
    p = dict(t_thermalization=20, t_measurement=40, probe_rate=10, swap_rate=10, overrelaxation_rate=10,
             report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    pb.pt_init(W.geometric_temperatures(0.02, 2.0, R), p, csmc.METROPOLIS, seed=4)
    while pb.pt_advance(-1)[1] != 3:
        pass
    st = pb.pt_state()
    assert np.all(np.isfinite(st["E"]))
    for r in (0, R - 1):
        assert np.allclose(np.linalg.norm(pb.get_spins(r), axis=1), 1.0, atol=1e-5)
    pb.close()


def test_triangular_j1j2_adjoint_full_size():
    uc = W.triangular_j1j2(1.0, 0.13)
    lat = csmc.Lattice((96, 96), uc, 1.0, seed=5)
    n = 128
    rng = np.random.default_rng(2)
    b = csmc.SpinBatch([lat] * n, precision=64)
    confs = [random_spins(lat.size, 1.0, rng) for _ in range(n)]
    for r in range(n):
        b.set_spins(r, confs[r])
    gs = b.grad_spins()
    gJ = b.grad_couplings()
    uc1 = W.triangular_j1j2(1.0, 0.0)
    uc1.basis.append(np.zeros(2))
    T1 = orc.Tables(uc1, (96, 96))
    T = orc.Tables(uc, (96, 96))
    for r in (0, 77, n - 1):
        assert _rel(gs[r], b.local_field(r)) < 1e-12
        dJ1 = sum(np.trace(gJ[r, 9 * t:9 * t + 9].reshape(3, 3)) for t in range(6))
        assert abs(dJ1 - orc.total_energy(T1, confs[r])) < 1e-10 * lat.size
    s = confs[3]
    eps = 1e-6
    for i in rng.choice(lat.size, 4, replace=False):
        for a in range(3):
            sp, sm = s.copy(), s.copy()
            sp[i, a] += eps
            sm[i, a] -= eps
            fd = (orc.total_energy(T, sp) - orc.total_energy(T, sm)) / (2 * eps)
            assert abs(gs[3][i, a] - fd) < 1e-5 * max(1.0, abs(fd)), (i, a, gs[3][i, a], fd)
    b.close()


def test_uneven_batch_full_size_energies():
    cells = W.uneven_batch(n=256, seed=3)
    lats = [csmc.Lattice(shape, uc, 1.0, seed=3 + k) for k, (uc, shape) in enumerate(cells)]
    b = csmc.SpinBatch(lats, precision=64)
    for k, lat in enumerate(lats):
        b.set_spins(k, lat._spins)
    E = b.total_energy()
    for k in range(0, 256, 5):
        uc, shape = cells[k]
        Eo = orc.total_energy(orc.Tables(uc, shape), lats[k]._spins)
        assert abs(E[k] - Eo) <= 1e-12 * lats[k].size, k
    b.close()
