"""Host plan of the tiled batch-wide sweeps (csmc_tile_plan, no device
needed) and a CPU replay of the tile kernel's index arithmetic.

The kernel (tile.hip) sweeps a lattice tile by tile from LDS: it stages a box
of cells around the tile from the device-order spin records, runs every
colour step over the step's region (the tile plus the halo sites later steps
read), finds partners by adding a per-(basis, slot) constant to the box
index, and stores the tile.  `replay_tiles` below walks the same loops with
the same integer arithmetic in Python and applies, at each update, an
order-independent hash of the site's reference index, its value and its
partners' values.  The result must equal the colour-pass schedule over the
reference tables (lattice.jl:15-22: partners from bsite, colours from the
colouring) bit for bit, and every index the kernel would form must stay in
its buffer.  A tile plan that dropped a halo site, used a wrong partner
offset, or mapped a cell to the wrong device record fails here."""
import ctypes as C

import numpy as np
import pytest

from csmc import _lib, workloads as W
from csmc.lattice import Lattice
from csmc.unitcell import Pyrochlore, UnitCell, addBasisSite, addBilinear, addZeemanCoupling

HEAD = ["D", "nb", "K2", "p", "cz", "cinv", "ncol", "ntiles"]
TILE_PAD = 0x7FFFFFFF
M64 = (1 << 64) - 1


def tile_plan(uc, shape, prec=32, budget=48 << 10, periodic=1, nsw=1):
    if len(uc.basis) == 0:  # the Lattice constructor's default site (lattice.jl:68-70)
        addBasisSite(uc, np.zeros(uc.D))
    s, keep = uc._pack()
    sh = np.ascontiguousarray(np.array(shape, np.int32))
    head = np.zeros(32, np.int32)
    n = C.c_int32(0)
    _lib.check(_lib.lib().csmc_tile_plan(C.byref(s), _lib.ptr(sh), periodic, prec, budget, nsw, _lib.ptr(head),
                                         None, 0, C.byref(n)))
    if head[1] == 0:
        return None, _lib.lib().csmc_last_error().decode()
    tab = np.zeros(n.value, np.int32)
    _lib.check(_lib.lib().csmc_tile_plan(C.byref(s), _lib.ptr(sh), periodic, prec, budget, nsw, _lib.ptr(head),
                                         _lib.ptr(tab), n.value, C.byref(n)))
    t = {k: int(head[i]) for i, k in enumerate(HEAD)}
    for i, k in enumerate(["L", "T", "lo", "B", "nt", "c"]):
        t[k] = [int(v) for v in head[8 + 3 * i: 11 + 3 * i]]
    t.update(boxcells=int(head[26]), nbox=int(head[27]), tcells=int(head[28]), ncell=int(head[29]),
             nsw=int(head[31]))
    nb, p, K2, ncol = t["nb"], t["p"], t["K2"], t["ncol"]
    o = 0
    t["eb"] = tab[o:o + nb].tolist(); o += nb
    t["runbase"] = tab[o:o + p * nb].tolist(); o += p * nb
    t["delta"] = tab[o:o + nb * K2].tolist(); o += nb * K2
    t["rect"] = tab[o:o + t["nsw"] * ncol * nb * 6].tolist(); o += t["nsw"] * ncol * nb * 6
    t["rawc"] = tab[o:o + ncol].tolist(); o += ncol
    assert o == len(tab)
    return t, ""


def mix(v):
    """splitmix64 finaliser on Python ints (mod 2^64)"""
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9 & M64
    v = (v ^ (v >> 27)) * 0x94D049BB133111EB & M64
    return v ^ (v >> 31)


def update(ref, own, partners, sweep):
    """a stand-in for one site update: depends on the site, the sweep (the
    Philox counter of a Metropolis update), its value and the multiset of its
    partners' values"""
    acc = (own * 0x9E3779B97F4A7C15 + ref * 0xD6E8FEB86659FD93 + sweep * 0xA0761D6478BD642F) & M64
    for v in partners:
        acc = (acc + mix(v)) & M64
    return mix(acc)


def device_order(colour, ncell):
    """device order (colour, basis, reference index) of the lattice"""
    N = len(colour)
    ref = np.arange(N)
    order = np.lexsort((ref, ref // ncell, colour))
    dev_of_ref = np.empty(N, np.int64)
    dev_of_ref[order] = np.arange(N)
    return order, dev_of_ref


def pass_schedule(vals, bsite, colour, ncol, sweeps):
    """the colour-pass schedule on reference-indexed values"""
    vals = list(vals)
    for sw in range(sweeps):
        for col in range(ncol):
            new = {}
            for s in np.flatnonzero(colour == col):
                new[s] = update(int(s), vals[s], [vals[j - 1] for j in bsite[s] if j > 0], sw)
            for s, v in new.items():
                vals[s] = v
    return vals


def imod(a, m):
    r = a % m  # Python's % is already non-negative for m > 0
    return r


def replay_tiles(t, G, N, sweeps):
    """the loops of k_tile over every tile of one system, reading G (device
    order) and writing a second buffer, once per launch of nsw sweeps"""
    nb, K2, p, cz, cinv = t["nb"], t["K2"], t["p"], t["cz"], t["cinv"]
    Lx, T, lo, B, nt, c = t["L"], t["T"], t["lo"], t["B"], t["nt"], t["c"]
    boxcells, nbox, tcells = t["boxcells"], t["nbox"], t["tcells"]
    eb, runbase, delta, rect, rawc = t["eb"], t["runbase"], t["delta"], t["rect"], t["rawc"]

    def dev_index(b, x):
        ci = (x[0] * Lx[1] + x[1]) * Lx[2] + x[2]
        if cz:
            d = runbase[eb[b] * nb + b] + ci
        else:
            r = imod(eb[b] + c[0] * x[0] + c[1] * x[1] + c[2] * x[2], p)
            d = runbase[r * nb + b] + (ci // Lx[2]) * (Lx[2] // p) + x[2] // p
        assert 0 <= d < N
        return d

    def ref_of(b, x):
        return ((b * Lx[0] + x[0]) * Lx[1] + x[1]) * Lx[2] + x[2]

    nsw, ncol = t["nsw"], t["ncol"]
    assert sweeps % nsw == 0
    for launch in range(sweeps // nsw):
        out = [None] * N
        for tile in range(t["ntiles"]):
            org, r = [0, 0, 0], tile
            for d in (2, 1, 0):
                q = r // nt[d]
                org[d] = (r - q * nt[d]) * T[d]
                r = q
            X = [0] * (nbox + 1)  # index nbox: the zero pad site
            for li in range(nbox):
                b = li // boxcells
                cc, x = li - b * boxcells, [0, 0, 0]
                for d in (2, 1, 0):
                    q = cc // B[d]
                    v = org[d] + lo[d] + (cc - q * B[d])
                    v += Lx[d] if v < 0 else 0
                    v -= Lx[d] if v >= Lx[d] else 0
                    x[d] = v
                    cc = q
                X[li] = G[dev_index(b, x)]
            for st in range(nsw * ncol):
                col, sw = st % ncol, launch * nsw + st // ncol
                r0 = rawc[col]
                for b in range(nb):
                    rc = rect[(st * nb + b) * 6:(st * nb + b) * 6 + 6]
                    lo0, hi0, lo1, hi1, lo2, hi2 = rc
                    if lo0 >= hi0 or lo1 >= hi1 or lo2 >= hi2:
                        continue
                    w2 = hi2 - lo2 if cz else (hi2 - lo2 + p - 1) // p
                    n1 = hi1 - lo1
                    new = {}
                    for w in range((hi0 - lo0) * n1 * w2):
                        row, j = w // w2, w % w2
                        q1 = row // n1
                        u = [lo0 + q1, lo1 + (row - q1 * n1), lo2 + j]
                        if not cz:
                            s = r0 - eb[b] - c[0] * (lo[0] + u[0]) - c[1] * (lo[1] + u[1])
                            u[2] = lo2 + imod(s * cinv - lo[2] - lo2, p) + j * p
                            if u[2] >= hi2:
                                continue
                        li = b * boxcells + (u[0] * B[1] + u[1]) * B[2] + u[2]
                        assert 0 <= li < nbox
                        parts = []
                        for k in range(K2):
                            dl = delta[b * K2 + k]
                            if dl == TILE_PAD:
                                continue
                            assert 0 <= li + dl < nbox, (li, dl)
                            parts.append(X[li + dl])
                        x = [(org[d] + u[d] + lo[d]) % Lx[d] for d in range(3)]
                        new[li] = update(ref_of(b, x), X[li], parts, sw)
                    for li, v in new.items():  # one colour class: independent updates
                        X[li] = v
            for oi in range(nb * tcells):
                b = oi // tcells
                cc, x, u = oi - b * tcells, [0, 0, 0], [0, 0, 0]
                for d in (2, 1, 0):
                    q = cc // T[d]
                    v = cc - q * T[d]
                    u[d] = v - lo[d]
                    x[d] = org[d] + v
                    cc = q
                li = b * boxcells + (u[0] * B[1] + u[1]) * B[2] + u[2]
                d = dev_index(b, x)
                assert out[d] is None, "a site stored twice"
                out[d] = X[li]
        assert all(v is not None for v in out), "a site not stored"
        G = out
    return G


def fdiv(a, d):
    """k_tile's quotient from a float reciprocal: trunc(float32(a) *
    float32(1 / d)), then one correction from the remainder"""
    rd = np.float32(1.0) / np.float32(d)
    q = int(np.float32(a) * rd)
    r = a - q * d
    return q + (-1 if r < 0 else (1 if r >= d else 0))


def walk3(c, stride, n1, n2, count):
    """the cells k_tile's Walk3 visits from cell c by `stride`, as
    (u0, u1, u2), for the first `count` steps"""
    n12 = n1 * n2
    u0 = fdiv(c, n12)
    rc = c - u0 * n12
    u1 = fdiv(rc, n2)
    u2 = rc - u1 * n2
    d0 = fdiv(stride, n12)
    rs = stride - d0 * n12
    d1 = fdiv(rs, n2)
    d2 = rs - d1 * n2
    out = []
    for _ in range(count):
        out.append((u0, u1, u2))
        u2 += d2
        c2 = 1 if u2 >= n2 else 0
        u2 -= n2 if c2 else 0
        u1 += d1 + c2
        c1 = 1 if u1 >= n1 else 0
        u1 -= n1 if c1 else 0
        u0 += d0 + c1
    return out


def xxz_chain():
    """a one-dimensional chain (two raw colours, x mod 2), diagonal couplings
    whose products round, and a field"""
    uc = UnitCell([1.0])
    addBasisSite(uc, np.zeros(1))
    for d in ((1,), (-1,)):
        addBilinear(uc, 1, 1, np.diag([0.7, 0.7, -1.3]), d)
    addZeemanCoupling(uc, 1, np.array([0.0, 0.1, 0.2]))
    return uc


def pyrochlore_heisenberg():
    uc = Pyrochlore()
    for b1, b2, o in W.PYRO_BONDS:
        addBilinear(uc, b1, b2, np.eye(3), o)
    return uc


CASES = [
    ("honeycomb_64x64", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64)),
    ("honeycomb_40x8", lambda: W.honeycomb_kitaev(1.0, -2.0), (40, 8)),
    ("honeycomb_4x4", lambda: W.honeycomb_kitaev(1.0, -2.0), (4, 4)),  # the box is the whole lattice
    ("square_64x64", lambda: W.square_heisenberg(), (64, 64)),
    ("triangular_dm_48x48", lambda: W.triangular_dm(-1.0, 0.1, (0.0, 0.0, 0.2)), (48, 48)),
    ("pyrochlore_8", pyrochlore_heisenberg, (8, 8, 8)),
    ("pyrochlore_12", pyrochlore_heisenberg, (12, 12, 12)),
    ("xxz_chain_512", xxz_chain, (512,)),
]


@pytest.mark.parametrize("nsw", [1, 2])
@pytest.mark.parametrize("name,make,shape", CASES, ids=[c[0] for c in CASES])
def test_tile_replay_equals_colour_passes(name, make, shape, nsw):
    uc = make()
    t, why = tile_plan(uc, shape, budget=(16 << 10) * nsw * nsw, nsw=nsw)
    if nsw > 1 and name in ("honeycomb_4x4", "pyrochlore_8"):
        # two sweeps' halo needs a box wider than the lattice (pyrochlore:
        # tiles of 2^3 cells own fewer than 256 sites)
        assert t is None and "no tile shape fits" in why
        return
    assert t is not None, why
    assert t["ntiles"] >= 2, "the case must cut the lattice into several tiles"
    lat = Lattice(shape, uc, 0.5, seed=1)
    tb = lat.tables()
    N = lat.size
    colour = np.asarray(tb["colour"])
    order, dev_of_ref = device_order(colour, t["ncell"])
    rng = np.random.default_rng(3)
    vals = [int(v) for v in rng.integers(0, 1 << 62, size=N)]
    want = pass_schedule(vals, np.asarray(tb["bsite"]), colour, t["ncol"], sweeps=2)
    G = [0] * N
    for s in range(N):
        G[dev_of_ref[s]] = vals[s]
    got = replay_tiles(t, G, N, sweeps=2)
    assert all(got[dev_of_ref[s]] == want[s] for s in range(N))


@pytest.mark.parametrize("nsw", [1, 2])
def test_planted_halo_mistake_is_caught(nsw):
    """the replay notices a plan whose update region misses one halo row
    (with two sweeps a launch: a row of the first sweep's halo)"""
    uc = W.square_heisenberg()
    t, _ = tile_plan(uc, (32, 32), budget=16 << 10, nsw=nsw)
    lat = Lattice((32, 32), uc, 0.5, seed=1)
    tb = lat.tables()
    colour = np.asarray(tb["colour"])
    order, dev_of_ref = device_order(colour, t["ncell"])
    vals = [int(v) for v in np.random.default_rng(4).integers(0, 1 << 62, size=lat.size)]
    want = pass_schedule(vals, np.asarray(tb["bsite"]), colour, t["ncol"], sweeps=nsw)
    G = [0] * lat.size
    for s in range(lat.size):
        G[dev_of_ref[s]] = vals[s]
    # colour 0's region one row shorter at its low end
    rect = list(t["rect"])
    rect[2] += 1
    bad = dict(t, rect=rect)
    assert replay_tiles(t, G, lat.size, sweeps=nsw) == [want[s] for s in order]
    got = replay_tiles(bad, G, lat.size, sweeps=nsw)
    assert any(got[dev_of_ref[s]] != want[s] for s in range(lat.size))


def test_plan_geometry():
    """tile and box sizes that the geometry fixes"""
    th, _ = tile_plan(W.honeycomb_kitaev(1.0, -2.0), (512, 512))
    # the colour of a honeycomb site is its basis; each site's partners are
    # one cell away at most, so the box is the 32 x 32 tile plus one cell
    # on each side
    assert th["cz"] == 1 and th["ncol"] == 2 and th["T"] == [1, 32, 32]
    assert th["B"] == [1, 34, 34] and th["lo"] == [0, -1, -1] and th["ntiles"] == 256
    assert th["nbox"] == 2 * 34 * 34
    # square lattice, checkerboard colouring c = (1, 1) mod 2: the colour-0
    # steps reach one cell past the tile and read one more
    t, _ = tile_plan(W.square_heisenberg(), (64, 64))
    assert t["cz"] == 0 and t["p"] == 2 and t["c"][1:] == [1, 1]
    assert t["lo"][1:] == [-2, -2]
    # FP64 records take twice the LDS: the tile shrinks under the same budget
    t64, _ = tile_plan(W.honeycomb_kitaev(1.0, -2.0), (512, 512), prec=64)
    assert (t64["nbox"] + 1) * 24 <= 48 << 10 and t64["nbox"] < th["nbox"]


def test_two_sweep_geometry():
    """two sweeps a launch: the honeycomb box grows by one cell a side (the
    halo of the first sweep's partners), and the regions shrink from the
    first step to the last, which covers the tile alone"""
    t, _ = tile_plan(W.honeycomb_kitaev(1.0, -2.0), (512, 512), nsw=2)
    assert t["nsw"] == 2 and t["T"] == [1, 32, 32] and t["B"] == [1, 36, 36] and t["lo"] == [0, -2, -2]
    nb, S = t["nb"], 2 * t["ncol"]
    size = []
    for st in range(S):
        cells = 0
        for b in range(nb):
            lo0, hi0, lo1, hi1, lo2, hi2 = t["rect"][(st * nb + b) * 6:(st * nb + b) * 6 + 6]
            cells += max(0, hi0 - lo0) * max(0, hi1 - lo1) * max(0, hi2 - lo2)
        size.append(cells)
    assert size == sorted(size, reverse=True) and size[-1] == 32 * 32
    # the last step's region is the tile itself, box cells [2, 34)
    last = [t["rect"][((S - 1) * nb + b) * 6:((S - 1) * nb + b) * 6 + 6] for b in range(nb)]
    assert [0, 1, 2, 34, 2, 34] in last


def test_sweeps_argument_checked():
    """csmc_tile_plan takes 1..8 sweeps a launch"""
    for bad in (0, 9):
        with pytest.raises(_lib.CsmcError):
            tile_plan(W.honeycomb_kitaev(1.0, -2.0), (64, 64), nsw=bad)


def test_no_plan_reasons():
    assert tile_plan(W.pyrochlore_mixed(), (8, 8, 8))[1] == "cubic or quartic terms"
    assert tile_plan(W.honeycomb_kitaev(1.0, -2.0), (64, 64), periodic=0)[1] == "open boundaries"
    # too small for a box that does not wrap onto itself
    assert tile_plan(W.honeycomb_kitaev(1.0, -2.0), (2, 2))[0] is None


@pytest.mark.parametrize("nsw", [1, 2])
@pytest.mark.parametrize("name,make,shape", CASES, ids=[c[0] for c in CASES])
def test_kernel_walks_equal_divmod(name, make, shape, nsw):
    """the strided cell walks of k_tile (box staging, each colour step's
    region, the tile store) visit the cells that row-major division gives,
    for workgroups of 256 and 512 threads"""
    t, _ = tile_plan(make(), shape, budget=(16 << 10) * nsw * nsw, nsw=nsw)
    if t is None:
        return
    nb, p, cz = t["nb"], t["p"], t["cz"]
    blocks = [(t["B"], t["boxcells"]), (t["T"], t["tcells"])]
    for st in range(nsw * t["ncol"]):
        for b in range(nb):
            lo0, hi0, lo1, hi1, lo2, hi2 = t["rect"][(st * nb + b) * 6:(st * nb + b) * 6 + 6]
            if lo0 >= hi0 or lo1 >= hi1 or lo2 >= hi2:
                continue
            w2 = hi2 - lo2 if cz else (hi2 - lo2 + p - 1) // p
            blocks.append(([hi0 - lo0, hi1 - lo1, w2], (hi0 - lo0) * (hi1 - lo1) * w2))
    for (n0, n1, n2), cnt in blocks:
        for nt in (256, 512):
            for tid in (0, 1, 37, nt - 1):
                if tid >= cnt:
                    continue
                k = (cnt - tid + nt - 1) // nt
                got = walk3(tid, nt, n1, n2, k)
                want = [(c // (n1 * n2), (c // n2) % n1, c % n2) for c in range(tid, cnt, nt)]
                assert got == want, (name, n0, n1, n2, nt, tid)


def test_fdiv_exact():
    """the float-reciprocal quotient is exact over the ranges k_tile uses"""
    rng = np.random.default_rng(5)
    for d in list(range(1, 300)) + [1023, 4096, 65535, (1 << 24) - 1]:
        for a in list(rng.integers(0, 1 << 24, size=200)) + [0, d - 1, d, (1 << 24) - 1]:
            assert fdiv(int(a), d) == int(a) // d, (a, d)
