// tile.cpp -- host planning of the tiled sweeps of periodic Bravais lattices.
//
// Every term of a lattice built from a unit cell joins a site to partners at
// fixed unit-cell offsets under the periodic `mod` of lattice.jl:101-109
// (tables built at lattice.jl:176-206), so in cell coordinates the partner
// of slot k of a basis-b site at cell x is basis b_k at cell x + o_k.  When
// the colouring is the lattice-periodic one, colour(b, x) = remap[(c . x +
// e_b) mod p], the device order (colour, basis, reference index) is also
// arithmetic: the run of (colour, basis) starts at a known index and a
// cell's rank in it is its cell index (c = 0) or, when c[D-1] is invertible
// mod p, row * (L[D-1] / p) + x[D-1] / p.  Sweeps can then run tile by
// tile from LDS: a tile of T cells is loaded with the halo that nsw sweeps
// of it read, the halo sites that the tile's later colour steps read are
// updated too (recomputed, with the same counter-based random numbers, so
// they equal what their own tile computes), and only the tile is stored.
#include "tile.h"

#include <algorithm>
#include <cstdlib>

namespace csmc {
namespace {

int pmod(long a, int m) {
    const long r = a % m;
    return (int)(r < 0 ? r + m : r);
}

struct Off {
    int o[3] = {0, 0, 0};
    int bj = -1;  // -1: the slot names no site
};

struct Rect {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool empty() const { return lo[0] >= hi[0] || lo[1] >= hi[1] || lo[2] >= hi[2]; }
    long cells() const { return empty() ? 0 : (long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]); }
    void unite(const Rect &r) {
        if (r.empty()) return;
        if (empty()) {
            *this = r;
            return;
        }
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], r.lo[d]);
            hi[d] = std::max(hi[d], r.hi[d]);
        }
    }
    Rect shifted(const int *o) const {
        Rect r = *this;
        for (int d = 0; d < 3; ++d) {
            r.lo[d] += o[d];
            r.hi[d] += o[d];
        }
        return r;
    }
};

}  // namespace

TilePlan plan_tiles(const HostLattice &L, size_t rec_bytes, size_t lds_budget, int nsw) {
    TilePlan P;
    auto no = [&](const char *w) {
        P.ok = false;
        P.why = w;
        return P;
    };
    if (!L.periodic) return no("open boundaries");
    if (L.K3 || L.K4) return no("cubic or quartic terms");
    if (L.K2 == 0) return no("no bilinear terms");
    const ColourPattern &cp = L.pattern;
    if (cp.p <= 0) return no("the colouring is not lattice-periodic");
    const int D = L.dim, nb = L.nb, K2 = L.K2, p = cp.p;
    if (D < 1 || D > 3) return no("dimension");
    if (nb > 64 || K2 > 64) return no("basis or slot count");
    if (nsw < 1 || nsw > 8) return no("sweeps per launch");
    // dimensions right-aligned in three positions: lattice dimension d is
    // position d + 3 - D, so position 2 is always the fastest (last) one and
    // unused positions have extent 1
    const int sh = 3 - D;
    int Ls[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d) Ls[d + sh] = L.shape[d];
    int cc[3] = {0, 0, 0};
    bool cz = true;
    for (int d = 0; d < D; ++d) {
        cc[d + sh] = pmod(cp.c[d], p);
        cz = cz && cc[d + sh] == 0;
    }
    int cinv = 0;
    if (!cz) {
        for (int v = 1; v < p; ++v)
            if ((cc[2] * v) % p == 1) cinv = v;
        if (cinv == 0) return no("the colour pattern is not invertible along the last dimension");
    }
    const long ncell = L.ncell, N = L.N;
    // the kernel's index products are 24-bit multiplies: the last extent,
    // the row count over all bases and the raw colour's sum c . x + e below
    // 2^24
    const long lim24 = 1L << 24;
    if (Ls[2] >= lim24 || N / Ls[2] >= lim24 || (long)p * (Ls[0] + Ls[1] + Ls[2] + 1) >= lim24)
        return no("lattice extents beyond 24-bit index arithmetic");
    auto cellidx = [&](const int *x) { return ((long)x[0] * Ls[1] + x[1]) * Ls[2] + x[2]; };
    // reference site index s = ((b L0 + x0) L1 + x1) L2 + x2
    auto decompose = [&](long s, int *b, int *x) {
        long r = s;
        for (int d = 2; d >= 0; --d) {
            x[d] = (int)(r % Ls[d]);
            r /= Ls[d];
        }
        *b = (int)r;
    };

    // the offset and partner basis of each (basis, slot), read at a cell in
    // the middle of the lattice, then checked at every site
    std::vector<Off> off((size_t)nb * K2);
    int mid[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) mid[d] = Ls[d] / 2;
    for (int b = 0; b < nb; ++b) {
        const long s = b * ncell + cellidx(mid);
        for (int k = 0; k < K2; ++k) {
            const int j = L.nbr2[(size_t)s * K2 + k];
            if (j < 0) continue;
            Off &f = off[(size_t)b * K2 + k];
            int y[3];
            decompose(j, &f.bj, y);
            for (int d = sh; d < 3; ++d) {
                int o = y[d] - mid[d];
                if (2 * o > Ls[d]) o -= Ls[d];
                if (2 * o <= -Ls[d]) o += Ls[d];
                if (2 * std::abs(o) >= Ls[d]) return no("a term spans half the lattice");
                f.o[d] = o;
            }
        }
    }
    for (long s = 0; s < N; ++s) {
        int b, x[3];
        decompose(s, &b, x);
        for (int k = 0; k < K2; ++k) {
            const int j = L.nbr2[(size_t)s * K2 + k];
            const Off &f = off[(size_t)b * K2 + k];
            if (f.bj < 0) {
                if (j >= 0) return no("the neighbour table is not a periodic stencil");
                continue;
            }
            int y[3];
            for (int d = 0; d < 3; ++d) y[d] = pmod((long)x[d] + f.o[d], Ls[d]);
            if (j != f.bj * ncell + cellidx(y)) return no("the neighbour table is not a periodic stencil");
        }
    }

    // colours: raw (c . x + e_b) mod p, compacted by remap
    const int ncol = L.ncol;
    std::vector<int> rawc(ncol, -1);
    for (int r = 0; r < p; ++r)
        if (cp.remap[r] >= 0) rawc[cp.remap[r]] = r;
    for (int col = 0; col < ncol; ++col)
        if (rawc[col] < 0) return no("colour map");
    auto raw_of = [&](const int *x, int b) {
        long v = cp.e[b];
        for (int d = 0; d < 3; ++d) v += (long)cc[d] * x[d];
        return pmod(v, p);
    };
    std::vector<int> runbase((size_t)p * nb, -1);
    for (const auto &r : L.runs) {
        int &rb = runbase[(size_t)rawc[r[3]] * nb + r[2]];
        if (rb < 0 || r[0] < rb) rb = r[0];
    }
    const int Llast = Ls[2];
    for (long s = 0; s < N; ++s) {
        int b, x[3];
        decompose(s, &b, x);
        const int r = raw_of(x, b);
        if (L.colour[s] != cp.remap[r]) return no("colouring");
        const long base = runbase[(size_t)r * nb + b];
        const long ci = cellidx(x);
        const long rank = cz ? ci : (ci / Llast) * (Llast / p) + x[2] / p;
        if (base < 0 || base + rank != L.dev_of_ref[s]) return no("the device order is not the arithmetic one");
    }

    // tile shapes: T[d] divides L[d], and c[d] T[d] = 0 mod p so that every
    // tile sees the same colour pattern
    std::vector<int> cand[3];
    for (int d = 0; d < 3; ++d) {
        if (d < sh) {
            cand[d].push_back(1);
            continue;
        }
        const int tmax = D == 3 ? 16 : D == 2 ? 64 : 4096;
        for (int t = 1; t <= std::min(Ls[d], tmax); ++t)
            if (Ls[d] % t == 0 && (cc[d] * t) % p == 0) cand[d].push_back(t);
    }
    bool found = false;
    double best = 1e300;
    for (int t0 : cand[0])
        for (int t1 : cand[1])
            for (int t2 : cand[2]) {
                const int T[3] = {t0, t1, t2};
                // update regions of the nsw * ncol colour steps, from the
                // last back: a step of the last sweep covers the tile; a
                // step also covers the partners of every site a later step
                // updates, as of their latest update before that step (this
                // sweep for an earlier colour, the previous sweep for a later
                // one), and the sites the next sweep's step of its colour
                // updates (an update reads the site's own value)
                const int S = nsw * ncol;
                std::vector<Rect> rect((size_t)S * nb);
                for (int col = 0; col < ncol; ++col)
                    for (int b = 0; b < nb; ++b)
                        if (!cz || cp.remap[pmod(cp.e[b], p)] == col) {
                            Rect &r = rect[((size_t)(nsw - 1) * ncol + col) * nb + b];
                            for (int d = 0; d < 3; ++d) r.hi[d] = T[d];
                        }
                for (int st = S - 1; st >= 0; --st) {
                    const int col = st % ncol;
                    for (int b = 0; b < nb; ++b) {
                        const Rect rc = rect[(size_t)st * nb + b];
                        if (rc.empty()) continue;
                        if (st >= ncol) rect[(size_t)(st - ncol) * nb + b].unite(rc);
                        for (int k = 0; k < K2; ++k) {
                            const Off &f = off[(size_t)b * K2 + k];
                            if (f.bj < 0) continue;
                            long v = (long)rawc[col] + cp.e[f.bj] - cp.e[b];
                            for (int d = 0; d < 3; ++d) v += (long)cc[d] * f.o[d];
                            const int pc = cp.remap[pmod(v, p)];
                            if (pc == col) continue;  // the site itself (a self term)
                            const int ps = st - col + pc - (pc > col ? ncol : 0);
                            if (ps >= 0) rect[(size_t)ps * nb + f.bj].unite(rc.shifted(f.o));
                        }
                    }
                }
                Rect box;
                for (int d = 0; d < 3; ++d) box.hi[d] = T[d];
                long work = 0;
                for (int st = 0; st < S; ++st)
                    for (int b = 0; b < nb; ++b) {
                        const Rect &rc = rect[(size_t)st * nb + b];
                        if (rc.empty()) continue;
                        work += rc.cells() / (cz ? 1 : p);
                        box.unite(rc);
                        for (int k = 0; k < K2; ++k)
                            if (off[(size_t)b * K2 + k].bj >= 0) box.unite(rc.shifted(off[(size_t)b * K2 + k].o));
                    }
                int B[3];
                bool fits = true;
                for (int d = 0; d < 3; ++d) {
                    B[d] = box.hi[d] - box.lo[d];
                    fits = fits && B[d] <= Ls[d];
                }
                if (!fits) continue;
                const long boxcells = (long)B[0] * B[1] * B[2];
                const long nbox = boxcells * nb;
                if ((size_t)(nbox + 1) * rec_bytes > lds_budget || nbox + 1 > (1L << 20)) continue;
                const long owned = (long)nb * T[0] * T[1] * T[2];
                if (owned < std::min<long>(256, N / 4)) continue;
                const double load = (double)nbox / owned / nsw, wk = (double)work / owned / nsw;
                const double score = load + 0.5 * wk;
                if (found && !(score < best - 1e-9 || (score < best + 1e-9 && owned > (long)nb * P.d.tcells)))
                    continue;
                found = true;
                best = score;
                P.load_per_site = load;
                P.work_per_site = wk;
                TileDev &t = P.d;
                t = TileDev{};
                t.D = D;
                t.nb = nb;
                t.K2 = K2;
                t.p = p;
                t.cz = cz ? 1 : 0;
                t.cinv = cinv;
                t.ncol = ncol;
                t.nsw = nsw;
                t.ncell = (int32_t)ncell;
                t.ntiles = 1;
                for (int d = 0; d < 3; ++d) {
                    t.L[d] = Ls[d];
                    t.T[d] = T[d];
                    t.lo[d] = box.lo[d];
                    t.B[d] = B[d];
                    t.nt[d] = Ls[d] / T[d];
                    t.c[d] = cc[d];
                    t.ntiles *= t.nt[d];
                }
                t.boxcells = (int32_t)boxcells;
                t.nbox = (int32_t)nbox;
                t.tcells = T[0] * T[1] * T[2];
                // tables
                P.tab.clear();
                for (int b = 0; b < nb; ++b) P.tab.push_back(pmod(cp.e[b], p));
                P.tab.insert(P.tab.end(), runbase.begin(), runbase.end());
                const long stride[3] = {(long)B[1] * B[2], B[2], 1};
                for (int b = 0; b < nb; ++b)
                    for (int k = 0; k < K2; ++k) {
                        const Off &f = off[(size_t)b * K2 + k];
                        if (f.bj < 0) {
                            P.tab.push_back(TILE_PAD);
                            continue;
                        }
                        long dl = (long)(f.bj - b) * boxcells;
                        for (int d = 0; d < 3; ++d) dl += f.o[d] * stride[d];
                        P.tab.push_back((int32_t)dl);
                    }
                for (int st = 0; st < S; ++st)
                    for (int b = 0; b < nb; ++b) {
                        const Rect &rc = rect[(size_t)st * nb + b];
                        for (int d = 0; d < 3; ++d) {
                            P.tab.push_back(rc.empty() ? 0 : rc.lo[d] - box.lo[d]);
                            P.tab.push_back(rc.empty() ? 0 : rc.hi[d] - box.lo[d]);
                        }
                    }
                for (int col = 0; col < ncol; ++col) P.tab.push_back(rawc[col]);
            }
    if (!found) return no("no tile shape fits (lattice too small or LDS budget)");
    P.ok = true;
    return P;
}

}  // namespace csmc
