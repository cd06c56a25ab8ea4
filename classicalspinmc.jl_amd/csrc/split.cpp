// split.cpp -- host planning of the split register-table PT kernel (split.h).
#include "split.h"

#include <algorithm>

namespace csmc {

namespace {

// first cell coordinate of a reference site index (lattice.cpp: site =
// ((b * s0 + x0) * s1 + x1) * s2 + x2)
int first_coord(const HostLattice &L, long ref) {
    const long inner = L.ncell / L.shape[0];
    return (int)((ref / inner) % L.shape[0]);
}

long waves(int n) { return (n + 63) / 64; }

}  // namespace

bool build_split(const HostLattice &L, int G, int nor, int threads, SplitPlan &out) {
    if (G < 1 || L.shape[0] < G || L.ncol < 1) return false;
    const int N = (int)L.N, K2 = L.K2, ncol = L.ncol;
    const int P = (nor + 1) * ncol;
    out = SplitPlan();
    out.G = G;
    out.nor = nor;
    out.P = P;
    out.ncol = ncol;
    out.K2 = K2;
    std::vector<int> colour(N), basis(N), part(N);
    for (int d = 0; d < N; ++d) {
        const int r = L.ref_of_dev[d];
        colour[d] = L.colour[r];
        basis[d] = L.basis_of(r);
        part[d] = (int)((long)first_coord(L, r) * G / L.shape[0]);
    }
    // one basis per colour class (the register kernel holds one coefficient
    // set per colour)
    std::vector<int> colbasis(ncol, -1);
    for (int d = 0; d < N; ++d) {
        int &cb = colbasis[colour[d]];
        if (cb < 0) cb = basis[d];
        else if (cb != basis[d]) return false;
    }
    out.colbasis = colbasis;
    auto nbr = [&](int d, int k) {
        const int j = L.nbr2[(size_t)L.ref_of_dev[d] * K2 + k];
        return j < 0 ? -1 : L.dev_of_ref[j];
    };
    // cost of the unsplit system: every phase updates the whole colour class
    std::vector<int> colsize(ncol, 0);
    for (int d = 0; d < N; ++d) ++colsize[colour[d]];
    for (int k = 0; k < P; ++k) out.work_full += waves(colsize[k % ncol]);

    std::vector<int> last(N), loc(N);
    for (int g = 0; g < G; ++g) {
        SplitPart sp;
        // last[d]: the last phase after which site d must hold its correct
        // value (P: at the window end); -2: not local
        std::fill(last.begin(), last.end(), -2);
        for (int d = 0; d < N; ++d)
            if (part[d] == g) last[d] = P;
        // the window-end energy reads the neighbours of the owned sites
        for (int d = 0; d < N; ++d)
            if (part[d] == g)
                for (int k = 0; k < K2; ++k) {
                    const int j = nbr(d, k);
                    if (j >= 0 && last[j] == -2) last[j] = P;
                }
        // backwards over the phases: a site updated at phase k (needed after
        // it and of the phase's colour) needs its neighbours after phase k-1
        std::vector<int> local;
        for (int d = 0; d < N; ++d)
            if (last[d] >= 0) local.push_back(d);
        for (int k = P - 1; k >= 0; --k) {
            const int c = k % ncol;
            const size_t n0 = local.size();
            for (size_t i = 0; i < n0; ++i) {
                const int d = local[i];
                if (colour[d] != c || last[d] < k) continue;
                for (int q = 0; q < K2; ++q) {
                    const int j = nbr(d, q);
                    if (j >= 0 && last[j] == -2) {
                        last[j] = k - 1;
                        local.push_back(j);
                    }
                }
            }
        }
        // local order: by colour; owned first (device order), then by last
        // phase needed, latest first
        std::sort(local.begin(), local.end(), [&](int a, int b) {
            if (colour[a] != colour[b]) return colour[a] < colour[b];
            const bool oa = part[a] == g, ob = part[b] == g;
            if (oa != ob) return oa;
            if (last[a] != last[b]) return last[a] > last[b];
            return a < b;
        });
        sp.n_loc = (int)local.size();
        sp.loc_dev = local;
        std::fill(loc.begin(), loc.end(), -1);
        for (int i = 0; i < sp.n_loc; ++i) loc[local[i]] = i;
        sp.col_start.assign(ncol + 1, 0);
        sp.own_cnt.assign(ncol, 0);
        for (int i = 0; i < sp.n_loc; ++i) {
            const int d = local[i];
            ++sp.col_start[colour[d] + 1];
            if (part[d] == g) {
                ++sp.own_cnt[colour[d]];
                ++sp.n_own;
            }
        }
        for (int c = 0; c < ncol; ++c) sp.col_start[c + 1] += sp.col_start[c];
        sp.phase_cnt.assign(P, 0);
        for (int k = 0; k < P; ++k) {
            const int c = k % ncol;
            int n = 0;
            for (int i = sp.col_start[c]; i < sp.col_start[c + 1]; ++i)
                if (last[local[i]] >= k) ++n;
            sp.phase_cnt[k] = n;
            sp.work += waves(n);
        }
        sp.nbl.assign((size_t)K2 * sp.n_loc, sp.n_loc);
        for (int i = 0; i < sp.n_loc; ++i)
            for (int k = 0; k < K2; ++k) {
                const int j = nbr(local[i], k);
                if (j >= 0 && loc[j] >= 0) sp.nbl[(size_t)k * sp.n_loc + i] = loc[j];
            }
        sp.colslot.assign(ncol + 1, 0);
        for (int c = 0; c < ncol; ++c) {
            sp.colslot[c] = (int)sp.slots.size();
            for (int f = sp.col_start[c]; f < sp.col_start[c + 1]; f += threads)
                sp.slots.push_back({f, sp.col_start[c + 1], colbasis[c], c});
        }
        sp.colslot[ncol] = (int)sp.slots.size();
        out.maxloc = std::max(out.maxloc, sp.n_loc);
        out.maxslots = std::max(out.maxslots, (int)sp.slots.size());
        out.work_max = std::max(out.work_max, sp.work);
        out.parts.push_back(std::move(sp));
    }
    // boundary sites: owned sites that some other part holds
    std::vector<char> held(N);
    for (int g = 0; g < G; ++g) {
        std::fill(held.begin(), held.end(), 0);
        for (int h = 0; h < G; ++h)
            if (h != g)
                for (int d : out.parts[h].loc_dev) held[d] = 1;
        SplitPart &sp = out.parts[g];
        for (int i = 0; i < sp.n_loc; ++i)
            if (part[sp.loc_dev[i]] == g && held[sp.loc_dev[i]]) sp.bnd.push_back(i);
        out.maxbnd = std::max(out.maxbnd, (int)sp.bnd.size());
    }
    // halo sites: where their owner publishes them, in that order
    std::vector<int> pos(N, -1);
    for (int g = 0; g < G; ++g)
        for (size_t k = 0; k < out.parts[g].bnd.size(); ++k)
            pos[out.parts[g].loc_dev[out.parts[g].bnd[k]]] = (int)k;
    for (int g = 0; g < G; ++g) {
        SplitPart &sp = out.parts[g];
        std::vector<std::pair<long, int>> h;
        for (int i = 0; i < sp.n_loc; ++i) {
            const int d = sp.loc_dev[i];
            if (part[d] != g) h.push_back({(long)part[d] * 3 * out.maxbnd + pos[d], i});
        }
        std::sort(h.begin(), h.end());
        for (auto &e : h) {
            sp.halo.push_back(e.second);
            sp.halo_src.push_back((int)e.first);
        }
        out.maxhalo = std::max(out.maxhalo, (int)sp.halo.size());
    }
    return true;
}

long check_split(const HostLattice &L, const SplitPlan &plan) {
    const int N = (int)L.N, K2 = L.K2, ncol = plan.ncol, P = plan.P;
    long bad = 0;
    std::vector<int> colour(N);
    for (int d = 0; d < N; ++d) colour[d] = L.colour[L.ref_of_dev[d]];
    for (int g = 0; g < plan.G; ++g) {
        const SplitPart &sp = plan.parts[g];
        const int n = sp.n_loc;
        std::vector<char> owned(n, 0);
        for (int c = 0; c < ncol; ++c)
            for (int i = sp.col_start[c]; i < sp.col_start[c] + sp.own_cnt[c]; ++i) owned[i] = 1;
        int nown = 0;
        for (int i = 0; i < n; ++i) nown += owned[i];
        if (nown != sp.n_own) ++bad;
        // a present neighbour missing from the local set reads as invalid
        auto nb_ok = [&](const std::vector<char> &v, int i, int k) {
            const int j = L.nbr2[(size_t)L.ref_of_dev[sp.loc_dev[i]] * K2 + k];
            const int l = sp.nbl[(size_t)k * n + i];
            if (j < 0) return l == n;
            return l < n && sp.loc_dev[l] == L.dev_of_ref[j] && v[l];
        };
        for (int start = 0; start < P; start += ncol) {
            std::vector<char> v(n, 1);
            for (int k = start; k < P; ++k) {
                const int c = k % ncol;
                std::vector<char> nv = v;
                for (int i = sp.col_start[c]; i < sp.col_start[c + 1]; ++i) {
                    if (i < sp.col_start[c] + sp.phase_cnt[k]) {
                        bool ok = true;
                        for (int q = 0; q < K2; ++q) ok = ok && nb_ok(v, i, q);
                        nv[i] = ok;
                    } else {
                        nv[i] = 0;  // its colour moved on without it
                    }
                }
                v.swap(nv);
                for (int i = 0; i < n; ++i)
                    if (owned[i] && !v[i]) ++bad;
            }
            for (int i = 0; i < n; ++i)
                if (owned[i])
                    for (int q = 0; q < K2; ++q)
                        if (!nb_ok(v, i, q)) ++bad;
        }
        // the halo and boundary lists agree with ownership
        for (int i : sp.halo)
            if (owned[i]) ++bad;
        if ((int)sp.halo.size() != n - sp.n_own) ++bad;
        for (int i : sp.bnd)
            if (!owned[i]) ++bad;
    }
    // every site owned by exactly one part, every halo site published by its owner
    std::vector<int> owner(N, -1);
    for (int g = 0; g < plan.G; ++g) {
        const SplitPart &sp = plan.parts[g];
        for (int c = 0; c < ncol; ++c)
            for (int i = sp.col_start[c]; i < sp.col_start[c] + sp.own_cnt[c]; ++i) {
                if (owner[sp.loc_dev[i]] >= 0) ++bad;
                owner[sp.loc_dev[i]] = g;
            }
    }
    for (int d = 0; d < N; ++d)
        if (owner[d] < 0) ++bad;
    // every halo site read from its owner's published entry for that site
    const int B = plan.maxbnd;
    for (int g = 0; g < plan.G; ++g) {
        const SplitPart &sp = plan.parts[g];
        if (sp.halo_src.size() != sp.halo.size()) ++bad;
        for (size_t j = 0; j < sp.halo.size() && j < sp.halo_src.size(); ++j) {
            const int src = sp.halo_src[j], h = src / (3 * B), k = src % (3 * B);
            if (h < 0 || h >= plan.G || h == g || k >= (int)plan.parts[h].bnd.size()) {
                ++bad;
                continue;
            }
            const SplitPart &o = plan.parts[h];
            if (o.loc_dev[o.bnd[k]] != sp.loc_dev[sp.halo[j]]) ++bad;
        }
    }
    return bad;
}

}  // namespace csmc
