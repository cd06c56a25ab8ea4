// lattice.cpp -- Lattice construction (lattice.jl:65-291) for libcsmc.
//
// The reference builds, for every site, one tuple entry per unit-cell term
// and finds each neighbour with an O(N) findfirst over the sorted site list
// (lattice.jl:196, 228-229, 273-275).  Here the position of a tuple in that
// sorted list is computed arithmetically (site order of lattice.jl:29-33:
// basis slowest, last lattice axis fastest), and the terms that touch a
// basis are compacted into "slots" so that a device site only walks the
// terms it is part of.  Couplings depend only on (basis, slot), never on the
// cell, so they are stored once per basis.
#include "lattice.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <queue>

namespace csmc {

namespace {

long site_of(const HostLattice &L, int b0, const int *x) {
    // b0 0-based basis, x 0-based cell coordinates (in range)
    long idx = b0;
    for (int d = 0; d < L.dim; ++d) idx = idx * L.shape[d] + x[d];
    return idx;
}

void cell_of(const HostLattice &L, long site, int *b0, int *x) {
    long r = site;
    for (int d = L.dim - 1; d >= 0; --d) {
        x[d] = (int)(r % L.shape[d]);
        r /= L.shape[d];
    }
    *b0 = (int)r;
}

// BC(index, offset) of lattice.jl:101-109 followed by the lookup; returns the
// 0-based neighbour or -1 when open boundaries put it outside the lattice.
long neighbour(const HostLattice &L, const int *x, const int *off, int bj0) {
    int y[3];
    for (int d = 0; d < L.dim; ++d) {
        long v = (long)x[d] + off[d];
        if (L.periodic) {
            v %= L.shape[d];
            if (v < 0) v += L.shape[d];
        } else if (v < 0 || v >= L.shape[d]) {
            return -1;
        }
        y[d] = (int)v;
    }
    if (bj0 < 0 || bj0 >= L.nb) return -1;
    return site_of(L, bj0, y);
}

// Source call of the field (or onsite) entry of each basis, restating the
// loop of lattice.jl:128-140 including its indexing quirk: for basis i that
// appears among the calls, column calls[i].basis receives call i.
int basis_sources(int nb, const std::vector<int> &calls, std::vector<int> &src, std::string &err,
                  const char *what) {
    std::vector<int> set(nb, 0);
    src.assign(nb, -1);
    int ncall = (int)calls.size();
    for (int i = 1; i <= nb; ++i) {
        bool in = std::find(calls.begin(), calls.end(), i) != calls.end();
        if (!in) {
            src[i - 1] = -1;
            set[i - 1] = 1;
            continue;
        }
        if (i > ncall) {
            err = std::string(what) + ": BoundsError (basis " + std::to_string(i) +
                  " has a call but the call list is shorter; lattice.jl:132,138)";
            return CSMC_EBOUNDS;
        }
        int col = calls[i - 1];
        if (col < 1 || col > nb) {
            err = std::string(what) + ": BoundsError (call " + std::to_string(i) +
                  " names basis " + std::to_string(col) + ")";
            return CSMC_EBOUNDS;
        }
        src[col - 1] = i - 1;
        set[col - 1] = 1;
    }
    for (int b = 0; b < nb; ++b)
        if (!set[b]) {
            err = std::string(what) + ": basis " + std::to_string(b + 1) +
                  " is left undef by the reference's per-basis assignment";
            return CSMC_EUNDEF;
        }
    return CSMC_OK;
}

// Role of basis b (1-based) in a term with bases tb[0..n-1]: the first
// match in order, as in the elseif chains of lattice.jl:184-193, 215-227,
// 251-272; -1 if the basis is not part of the term.
int role_of(int b, const int *tb, int n) {
    for (int r = 0; r < n; ++r)
        if (tb[r] == b) return r;
    return -1;
}

// Offsets of the partner sites seen from a site in role r; o[k] are the
// unit-cell offsets of term positions 1..n-1 (position 0 at the origin).
void role_offsets(int dim, int n, int r, const int *o, int *out) {
    // position p's offset relative to position r: o[p] - o[r] (o[0] = 0)
    // partner order after the role swap: position r takes the place of the
    // site, position 0 takes position r's place (lattice.jl:217-226,253-271).
    int pos[4];
    for (int p = 0; p < n; ++p) pos[p] = p;
    pos[0] = r;
    pos[r] = 0;
    for (int q = 1; q < n; ++q) {
        int p = pos[q];
        for (int d = 0; d < dim; ++d) {
            int op = (p == 0) ? 0 : o[(p - 1) * dim + d];
            int orr = (r == 0) ? 0 : o[(r - 1) * dim + d];
            out[(q - 1) * dim + d] = op - orr;
        }
    }
}

void load_unit_cell(const csmc_unit_cell *uc, HostLattice &L) {
    int D = uc->dim;
    L.dim = D;
    L.nb = uc->n_basis;
    L.avec.assign(uc->lattice_vectors, uc->lattice_vectors + D * D);
    L.basis.assign(uc->basis, uc->basis + (size_t)uc->n_basis * D);
    auto cpi = [](const int32_t *p, long n) { return std::vector<int>(p, p + (p ? n : 0)); };
    auto cpd = [](const double *p, long n) { return std::vector<double>(p, p + (p ? n : 0)); };
    L.fb = cpi(uc->field_b, uc->n_field);
    L.fh = cpd(uc->field_h, 3L * uc->n_field);
    L.ob = cpi(uc->onsite_b, uc->n_onsite);
    L.om = cpd(uc->onsite_m, 9L * uc->n_onsite);
    L.n2 = uc->n_bilinear;
    L.b2 = cpi(uc->bilinear_b, 2L * L.n2);
    L.m2 = cpd(uc->bilinear_m, 9L * L.n2);
    L.off2 = cpi(uc->bilinear_off, (long)D * L.n2);
    L.n3 = uc->n_cubic;
    L.b3 = cpi(uc->cubic_b, 3L * L.n3);
    L.t3 = cpd(uc->cubic_t, 27L * L.n3);
    L.off3 = cpi(uc->cubic_off, 2L * D * L.n3);
    L.n4 = uc->n_quartic;
    L.b4 = cpi(uc->quartic_b, 4L * L.n4);
    L.t4 = cpd(uc->quartic_t, 81L * L.n4);
    L.off4 = cpi(uc->quartic_off, 3L * D * L.n4);
}

// permuted coefficient for role r of an n-site tensor stored column-major:
// B[i_0 .. i_{n-1}] = A[i with positions 0 and r exchanged] (permutedims
// with the involution that swaps dimensions 1 and r+1).
void permute_tensor(int n, int r, const double *A, double *B) {
    int tot = 1;
    for (int k = 0; k < n; ++k) tot *= 3;
    for (int lin = 0; lin < tot; ++lin) {
        int idx[4], t = lin;
        for (int k = 0; k < n; ++k) {
            idx[k] = t % 3;
            t /= 3;
        }
        std::swap(idx[0], idx[r]);
        int src = 0, mul = 1;
        for (int k = 0; k < n; ++k) {
            src += idx[k] * mul;
            mul *= 3;
        }
        B[lin] = A[src];
    }
}

void build_coefficients(HostLattice &L) {
    int nb = L.nb;
    L.coef_onsite = 0;
    L.coef_field = (long)nb * 9;
    L.coef_c2 = L.coef_field + (long)nb * 3;
    L.coef_c3 = L.coef_c2 + (long)nb * L.K2 * 9;
    L.coef_c4 = L.coef_c3 + (long)nb * L.K3 * 27;
    L.coef_len = L.coef_c4 + (long)nb * L.K4 * 81;
    L.coef.assign(L.coef_len, 0.0);
    for (int b = 0; b < nb; ++b) {
        if (L.osrc[b] >= 0)
            std::memcpy(&L.coef[L.coef_onsite + b * 9], &L.om[L.osrc[b] * 9], 9 * sizeof(double));
        if (L.fsrc[b] >= 0)
            std::memcpy(&L.coef[L.coef_field + b * 3], &L.fh[L.fsrc[b] * 3], 3 * sizeof(double));
        for (int k = 0; k < L.K2; ++k) {
            int t = L.slot2[b * L.K2 + k];
            if (t < 0) continue;
            double *dst = &L.coef[L.coef_c2 + ((long)b * L.K2 + k) * 9];
            const double *M = &L.m2[t * 9];
            if (L.role2[b * L.K2 + k] == 0) {
                std::memcpy(dst, M, 9 * sizeof(double));
            } else {  // transposeJ (interaction_matrix.jl:27-31)
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) dst[a * 3 + c] = M[c * 3 + a];
            }
        }
        for (int k = 0; k < L.K3; ++k) {
            int t = L.slot3[b * L.K3 + k];
            if (t < 0) continue;
            permute_tensor(3, L.role3[b * L.K3 + k], &L.t3[t * 27],
                           &L.coef[L.coef_c3 + ((long)b * L.K3 + k) * 27]);
        }
        for (int k = 0; k < L.K4; ++k) {
            int t = L.slot4[b * L.K4 + k];
            if (t < 0) continue;
            permute_tensor(4, L.role4[b * L.K4 + k], &L.t4[t * 81],
                           &L.coef[L.coef_c4 + ((long)b * L.K4 + k) * 81]);
        }
    }
}

}  // namespace

int build_lattice(const csmc_unit_cell *uc, const int32_t *shape, int periodic, double S,
                  HostLattice &L, std::string &err) {
    if (!uc || !shape) {
        err = "null unit cell or shape";
        return CSMC_EINVAL;
    }
    if (periodic != 0 && periodic != 1) {
        // the reference only evaluates BC() while walking terms
        // (lattice.jl:101-109, 195), so an unknown option fails only then
        if (uc->n_bilinear + uc->n_cubic + uc->n_quartic > 0) {
            err = "Invalid boundary condition option";
            return CSMC_EBC;
        }
        periodic = 1;
    }
    if (uc->dim < 1 || uc->dim > 3 || uc->n_basis < 1) {
        err = "unit cell needs 1 <= dim <= 3 and at least one basis site";
        return CSMC_EINVAL;
    }
    L = HostLattice();
    load_unit_cell(uc, L);
    L.periodic = periodic;
    L.S = S;
    L.ncell = 1;
    for (int d = 0; d < L.dim; ++d) {
        if (shape[d] < 1) {
            err = "lattice shape entries must be >= 1";
            return CSMC_EINVAL;
        }
        L.shape[d] = shape[d];
        L.ncell *= shape[d];
    }
    L.N = L.ncell * L.nb;
    if (L.N > (1L << 30)) {
        err = "lattice too large for int32 device indices";
        return CSMC_EINVAL;
    }
    for (int t = 0; t < L.n2; ++t)
        for (int q = 0; q < 2; ++q)
            if (L.b2[2 * t + q] < 1 || L.b2[2 * t + q] > L.nb) {
                err = "bilinear term " + std::to_string(t + 1) + " names a basis out of range";
                return CSMC_EBOUNDS;
            }
    for (int t = 0; t < L.n3; ++t)
        for (int q = 0; q < 3; ++q)
            if (L.b3[3 * t + q] < 1 || L.b3[3 * t + q] > L.nb) {
                err = "cubic term " + std::to_string(t + 1) + " names a basis out of range";
                return CSMC_EBOUNDS;
            }
    for (int t = 0; t < L.n4; ++t)
        for (int q = 0; q < 4; ++q)
            if (L.b4[4 * t + q] < 1 || L.b4[4 * t + q] > L.nb) {
                err = "quartic term " + std::to_string(t + 1) + " names a basis out of range";
                return CSMC_EBOUNDS;
            }

    int rc = basis_sources(L.nb, L.fb, L.fsrc, err, "Zeeman field");
    if (rc) return rc;
    rc = basis_sources(L.nb, L.ob, L.osrc, err, "onsite");
    if (rc) return rc;

    // compaction: active terms per basis, in term order
    auto compact = [&](int nterm, int width, const std::vector<int> &tb, std::vector<int> &slot,
                       std::vector<int> &role, int &K) {
        std::vector<std::vector<std::pair<int, int>>> act(L.nb);
        for (int b = 1; b <= L.nb; ++b)
            for (int t = 0; t < nterm; ++t) {
                int r = role_of(b, &tb[(size_t)t * width], width);
                if (r >= 0) act[b - 1].push_back({t, r});
            }
        K = 0;
        for (auto &a : act) K = std::max<int>(K, (int)a.size());
        slot.assign((size_t)L.nb * K, -1);
        role.assign((size_t)L.nb * K, 0);
        for (int b = 0; b < L.nb; ++b)
            for (size_t k = 0; k < act[b].size(); ++k) {
                slot[b * K + k] = act[b][k].first;
                role[b * K + k] = act[b][k].second;
            }
    };
    compact(L.n2, 2, L.b2, L.slot2, L.role2, L.K2);
    compact(L.n3, 3, L.b3, L.slot3, L.role3, L.K3);
    compact(L.n4, 4, L.b4, L.slot4, L.role4, L.K4);

    // neighbours
    L.nbr2.assign((size_t)L.N * L.K2, -1);
    L.nbr3.assign((size_t)L.N * L.K3 * 2, -1);
    L.nbr4.assign((size_t)L.N * L.K4 * 3, -1);
    const int D = L.dim;
    for (long s = 0; s < L.N; ++s) {
        int b0, x[3];
        cell_of(L, s, &b0, x);
        for (int k = 0; k < L.K2; ++k) {
            int t = L.slot2[b0 * L.K2 + k];
            if (t < 0) continue;
            int r = L.role2[b0 * L.K2 + k];
            const int *tb = &L.b2[2 * t];
            int off[3], bj;
            if (tb[0] == tb[1] || r == 0) {
                bj = (tb[0] == tb[1]) ? b0 + 1 : tb[1];
                for (int d = 0; d < D; ++d) off[d] = L.off2[t * D + d];
            } else {
                bj = tb[0];
                for (int d = 0; d < D; ++d) off[d] = -L.off2[t * D + d];
            }
            L.nbr2[s * L.K2 + k] = (int)neighbour(L, x, off, bj - 1);
        }
        for (int k = 0; k < L.K3; ++k) {
            int t = L.slot3[b0 * L.K3 + k];
            if (t < 0) continue;
            int r = L.role3[b0 * L.K3 + k];
            const int *tb = &L.b3[3 * t];
            int roff[6];
            role_offsets(D, 3, r, &L.off3[(size_t)t * 2 * D], roff);
            int pb[3] = {tb[0], tb[1], tb[2]};
            std::swap(pb[0], pb[r]);
            long j = neighbour(L, x, &roff[0], pb[1] - 1);
            long kk = neighbour(L, x, &roff[D], pb[2] - 1);
            if (j >= 0 && kk >= 0) {
                L.nbr3[(s * L.K3 + k) * 2] = (int)j;
                L.nbr3[(s * L.K3 + k) * 2 + 1] = (int)kk;
            }
        }
        for (int k = 0; k < L.K4; ++k) {
            int t = L.slot4[b0 * L.K4 + k];
            if (t < 0) continue;
            int r = L.role4[b0 * L.K4 + k];
            const int *tb = &L.b4[4 * t];
            int roff[9];
            role_offsets(D, 4, r, &L.off4[(size_t)t * 3 * D], roff);
            int pb[4] = {tb[0], tb[1], tb[2], tb[3]};
            std::swap(pb[0], pb[r]);
            long j = neighbour(L, x, &roff[0], pb[1] - 1);
            long kk = neighbour(L, x, &roff[D], pb[2] - 1);
            long l = neighbour(L, x, &roff[2 * D], pb[3] - 1);
            if (j >= 0 && kk >= 0 && l >= 0) {
                L.nbr4[(s * L.K4 + k) * 3] = (int)j;
                L.nbr4[(s * L.K4 + k) * 3 + 1] = (int)kk;
                L.nbr4[(s * L.K4 + k) * 3 + 2] = (int)l;
            }
        }
    }
    // self references (a term that names the site itself, e.g. L=1 or the
    // repeated-site biquadratic term)
    L.has_self = 0;
    for (long s = 0; s < L.N && !L.has_self; ++s) {
        for (int k = 0; k < L.K2; ++k) L.has_self |= (L.nbr2[s * L.K2 + k] == s);
        for (int k = 0; k < L.K3 * 2; ++k) L.has_self |= (L.nbr3[s * L.K3 * 2 + k] == s);
        for (int k = 0; k < L.K4 * 3; ++k) L.has_self |= (L.nbr4[s * L.K4 * 3 + k] == s);
    }
    build_coefficients(L);
    L.grad_len = 3 * (int)L.fb.size() + 9 * (int)L.ob.size() + 9 * L.n2 + 27 * L.n3 + 81 * L.n4;
    colour_lattice(L);
    if (!colouring_valid(L, err)) return CSMC_ESTATE;
    return CSMC_OK;
}

int update_couplings(const csmc_unit_cell *uc, HostLattice &L, std::string &err) {
    bool same = uc->dim == L.dim && uc->n_basis == L.nb && uc->n_field == (int)L.fb.size() &&
                uc->n_onsite == (int)L.ob.size() && uc->n_bilinear == L.n2 &&
                uc->n_cubic == L.n3 && uc->n_quartic == L.n4;
    if (same) {
        for (int i = 0; i < uc->n_field; ++i) same &= uc->field_b[i] == L.fb[i];
        for (int i = 0; i < uc->n_onsite; ++i) same &= uc->onsite_b[i] == L.ob[i];
        for (int i = 0; i < 2 * L.n2; ++i) same &= uc->bilinear_b[i] == L.b2[i];
        for (int i = 0; i < L.dim * L.n2; ++i) same &= uc->bilinear_off[i] == L.off2[i];
        for (int i = 0; i < 3 * L.n3; ++i) same &= uc->cubic_b[i] == L.b3[i];
        for (int i = 0; i < 2 * L.dim * L.n3; ++i) same &= uc->cubic_off[i] == L.off3[i];
        for (int i = 0; i < 4 * L.n4; ++i) same &= uc->quartic_b[i] == L.b4[i];
        for (int i = 0; i < 3 * L.dim * L.n4; ++i) same &= uc->quartic_off[i] == L.off4[i];
    }
    if (!same) {
        err = "update_couplings: unit cell term structure differs from the lattice's";
        return CSMC_EINVAL;
    }
    L.fh.assign(uc->field_h, uc->field_h + 3 * L.fb.size());
    L.om.assign(uc->onsite_m, uc->onsite_m + 9 * L.ob.size());
    L.m2.assign(uc->bilinear_m, uc->bilinear_m + 9L * L.n2);
    L.t3.assign(uc->cubic_t, uc->cubic_t + 27L * L.n3);
    L.t4.assign(uc->quartic_t, uc->quartic_t + 81L * L.n4);
    build_coefficients(L);
    return CSMC_OK;
}

void export_tables(const HostLattice &L, double *pos, double *field, double *onsite,
                   int64_t *bsite, double *bmat, int64_t *csite, double *ctens, int64_t *qsite,
                   double *qtens) {
    const int D = L.dim;
    for (long s = 0; s < L.N; ++s) {
        int b0, x[3];
        cell_of(L, s, &b0, x);
        if (pos)  // compute_site_positions (lattice.jl:38-51)
            for (int k = 0; k < D; ++k) {
                double acc = 0.0;
                for (int d = 0; d < D; ++d) acc += (double)x[d] * L.avec[d * D + k];
                pos[s * D + k] = acc + L.basis[b0 * D + k];
            }
        if (field)
            std::memcpy(&field[s * 3], &L.coef[L.coef_field + b0 * 3], 3 * sizeof(double));
        if (onsite)
            std::memcpy(&onsite[s * 9], &L.coef[L.coef_onsite + b0 * 9], 9 * sizeof(double));
        // expand compacted slots back to one entry per unit-cell term
        if (bsite || bmat) {
            for (int t = 0; t < L.n2; ++t) {
                if (bsite) bsite[s * L.n2 + t] = 0;
                if (bmat) std::memset(&bmat[(s * L.n2 + t) * 9], 0, 9 * sizeof(double));
            }
            for (int k = 0; k < L.K2; ++k) {
                int t = L.slot2[b0 * L.K2 + k];
                int j = L.nbr2[s * L.K2 + k];
                if (t < 0 || j < 0) continue;
                if (bsite) bsite[s * L.n2 + t] = j + 1;
                if (bmat)
                    std::memcpy(&bmat[(s * L.n2 + t) * 9], &L.coef[L.coef_c2 + ((long)b0 * L.K2 + k) * 9],
                                9 * sizeof(double));
            }
        }
        if (csite || ctens) {
            for (int t = 0; t < L.n3; ++t) {
                if (csite) csite[(s * L.n3 + t) * 2] = csite[(s * L.n3 + t) * 2 + 1] = 0;
                if (ctens) std::memset(&ctens[(s * L.n3 + t) * 27], 0, 27 * sizeof(double));
            }
            for (int k = 0; k < L.K3; ++k) {
                int t = L.slot3[b0 * L.K3 + k];
                int j = L.nbr3[(s * L.K3 + k) * 2];
                if (t < 0 || j < 0) continue;
                if (csite) {
                    csite[(s * L.n3 + t) * 2] = j + 1;
                    csite[(s * L.n3 + t) * 2 + 1] = L.nbr3[(s * L.K3 + k) * 2 + 1] + 1;
                }
                if (ctens)
                    std::memcpy(&ctens[(s * L.n3 + t) * 27],
                                &L.coef[L.coef_c3 + ((long)b0 * L.K3 + k) * 27], 27 * sizeof(double));
            }
        }
        if (qsite || qtens) {
            for (int t = 0; t < L.n4; ++t) {
                if (qsite) qsite[(s * L.n4 + t) * 3] = qsite[(s * L.n4 + t) * 3 + 1] = qsite[(s * L.n4 + t) * 3 + 2] = 0;
                if (qtens) std::memset(&qtens[(s * L.n4 + t) * 81], 0, 81 * sizeof(double));
            }
            for (int k = 0; k < L.K4; ++k) {
                int t = L.slot4[b0 * L.K4 + k];
                int j = L.nbr4[(s * L.K4 + k) * 3];
                if (t < 0 || j < 0) continue;
                if (qsite)
                    for (int q = 0; q < 3; ++q) qsite[(s * L.n4 + t) * 3 + q] = L.nbr4[(s * L.K4 + k) * 3 + q] + 1;
                if (qtens)
                    std::memcpy(&qtens[(s * L.n4 + t) * 81],
                                &L.coef[L.coef_c4 + ((long)b0 * L.K4 + k) * 81], 81 * sizeof(double));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Colouring.  Two sites conflict when one appears in a term of the other;
// a colour class is then an independent set, and a parallel pass over a
// class gives exactly the reference's sequential loop restricted to it.
// ---------------------------------------------------------------------------
namespace {

void conflict_graph(const HostLattice &L, std::vector<long> &off, std::vector<int> &adj) {
    std::vector<std::vector<int>> nb(L.N);
    auto add = [&](long i, long j) {
        if (j < 0 || j == i) return;
        nb[i].push_back((int)j);
        nb[j].push_back((int)i);
    };
    for (long s = 0; s < L.N; ++s) {
        for (int k = 0; k < L.K2; ++k) add(s, L.nbr2[s * L.K2 + k]);
        for (int k = 0; k < L.K3 * 2; ++k) add(s, L.nbr3[s * L.K3 * 2 + k]);
        for (int k = 0; k < L.K4 * 3; ++k) add(s, L.nbr4[s * L.K4 * 3 + k]);
    }
    off.assign(L.N + 1, 0);
    for (long s = 0; s < L.N; ++s) {
        auto &v = nb[s];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        off[s + 1] = off[s] + (long)v.size();
    }
    adj.resize(off[L.N]);
    for (long s = 0; s < L.N; ++s) std::copy(nb[s].begin(), nb[s].end(), adj.begin() + off[s]);
}

// BFS two-colouring; false if an odd cycle exists
bool two_colour(long N, const std::vector<long> &off, const std::vector<int> &adj,
                std::vector<int> &col) {
    col.assign(N, -1);
    std::vector<int> q;
    for (long s = 0; s < N; ++s) {
        if (col[s] >= 0) continue;
        col[s] = 0;
        q.assign(1, (int)s);
        for (size_t h = 0; h < q.size(); ++h) {
            int u = q[h];
            for (long e = off[u]; e < off[u + 1]; ++e) {
                int v = adj[e];
                if (col[v] < 0) {
                    col[v] = 1 - col[u];
                    q.push_back(v);
                } else if (col[v] == col[u]) {
                    return false;
                }
            }
        }
    }
    return true;
}

// DSatur (Brelaz 1979): colour the vertex with the most distinct neighbour
// colours next, ties by degree.
int dsatur(long N, const std::vector<long> &off, const std::vector<int> &adj,
           std::vector<int> &col) {
    col.assign(N, -1);
    std::vector<std::vector<char>> seen(N);
    std::vector<int> sat(N, 0);
    using Key = std::tuple<int, int, int>;  // sat, degree, -index
    std::priority_queue<Key> pq;
    for (long s = 0; s < N; ++s) pq.push({0, (int)(off[s + 1] - off[s]), -(int)s});
    int ncol = 0;
    while (!pq.empty()) {
        auto [st, dg, ni] = pq.top();
        pq.pop();
        int u = -ni;
        if (col[u] >= 0 || st != sat[u]) continue;
        int c = 0;
        while (c < (int)seen[u].size() && seen[u][c]) ++c;
        col[u] = c;
        ncol = std::max(ncol, c + 1);
        for (long e = off[u]; e < off[u + 1]; ++e) {
            int v = adj[e];
            if (col[v] >= 0) continue;
            if ((int)seen[v].size() <= c) seen[v].resize(c + 1, 0);
            if (!seen[v][c]) {
                seen[v][c] = 1;
                ++sat[v];
                pq.push({sat[v], (int)(off[v + 1] - off[v]), -v});
            }
        }
    }
    return ncol;
}

// greedy in reference order
int greedy(long N, const std::vector<long> &off, const std::vector<int> &adj,
           std::vector<int> &col) {
    col.assign(N, -1);
    int ncol = 0;
    std::vector<int> mark;
    for (long u = 0; u < N; ++u) {
        mark.assign(ncol + 1, 0);
        for (long e = off[u]; e < off[u + 1]; ++e)
            if (col[adj[e]] >= 0 && col[adj[e]] <= ncol) mark[col[adj[e]]] = 1;
        int c = 0;
        while (mark[c]) ++c;
        col[u] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// Lattice-periodic colouring colour(b, x) = (c.x + e_b) mod p: every class
// is a union of whole sublattice strata, so classes are balanced and each
// (colour, basis) run has one coefficient set.  A pattern is valid when no
// term partner (basis bj at offset o from basis b) gets the same colour,
// i.e. (c.o + e_bj - e_b) mod p != 0, and, for periodic boundaries, when
// it is well defined on the torus (p | c_d L_d).  Returns p, or 0.
int pattern_colour(const HostLattice &L, std::vector<int> &col, ColourPattern &pat) {
    const int D = L.dim, nb = L.nb;
    struct Con {
        int b, bj, o[3];
    };
    std::vector<Con> cons;
    auto add = [&](int b, int bj, const int *o) {
        Con c{b, bj, {0, 0, 0}};
        bool zero = (b == bj);
        for (int d = 0; d < D; ++d) {
            c.o[d] = o[d];
            zero &= (o[d] == 0);
        }
        if (!zero) cons.push_back(c);  // the site itself is not a conflict
    };
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < L.K2; ++k) {
            int t = L.slot2[b * L.K2 + k];
            if (t < 0) continue;
            const int *tb = &L.b2[2 * t];
            int o[3];
            bool fwd = tb[0] == tb[1] || L.role2[b * L.K2 + k] == 0;
            for (int d = 0; d < D; ++d) o[d] = (fwd ? 1 : -1) * L.off2[t * D + d];
            add(b, fwd ? (tb[0] == tb[1] ? b : tb[1] - 1) : tb[0] - 1, o);
        }
        for (int k = 0; k < L.K3; ++k) {
            int t = L.slot3[b * L.K3 + k];
            if (t < 0) continue;
            int r = L.role3[b * L.K3 + k], roff[6];
            role_offsets(D, 3, r, &L.off3[(size_t)t * 2 * D], roff);
            int pb[3] = {L.b3[3 * t], L.b3[3 * t + 1], L.b3[3 * t + 2]};
            std::swap(pb[0], pb[r]);
            for (int q = 1; q < 3; ++q) add(b, pb[q] - 1, &roff[(q - 1) * D]);
        }
        for (int k = 0; k < L.K4; ++k) {
            int t = L.slot4[b * L.K4 + k];
            if (t < 0) continue;
            int r = L.role4[b * L.K4 + k], roff[9];
            role_offsets(D, 4, r, &L.off4[(size_t)t * 3 * D], roff);
            int pb[4] = {L.b4[4 * t], L.b4[4 * t + 1], L.b4[4 * t + 2], L.b4[4 * t + 3]};
            std::swap(pb[0], pb[r]);
            for (int q = 1; q < 4; ++q) add(b, pb[q] - 1, &roff[(q - 1) * D]);
        }
    }
    for (int p = 1; p <= 8; ++p) {
        long ncand = 1;
        for (int k = 0; k < D + nb - 1; ++k) ncand *= p;
        if (ncand > 400000) break;
        std::vector<int> c(D), e(nb);
        for (long idx = 0; idx < ncand; ++idx) {
            long r = idx;
            for (int d = 0; d < D; ++d) {
                c[d] = (int)(r % p);
                r /= p;
            }
            e[0] = 0;
            for (int b = 1; b < nb; ++b) {
                e[b] = (int)(r % p);
                r /= p;
            }
            bool ok = true;
            if (L.periodic)
                for (int d = 0; d < D && ok; ++d) ok = ((long)c[d] * L.shape[d]) % p == 0;
            for (size_t q = 0; q < cons.size() && ok; ++q) {
                long v = e[cons[q].bj] - e[cons[q].b];
                for (int d = 0; d < D; ++d) v += (long)c[d] * cons[q].o[d];
                ok = ((v % p) + p) % p != 0;
            }
            if (!ok) continue;
            col.assign(L.N, 0);
            for (long s = 0; s < L.N; ++s) {
                int b0, x[3];
                cell_of(L, s, &b0, x);
                long v = e[b0];
                for (int d = 0; d < D; ++d) v += (long)c[d] * x[d];
                col[s] = (int)(v % p);
            }
            // compact away colours that no site uses (p larger than needed)
            std::vector<int> used(p, 0), remap(p, -1);
            for (long s = 0; s < L.N; ++s) used[col[s]] = 1;
            int n = 0;
            for (int q = 0; q < p; ++q)
                if (used[q]) remap[q] = n++;
            for (long s = 0; s < L.N; ++s) col[s] = remap[col[s]];
            pat.p = p;
            for (int d = 0; d < 3; ++d) pat.c[d] = d < D ? c[d] : 0;
            pat.e = e;
            pat.remap = remap;
            return n;
        }
    }
    return 0;
}

}  // namespace

void colour_lattice(HostLattice &L) {
    std::vector<long> off;
    std::vector<int> adj;
    conflict_graph(L, off, adj);
    std::vector<int> best;
    int nbest;
    std::vector<int> pat;
    ColourPattern cp;
    int npat = pattern_colour(L, pat, cp);
    L.pattern = ColourPattern{};
    if (off[L.N] == 0) {
        best.assign(L.N, 0);
        nbest = 1;
    } else {
        std::vector<int> c2;
        if (two_colour(L.N, off, adj, best)) {
            nbest = 2;
        } else {
            nbest = dsatur(L.N, off, adj, best);
            int ng = greedy(L.N, off, adj, c2);
            if (ng < nbest) {
                nbest = ng;
                best.swap(c2);
            }
        }
        if (npat > 0 && npat <= nbest) {
            best.swap(pat);
            nbest = npat;
            L.pattern = cp;
        }
    }
    L.colour = best;
    L.ncol = nbest;
    // device order: (colour, basis, reference index)
    L.ref_of_dev.resize(L.N);
    std::iota(L.ref_of_dev.begin(), L.ref_of_dev.end(), 0);
    std::stable_sort(L.ref_of_dev.begin(), L.ref_of_dev.end(), [&](int a, int b) {
        if (L.colour[a] != L.colour[b]) return L.colour[a] < L.colour[b];
        return L.basis_of(a) < L.basis_of(b);
    });
    L.dev_of_ref.resize(L.N);
    for (long d = 0; d < L.N; ++d) L.dev_of_ref[L.ref_of_dev[d]] = (int)d;
    L.runs.clear();
    L.colrun.assign(L.ncol + 1, 0);
    long d = 0;
    for (int c = 0; c < L.ncol; ++c) {
        L.colrun[c] = (int)L.runs.size();
        while (d < L.N && L.colour[L.ref_of_dev[d]] == c) {
            int b = L.basis_of(L.ref_of_dev[d]);
            long e = d;
            while (e < L.N && L.colour[L.ref_of_dev[e]] == c && L.basis_of(L.ref_of_dev[e]) == b) ++e;
            L.runs.push_back({(int)d, (int)e, b, c});
            d = e;
        }
    }
    L.colrun[L.ncol] = (int)L.runs.size();
}

bool colouring_valid(const HostLattice &L, std::string &err) {
    auto chk = [&](long s, int j) {
        if (j < 0 || j == s) return true;
        if (L.colour[j] == L.colour[s]) {
            err = "colouring conflict between sites " + std::to_string(s + 1) + " and " +
                  std::to_string(j + 1);
            return false;
        }
        return true;
    };
    for (long s = 0; s < L.N; ++s) {
        for (int k = 0; k < L.K2; ++k)
            if (!chk(s, L.nbr2[s * L.K2 + k])) return false;
        // a site's update reads exactly the sites of its own entries
        // (get_local_field / energy walk only the site's tuple), so the
        // owner-member pairs are the only dependencies
        for (int k = 0; k < L.K3 * 2; ++k)
            if (!chk(s, L.nbr3[s * L.K3 * 2 + k])) return false;
        for (int k = 0; k < L.K4 * 3; ++k)
            if (!chk(s, L.nbr4[s * L.K4 * 3 + k])) return false;
    }
    return true;
}

}  // namespace csmc
