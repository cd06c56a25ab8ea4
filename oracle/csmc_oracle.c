/*
 * csmc_oracle.c -- CPU FP64 oracle for the ClassicalSpinMC.jl sweep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is imported by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (libcsmc.so) never links or calls it.
 *
 * It is a plain-C restatement of the reference algorithm (Julia package
 * ClassicalSpinMC.jl).  Every function cites the reference file:line it follows.
 * Conventions kept from the reference:
 *   - per-site tables use 1-based site indices, 0 = absent (lattice.jl:16-21);
 *   - sites are ordered lexicographically by (basis, x1, .., xD)
 *     (lattice.jl:29-33);
 *   - InteractionMatrix is row-major m11..m33 (interaction_matrix.jl:2-12);
 *   - cubic/quartic tensors are Julia column-major: C[a,b,c] at
 *     a + 3b + 9c (0-based), R[a,b,c,d] at a + 3b + 9c + 27d;
 *   - spins are 3xN column-major, i.e. s[3*i + comp] (lattice.jl:6).
 *
 * Parity pins (what each function is checked against in tests/):
 *   orc_build_tables   KATs test/latticetests.jl:17-18,30; literal findfirst
 *                      restatement (search=1) vs arithmetic index (search=0);
 *                      brute-force geometric bond enumeration on tiny lattices.
 *   orc_local_field /  Zeeman KAT E=-1, H=(-1,0,0); FM E/N=-2; closed forms
 *   orc_site_energy /  (FM ground state -2S^2|J|-hS, chirality of an
 *   orc_total_energy   orthonormal triad = +-1, biquadratic B(s.s')^2);
 *                      central finite differences of orc_total_energy.
 *                      Cubic/quartic role permutations (lattice.jl:209-286):
 *                      on a pyrochlore cell with generic non-symmetric
 *                      tensors on distinct-basis faces and tetrahedra, the
 *                      total energy equals an independent term-by-term sum
 *                      and the local field equals its central finite
 *                      difference; the first-matching-role rule for terms
 *                      that repeat a basis is checked against a hand-built
 *                      field and the 2/3 (cubic) and 1/2 (quartic) energy
 *                      weights it implies.
 *   orc_overrelax_*    involution, |s| and s.H preserved, E conserved
 *                      (bilinear + Zeeman).
 *   orc_philox4x32_10  Random123 known-answer vectors.
 *   orc_binner_*       direct blocking analysis in numpy.
 *   orc_metropolis_*   single spin in a field: <s.h> = S L(beta h S) (Langevin),
 *                      with the uniform proposal and with the cone proposal
 *                      (metropolis.jl:84-87) at two cone widths.
 *   orc_metropolis_    weight 0 gives the plain cone chain draw for draw; a
 *   constraint_coloured  free spin with c = s_z - m0 and a strong weight is
 *                      driven to <s_z> = m0; lambda follows its recurrence
 *                      on a hand-worked two-proposal case.
 *   orc_pt_exchange_*  the acceptance formula monte_carlo.jl:326-331 checked
 *                      against detailed balance on a two-level toy.
 *   orc_equal_time_    closed forms: a ferromagnet along z gives S^zz(0) = N S^2
 *   correlations       and 0 at the other allowed momenta; a Neel state on the
 *                      square lattice gives S^zz(pi,pi) = N S^2; Parseval over
 *                      the ncell allowed momenta of a Bravais lattice,
 *                      sum_k S^aa(k) = sum_i s_ia^2; S^ab = S^ba and S(k) =
 *                      S(-k); a direct numpy FFT of the spins on the grid.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_EBOUNDS 1   /* Julia would raise BoundsError                    */
#define ORC_EUNDEF 2    /* Julia would read an undef entry                  */
#define ORC_EBC 3       /* error("Invalid boundary condition option")      */
#define ORC_EARG 4

/* ------------------------------------------------------------------------ */
/* Site ordering (lattice.jl:29-33): sort!(product(1:nb, 1:L1, .., 1:LD)).   */
/* Lexicographic tuple order => basis slowest, last lattice axis fastest.   */
/* ------------------------------------------------------------------------ */
static long orc_index_of(int dim, const int *shape, int b, const int *x)
{
    /* b and x are 1-based; returns 1-based site index */
    long idx = b - 1;
    for (int d = 0; d < dim; ++d) idx = idx * shape[d] + (x[d] - 1);
    return idx + 1;
}

static void orc_tuple_of(int dim, const int *shape, long site, int *b, int *x)
{
    /* inverse of orc_index_of; site 1-based */
    long r = site - 1;
    for (int d = dim - 1; d >= 0; --d) {
        x[d] = (int)(r % shape[d]) + 1;
        r /= shape[d];
    }
    *b = (int)r + 1;
}

/* BC(index, offset) of lattice.jl:101-109: periodic wraps with a floored mod,
 * open returns the raw coordinate (which findfirst then fails to match). */
static void orc_bc(int dim, const int *shape, int periodic, const int *x,
                   const int *off, int *out)
{
    for (int d = 0; d < dim; ++d) {
        long v = (long)x[d] + off[d];
        if (periodic) {
            long m = (v - 1) % shape[d];
            if (m < 0) m += shape[d];
            out[d] = (int)m + 1;
        } else {
            out[d] = (int)v;
        }
    }
}

/* Literal findfirst over the sorted index list (lattice.jl:196): scan every
 * site's tuple.  O(N) per lookup, as in the reference. */
static long orc_find_literal(int dim, const int *shape, long N, int b, const int *x)
{
    int tb, tx[3];
    for (long s = 1; s <= N; ++s) {
        orc_tuple_of(dim, shape, s, &tb, tx);
        if (tb != b) continue;
        int eq = 1;
        for (int d = 0; d < dim; ++d) eq &= (tx[d] == x[d]);
        if (eq) return s;
    }
    return 0;
}

/* Arithmetic lookup: the position the sorted list would hold (0 if the tuple
 * is outside the lattice, which is what findfirst returns for open BC). */
static long orc_find_arith(int dim, const int *shape, int nb, int b, const int *x)
{
    if (b < 1 || b > nb) return 0;
    for (int d = 0; d < dim; ++d)
        if (x[d] < 1 || x[d] > shape[d]) return 0;
    return orc_index_of(dim, shape, b, x);
}

static long orc_find(int search, int dim, const int *shape, int nb, long N, int b,
                     const int *x)
{
    return search ? orc_find_literal(dim, shape, N, b, x)
                  : orc_find_arith(dim, shape, nb, b, x);
}

/* ------------------------------------------------------------------------ */
/* Lattice construction (lattice.jl:65-291).                                 */
/*                                                                           */
/* Inputs (unit cell, unit_cell.jl:3-18):                                    */
/*   avec[dim*dim]   lattice vector d at avec[d*dim .. d*dim+dim-1]         */
/*   basis[nb*dim]                                                           */
/*   field:   nf entries (f_b[i], f_h[3i..3i+2])                             */
/*   onsite:  no entries (o_b[i], o_m[9i..])                                 */
/*   bilinear: n2 entries (b2[2i], b2[2i+1]), m2[9i..], off2[dim*i..]        */
/*   cubic:    n3 entries b3[3i..], t3[27i..], off3[2*dim*i..] (o2 then o3)  */
/*   quartic:  n4 entries b4[4i..], t4[81i..], off4[3*dim*i..]               */
/* Outputs (Lattice fields, lattice.jl:10-22), caller-allocated:             */
/*   pos[N*dim], field[N*3], onsite[N*9],                                    */
/*   bsite[N*n2], bmat[N*n2*9], csite[N*n3*2], ctens[N*n3*27],               */
/*   qsite[N*n4*3], qtens[N*n4*81]                                           */
/* ------------------------------------------------------------------------ */
int orc_build_tables(int dim, const int *shape, int periodic, int search,
                     const double *avec, int nb, const double *basis,
                     int nf, const int *f_b, const double *f_h,
                     int no, const int *o_b, const double *o_m,
                     int n2, const int *b2, const double *m2, const int *off2,
                     int n3, const int *b3, const double *t3, const int *off3,
                     int n4, const int *b4, const double *t4, const int *off4,
                     double *pos, double *field, double *onsite,
                     long *bsite, double *bmat, long *csite, double *ctens,
                     long *qsite, double *qtens)
{
    if (dim < 1 || dim > 3 || nb < 1) return ORC_EARG;
    if (periodic != 0 && periodic != 1) return ORC_EBC;
    long ncell = 1;
    for (int d = 0; d < dim; ++d) {
        if (shape[d] < 1) return ORC_EARG;
        ncell *= shape[d];
    }
    long N = ncell * nb;

    /* compute_site_positions (lattice.jl:38-51) */
    for (long s = 1; s <= N; ++s) {
        int b, x[3];
        orc_tuple_of(dim, shape, s, &b, x);
        for (int k = 0; k < dim; ++k) {
            double acc = 0.0;
            for (int d = 0; d < dim; ++d) acc += (double)(x[d] - 1) * avec[d * dim + k];
            pos[(s - 1) * dim + k] = acc + basis[(b - 1) * dim + k];
        }
    }

    /* Zeeman field and onsite per basis (lattice.jl:120-140), quirk kept:
     * for basis i with a call, column f_indices[i] (indexed by the CALL
     * number i) receives f_[i].  BoundsError when i exceeds the call count;
     * an entry never written is Julia's undef. */
    double *fcol = (double *)calloc((size_t)nb * 3, sizeof(double));
    double *ocol = (double *)calloc((size_t)nb * 9, sizeof(double));
    int *fset = (int *)calloc((size_t)nb, sizeof(int));
    int *oset = (int *)calloc((size_t)nb, sizeof(int));
    int rc = ORC_OK;
    for (int i = 1; i <= nb && rc == ORC_OK; ++i) {
        int in_f = 0, in_o = 0;
        for (int k = 0; k < nf; ++k) in_f |= (f_b[k] == i);
        for (int k = 0; k < no; ++k) in_o |= (o_b[k] == i);
        if (!in_f) {
            fcol[(i - 1) * 3 + 0] = fcol[(i - 1) * 3 + 1] = fcol[(i - 1) * 3 + 2] = 0.0;
            fset[i - 1] = 1;
        } else {
            if (i > nf) { rc = ORC_EBOUNDS; break; }
            int col = f_b[i - 1];
            if (col < 1 || col > nb) { rc = ORC_EBOUNDS; break; }
            memcpy(&fcol[(col - 1) * 3], &f_h[(i - 1) * 3], 3 * sizeof(double));
            fset[col - 1] = 1;
        }
        if (!in_o) {
            memset(&ocol[(i - 1) * 9], 0, 9 * sizeof(double));
            oset[i - 1] = 1;
        } else {
            if (i > no) { rc = ORC_EBOUNDS; break; }
            int col = o_b[i - 1];
            if (col < 1 || col > nb) { rc = ORC_EBOUNDS; break; }
            memcpy(&ocol[(col - 1) * 9], &o_m[(i - 1) * 9], 9 * sizeof(double));
            oset[col - 1] = 1;
        }
    }
    for (int i = 0; i < nb && rc == ORC_OK; ++i)
        if (!fset[i] || !oset[i]) rc = ORC_EUNDEF;
    if (rc != ORC_OK) {
        free(fcol); free(ocol); free(fset); free(oset);
        return rc;
    }

    for (long s = 1; s <= N; ++s) {
        int bi, x[3];
        orc_tuple_of(dim, shape, s, &bi, x);
        long si = s - 1;
        memcpy(&field[si * 3], &fcol[(bi - 1) * 3], 3 * sizeof(double));
        memcpy(&onsite[si * 9], &ocol[(bi - 1) * 9], 9 * sizeof(double));

        /* bilinear (lattice.jl:176-206) */
        for (int t = 0; t < n2; ++t) {
            long *ps = &bsite[si * n2 + t];
            double *pm = &bmat[(si * n2 + t) * 9];
            int tb1 = b2[2 * t], tb2 = b2[2 * t + 1];
            if (bi != tb1 && bi != tb2) {
                *ps = 0;
                memset(pm, 0, 9 * sizeof(double));
                continue;
            }
            int bj, sign;
            double M[9];
            memcpy(M, &m2[t * 9], 9 * sizeof(double));
            if (tb1 == tb2) { bj = bi; sign = 1; }
            else if (tb1 == bi) { bj = tb2; sign = 1; }
            else {
                /* b2 == site basis: reverse bond with transposeJ
                 * (interaction_matrix.jl:27-31) */
                bj = tb1; sign = -1;
                double T[9];
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) T[a * 3 + c] = M[c * 3 + a];
                memcpy(M, T, sizeof(T));
            }
            int off[3], nx[3];
            for (int d = 0; d < dim; ++d) off[d] = sign * off2[t * dim + d];
            orc_bc(dim, shape, periodic, x, off, nx);
            long j = orc_find(search, dim, shape, nb, N, bj, nx);
            if (j) {
                *ps = j;
                memcpy(pm, M, 9 * sizeof(double));
            } else {
                *ps = 0;
                memset(pm, 0, 9 * sizeof(double));
            }
        }

        /* cubic (lattice.jl:209-239) */
        for (int t = 0; t < n3; ++t) {
            long *ps = &csite[(si * n3 + t) * 2];
            double *pt = &ctens[(si * n3 + t) * 27];
            int tb[3] = {b3[3 * t], b3[3 * t + 1], b3[3 * t + 2]};
            const double *J = &t3[t * 27];
            int jo[3], ko[3];
            for (int d = 0; d < dim; ++d) {
                jo[d] = off3[(t * 2 + 0) * dim + d];
                ko[d] = off3[(t * 2 + 1) * dim + d];
            }
            if (bi != tb[0] && bi != tb[1] && bi != tb[2]) {
                ps[0] = ps[1] = 0;
                memset(pt, 0, 27 * sizeof(double));
                continue;
            }
            int bj, bk;
            double P[27];
            if (tb[0] == bi) {
                bj = tb[1]; bk = tb[2];
                memcpy(P, J, sizeof(P));
            } else if (tb[1] == bi) {
                /* [b2, b1, b3]: k -= j; j = -j; permutedims(J,[2,1,3]) */
                bj = tb[0]; bk = tb[2];
                for (int d = 0; d < dim; ++d) { ko[d] -= jo[d]; jo[d] = -jo[d]; }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c)
                            P[a + 3 * b + 9 * c] = J[b + 3 * a + 9 * c];
            } else {
                /* [b3, b2, b1]: j -= k; k = -k; permutedims(J,[3,2,1]) */
                bj = tb[1]; bk = tb[0];
                for (int d = 0; d < dim; ++d) { jo[d] -= ko[d]; ko[d] = -ko[d]; }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c)
                            P[a + 3 * b + 9 * c] = J[c + 3 * b + 9 * a];
            }
            int nx[3];
            orc_bc(dim, shape, periodic, x, jo, nx);
            long j = orc_find(search, dim, shape, nb, N, bj, nx);
            orc_bc(dim, shape, periodic, x, ko, nx);
            long k = orc_find(search, dim, shape, nb, N, bk, nx);
            if (!j || !k) {
                ps[0] = ps[1] = 0;
                memset(pt, 0, 27 * sizeof(double));
            } else {
                ps[0] = j; ps[1] = k;
                memcpy(pt, P, sizeof(P));
            }
        }

        /* quartic (lattice.jl:243-286) */
        for (int t = 0; t < n4; ++t) {
            long *ps = &qsite[(si * n4 + t) * 3];
            double *pt = &qtens[(si * n4 + t) * 81];
            int tb[4] = {b4[4 * t], b4[4 * t + 1], b4[4 * t + 2], b4[4 * t + 3]};
            const double *J = &t4[t * 81];
            int jo[3], ko[3], lo[3];
            for (int d = 0; d < dim; ++d) {
                jo[d] = off4[(t * 3 + 0) * dim + d];
                ko[d] = off4[(t * 3 + 1) * dim + d];
                lo[d] = off4[(t * 3 + 2) * dim + d];
            }
            if (bi != tb[0] && bi != tb[1] && bi != tb[2] && bi != tb[3]) {
                ps[0] = ps[1] = ps[2] = 0;
                memset(pt, 0, 81 * sizeof(double));
                continue;
            }
            int bj, bk, bl;
            double P[81];
            int perm; /* which role the site takes: 0..3 */
            if (tb[0] == bi) { bj = tb[1]; bk = tb[2]; bl = tb[3]; perm = 0; }
            else if (tb[1] == bi) {
                bj = tb[0]; bk = tb[2]; bl = tb[3]; perm = 1;
                for (int d = 0; d < dim; ++d) {
                    jo[d] = -jo[d]; ko[d] += jo[d]; lo[d] += jo[d];
                }
            } else if (tb[2] == bi) {
                bj = tb[1]; bk = tb[0]; bl = tb[3]; perm = 2;
                for (int d = 0; d < dim; ++d) {
                    ko[d] = -ko[d]; lo[d] += ko[d]; jo[d] += ko[d];
                }
            } else {
                bj = tb[1]; bk = tb[2]; bl = tb[0]; perm = 3;
                for (int d = 0; d < dim; ++d) {
                    lo[d] = -lo[d]; jo[d] += lo[d]; ko[d] += lo[d];
                }
            }
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                        for (int e = 0; e < 3; ++e) {
                            int src;
                            /* permutedims with an involutive perm:
                             * B[i_p1, i_p2, ..] = A[i1, i2, ..] */
                            switch (perm) {
                            case 0: src = a + 3 * b + 9 * c + 27 * e; break;
                            case 1: src = b + 3 * a + 9 * c + 27 * e; break;
                            case 2: src = c + 3 * b + 9 * a + 27 * e; break;
                            default: src = e + 3 * b + 9 * c + 27 * a; break;
                            }
                            P[a + 3 * b + 9 * c + 27 * e] = J[src];
                        }
            int nx[3];
            orc_bc(dim, shape, periodic, x, jo, nx);
            long j = orc_find(search, dim, shape, nb, N, bj, nx);
            orc_bc(dim, shape, periodic, x, ko, nx);
            long k = orc_find(search, dim, shape, nb, N, bk, nx);
            orc_bc(dim, shape, periodic, x, lo, nx);
            long l = orc_find(search, dim, shape, nb, N, bl, nx);
            if (!j || !k || !l) {
                ps[0] = ps[1] = ps[2] = 0;
                memset(pt, 0, 81 * sizeof(double));
            } else {
                ps[0] = j; ps[1] = k; ps[2] = l;
                memcpy(pt, P, sizeof(P));
            }
        }
    }
    free(fcol); free(ocol); free(fset); free(oset);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Hamiltonian (hamiltonian.jl)                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    long N;
    int n2, n3, n4;
    const long *bsite; const double *bmat;
    const long *csite; const double *ctens;
    const long *qsite; const double *qtens;
    const double *field; const double *onsite;
} orc_lat;

static inline const double *spin_at(const double *s, long site1) { return &s[3 * (site1 - 1)]; }

/* get_local_field (hamiltonian.jl:3-67); point is 1-based */
void orc_local_field(const orc_lat *L, const double *spins, long point, double *H)
{
    long i = point - 1;
    const double *s = spin_at(spins, point);
    const double *o = &L->onsite[i * 9];
    double Hx = 0.0, Hy = 0.0, Hz = 0.0;
    Hx += 2 * (o[0] * s[0] + o[1] * s[1] + o[2] * s[2]);
    Hy += 2 * (o[3] * s[0] + o[4] * s[1] + o[5] * s[2]);
    Hz += 2 * (o[6] * s[0] + o[7] * s[1] + o[8] * s[2]);
    for (int n = 0; n < L->n2; ++n) {
        long j = L->bsite[i * L->n2 + n];
        if (j == 0) continue;
        const double *J = &L->bmat[(i * L->n2 + n) * 9];
        const double *sj = spin_at(spins, j);
        Hx += J[0] * sj[0] + J[1] * sj[1] + J[2] * sj[2];
        Hy += J[3] * sj[0] + J[4] * sj[1] + J[5] * sj[2];
        Hz += J[6] * sj[0] + J[7] * sj[1] + J[8] * sj[2];
    }
    for (int n = 0; n < L->n3; ++n) {
        long j = L->csite[(i * L->n3 + n) * 2], k = L->csite[(i * L->n3 + n) * 2 + 1];
        if (j == 0 && k == 0) continue;
        const double *C = &L->ctens[(i * L->n3 + n) * 27];
        const double *sj = spin_at(spins, j), *sk = spin_at(spins, k);
        double h3[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) h3[a] += C[a + 3 * b + 9 * c] * sj[b] * sk[c];
        Hx += h3[0]; Hy += h3[1]; Hz += h3[2];
    }
    for (int n = 0; n < L->n4; ++n) {
        const long *q = &L->qsite[(i * L->n4 + n) * 3];
        if (q[0] == 0 && q[1] == 0 && q[2] == 0) continue;
        const double *R = &L->qtens[(i * L->n4 + n) * 81];
        const double *sj = spin_at(spins, q[0]), *sk = spin_at(spins, q[1]), *sl = spin_at(spins, q[2]);
        double h4[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d)
                        h4[a] += R[a + 3 * b + 9 * c + 27 * d] * sj[b] * sk[c] * sl[d];
        Hx += h4[0]; Hy += h4[1]; Hz += h4[2];
    }
    const double *h = &L->field[i * 3];
    H[0] = Hx - h[0];
    H[1] = Hy - h[1];
    H[2] = Hz - h[2];
}

/* energy(lattice, point) (hamiltonian.jl:139-196): no 1/2, 1/3, 1/4 */
/* get_local_field for every site (points 1..N): H[N][3] */
void orc_local_fields(const orc_lat *L, const double *spins, double *H)
{
    for (long p = 1; p <= L->N; ++p) orc_local_field(L, spins, p, H + 3 * (p - 1));
}

double orc_site_energy(const orc_lat *L, const double *spins, long point)
{
    long i = point - 1;
    const double *s = spin_at(spins, point);
    const double *o = &L->onsite[i * 9];
    double E = 0.0;
    E += s[0] * (o[0] * s[0] + o[1] * s[1] + o[2] * s[2]) +
         s[1] * (o[3] * s[0] + o[4] * s[1] + o[5] * s[2]) +
         s[2] * (o[6] * s[0] + o[7] * s[1] + o[8] * s[2]);
    for (int n = 0; n < L->n2; ++n) {
        long j = L->bsite[i * L->n2 + n];
        if (j == 0) continue;
        const double *J = &L->bmat[(i * L->n2 + n) * 9];
        const double *sj = spin_at(spins, j);
        E += s[0] * (J[0] * sj[0] + J[1] * sj[1] + J[2] * sj[2]) +
             s[1] * (J[3] * sj[0] + J[4] * sj[1] + J[5] * sj[2]) +
             s[2] * (J[6] * sj[0] + J[7] * sj[1] + J[8] * sj[2]);
    }
    for (int n = 0; n < L->n3; ++n) {
        long j = L->csite[(i * L->n3 + n) * 2], k = L->csite[(i * L->n3 + n) * 2 + 1];
        if (j == 0 && k == 0) continue;
        const double *C = &L->ctens[(i * L->n3 + n) * 27];
        const double *sj = spin_at(spins, j), *sk = spin_at(spins, k);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) E += C[a + 3 * b + 9 * c] * s[a] * sj[b] * sk[c];
    }
    for (int n = 0; n < L->n4; ++n) {
        const long *q = &L->qsite[(i * L->n4 + n) * 3];
        if (q[0] == 0 && q[1] == 0 && q[2] == 0) continue;
        const double *R = &L->qtens[(i * L->n4 + n) * 81];
        const double *sj = spin_at(spins, q[0]), *sk = spin_at(spins, q[1]), *sl = spin_at(spins, q[2]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d)
                        E += R[a + 3 * b + 9 * c + 27 * d] * s[a] * sj[b] * sk[c] * sl[d];
    }
    const double *h = &L->field[i * 3];
    return E - (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
}

/* total_energy (hamiltonian.jl:70-132).  parts[5] receives the five
 * accumulators (bilinear, cubic, quartic, zeeman, onsite) before the
 * 1/2, 1/3, 1/4 weights of hamiltonian.jl:131. */
double orc_total_energy(const orc_lat *L, const double *spins, double *parts)
{
    double Eb = 0.0, Ec = 0.0, Eq = 0.0, Ez = 0.0, Eo = 0.0;
    for (long p = 1; p <= L->N; ++p) {
        long i = p - 1;
        const double *s = spin_at(spins, p);
        const double *h = &L->field[i * 3];
        const double *o = &L->onsite[i * 9];
        Ez -= (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
        Eo += s[0] * (o[0] * s[0] + o[1] * s[1] + o[2] * s[2]) +
              s[1] * (o[3] * s[0] + o[4] * s[1] + o[5] * s[2]) +
              s[2] * (o[6] * s[0] + o[7] * s[1] + o[8] * s[2]);
        for (int n = 0; n < L->n2; ++n) {
            long j = L->bsite[i * L->n2 + n];
            if (j == 0) continue;
            const double *J = &L->bmat[(i * L->n2 + n) * 9];
            const double *sj = spin_at(spins, j);
            Eb += s[0] * (J[0] * sj[0] + J[1] * sj[1] + J[2] * sj[2]) +
                  s[1] * (J[3] * sj[0] + J[4] * sj[1] + J[5] * sj[2]) +
                  s[2] * (J[6] * sj[0] + J[7] * sj[1] + J[8] * sj[2]);
        }
        for (int n = 0; n < L->n3; ++n) {
            long j = L->csite[(i * L->n3 + n) * 2], k = L->csite[(i * L->n3 + n) * 2 + 1];
            if (j == 0 && k == 0) continue;
            const double *C = &L->ctens[(i * L->n3 + n) * 27];
            const double *sj = spin_at(spins, j), *sk = spin_at(spins, k);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) Ec += C[a + 3 * b + 9 * c] * s[a] * sj[b] * sk[c];
        }
        for (int n = 0; n < L->n4; ++n) {
            const long *q = &L->qsite[(i * L->n4 + n) * 3];
            if (q[0] == 0 && q[1] == 0 && q[2] == 0) continue;
            const double *R = &L->qtens[(i * L->n4 + n) * 81];
            const double *sj = spin_at(spins, q[0]), *sk = spin_at(spins, q[1]),
                         *sl = spin_at(spins, q[2]);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d)
                            Eq += R[a + 3 * b + 9 * c + 27 * d] * s[a] * sj[b] * sk[c] * sl[d];
        }
    }
    if (parts) { parts[0] = Eb; parts[1] = Ec; parts[2] = Eq; parts[3] = Ez; parts[4] = Eo; }
    return Eb / 2 + Ec / 3 + Eq / 4 + Ez + Eo;
}

/* ------------------------------------------------------------------------ */
/* Sweep primitives                                                          */
/* ------------------------------------------------------------------------ */

/* One overrelaxation site visit (monte_carlo.jl:126-139 loop body). */
void orc_overrelax_site(const orc_lat *L, double *spins, long point)
{
    double H[3];
    double *s = &spins[3 * (point - 1)];
    double si[3] = {s[0], s[1], s[2]};
    orc_local_field(L, spins, point, H);
    if (H[0] == 0.0 && H[1] == 0.0 && H[2] == 0.0) return;
    double proj = 2.0 * (si[0] * H[0] + si[1] * H[1] + si[2] * H[2]) /
                  (H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    s[0] = -si[0] + proj * H[0];
    s[1] = -si[1] + proj * H[1];
    s[2] = -si[2] + proj * H[2];
}

/* overrelaxation! over an explicit visiting order (1-based sites).  With
 * order = 1..N this is monte_carlo.jl:127 verbatim; with the sites of one
 * colour class it is the reference loop restricted to that class. */
void orc_overrelax_order(const orc_lat *L, double *spins, const long *order, long n)
{
    for (long k = 0; k < n; ++k) orc_overrelax_site(L, spins, order[k]);
}

/* deterministic update of one site (monte_carlo.jl:201-213 loop body) */
void orc_deterministic_site(const orc_lat *L, double *spins, long point, double S)
{
    double H[3];
    orc_local_field(L, spins, point, H);
    if (H[0] == 0.0 && H[1] == 0.0 && H[2] == 0.0) return;
    double nrm = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    double *s = &spins[3 * (point - 1)];
    s[0] = -H[0] / nrm * S;
    s[1] = -H[1] / nrm * S;
    s[2] = -H[2] / nrm * S;
}

void orc_deterministic_order(const orc_lat *L, double *spins, const long *order, long n, double S)
{
    for (long k = 0; k < n; ++k) orc_deterministic_site(L, spins, order[k], S);
}

/* random_spin_orientation (lattice.jl:306-311) from two uniforms */
void orc_random_spin(double S, double u1, double u2, double *out)
{
    double phi = 2.0 * M_PI * u1;
    double z = 2.0 * u2 - 1.0;
    double r = sqrt(1.0 - z * z);
    out[0] = S * (r * cos(phi));
    out[1] = S * (r * sin(phi));
    out[2] = S * z;
}

/* gaussian_move (metropolis.jl:84-87): s + sigma*random_spin, renormalised */
void orc_cone_move(double S, const double *spin, double sigma, double u1, double u2, double *out)
{
    double r[3];
    orc_random_spin(S, u1, u2, r);
    double n0 = spin[0] + sigma * r[0], n1 = spin[1] + sigma * r[1], n2 = spin[2] + sigma * r[2];
    double nrm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    out[0] = n0 / nrm * S;
    out[1] = n1 / nrm * S;
    out[2] = n2 / nrm * S;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with the published */
/* multipliers and Weyl key increments.                                      */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t *ctr_in, const uint32_t *key_in, uint32_t *out)
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in the open interval (0,1) from 32 random bits */
static inline double orc_u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

/* counter layout shared by the coloured schedule:
 *   key = (seed lo, seed hi)
 *   ctr = (0-based reference site index, stream, sweep counter, tag)   */
#define ORC_TAG_METRO 0x4D455452u /* "METR" */
#define ORC_TAG_EXCH 0x45584348u  /* "EXCH" */

/* ------------------------------------------------------------------------ */
/* Metropolis                                                                */
/* ------------------------------------------------------------------------ */

/* One proposal at `point` (metropolis.jl:70-78 body, calculate_energy_diff!
 * metropolis.jl:94-110): E_old, write proposal, E_new, accept if dE<0 or
 * u < exp(-dE/T), else restore.  kind: 0 = uniform (Metropolis), 1 = cone
 * (adaptive / fixed cone) with width sigma. Returns 1 if accepted. */
static int orc_propose(const orc_lat *L, double *spins, long point, double S, double T,
                       int kind, double sigma, double u1, double u2, double u3, double *dE_out)
{
    double *s = &spins[3 * (point - 1)];
    double old[3] = {s[0], s[1], s[2]};
    double E_old = orc_site_energy(L, spins, point);
    double nw[3];
    if (kind == 0) orc_random_spin(S, u1, u2, nw);
    else orc_cone_move(S, old, sigma, u1, u2, nw);
    s[0] = nw[0]; s[1] = nw[1]; s[2] = nw[2];
    double E_new = orc_site_energy(L, spins, point);
    double dE = E_new - E_old;
    if (dE_out) *dE_out = dE;
    int accept = (dE < 0) ? 1 : (u3 < exp(-dE / T));
    if (!accept) { s[0] = old[0]; s[1] = old[1]; s[2] = old[2]; }
    return accept;
}

/* Coloured Metropolis sweep with the Philox stream: the classes are visited
 * in order, each site of a class proposed once.  Within a class the sites
 * are independent, so the visiting order inside a class does not change the
 * result; this is the same chain a parallel colour pass draws.
 * order[] lists 1-based sites class after class.  decisions[k] (optional)
 * receives accept(1)/reject(0) for order[k]. Returns the accepted count. */
long orc_metropolis_coloured(const orc_lat *L, double *spins, const long *order, long n,
                             double S, double T, int kind, double sigma, uint64_t seed,
                             uint32_t stream, uint32_t sweep_ctr, unsigned char *decisions,
                             double *dE_out)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    long acc = 0;
    for (long k = 0; k < n; ++k) {
        long p = order[k];
        uint32_t ctr[4] = {(uint32_t)(p - 1), stream, sweep_ctr, ORC_TAG_METRO};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        double dE;
        int a = orc_propose(L, spins, p, S, T, kind, sigma, orc_u01(r[0]), orc_u01(r[1]),
                            orc_u01(r[2]), &dE);
        if (decisions) decisions[k] = (unsigned char)a;
        if (dE_out) dE_out[k] = dE;
        acc += a;
    }
    return acc;
}

/* metropolis_constraint_adaptive! (metropolis.jl:155-179) over the sites of
 * `order` (1-based, class after class), for the site-sum constraint
 *     c(lattice) = sum_i (a_b . s_i + s_i . Q_b s_i) - c0,  b = (i-1) / ncell,
 * a[cnb][3], Q[cnb][9] row-major (bases >= cnb contribute 0).  Per proposal,
 * literally as the reference: the cone move with width *sigma and the Philox
 * draws of orc_metropolis_coloured, dE = energy(new) - energy(old), c
 * evaluated over the whole lattice with the new spin in place (:164),
 * new_lambda = lambda - weight * c (:165), acceptance on
 * dE - new_lambda * c (:167), restore on rejection, lambda = new_lambda on
 * acceptance (:172).  After the n proposals the width law (:176-177):
 * f = 0.5 / (1 - rate); sigma = rate < 0.5 ? sigma * f : sigma0.
 * decisions[k] (optional) receives the decision for order[k].  Returns the
 * accepted count. */
static double orc_constraint_value(const orc_lat *L, const double *spins, long ncell, const double *ca,
                                   const double *cQ, int cnb, double c0)
{
    double c = 0.0;
    for (long i = 0; i < L->N; ++i) {
        long b = i / ncell;
        if (b >= cnb) continue;
        const double *s = &spins[3 * i], *q = &cQ[9 * b], *a = &ca[3 * b];
        double lin = a[0] * s[0] + a[1] * s[1] + a[2] * s[2];
        double quad = 0.0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) quad += s[x] * q[3 * x + y] * s[y];
        c += lin + quad;
    }
    return c - c0;
}

long orc_metropolis_constraint_coloured(const orc_lat *L, double *spins, const long *order, long n,
                                        double S, double T, double *sigma, double sigma0, uint64_t seed,
                                        uint32_t stream, uint32_t sweep_ctr, long ncell, const double *ca,
                                        const double *cQ, int cnb, double c0, double weight, double *lambda,
                                        unsigned char *decisions)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    long acc = 0;
    for (long k = 0; k < n; ++k) {
        long p = order[k];
        uint32_t ctr[4] = {(uint32_t)(p - 1), stream, sweep_ctr, ORC_TAG_METRO};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        double *s = &spins[3 * (p - 1)];
        double old[3] = {s[0], s[1], s[2]};
        double E_old = orc_site_energy(L, spins, p);
        double nw[3];
        orc_cone_move(S, old, *sigma, orc_u01(r[0]), orc_u01(r[1]), nw);
        s[0] = nw[0]; s[1] = nw[1]; s[2] = nw[2];
        double dE = orc_site_energy(L, spins, p) - E_old;
        double c = orc_constraint_value(L, spins, ncell, ca, cQ, cnb, c0);
        double new_lambda = *lambda - weight * c;
        double x = dE - new_lambda * c;
        int accept = (x < 0) ? 1 : (orc_u01(r[2]) < exp(-x / T));
        if (!accept) {
            s[0] = old[0]; s[1] = old[1]; s[2] = old[2];
        } else {
            ++acc;
            *lambda = new_lambda;
        }
        if (decisions) decisions[k] = (unsigned char)accept;
    }
    double rate = (double)acc / (double)n;
    double f = 0.5 / (1.0 - rate);
    *sigma = rate < 0.5 ? *sigma * f : sigma0;
    return acc;
}

/* The reference's random-site sweep (metropolis.jl:65-82): N proposals at
 * uniformly drawn sites, with replacement.  Julia's GLOBAL_RNG is replaced
 * by a Philox stream (site = floor(u*N)+1). Returns the accepted count. */
long orc_metropolis_random(const orc_lat *L, double *spins, double S, double T, int kind,
                           double sigma, uint64_t seed, uint32_t stream, uint32_t sweep_ctr)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    long acc = 0;
    for (long k = 0; k < L->N; ++k) {
        uint32_t ctr[4] = {(uint32_t)k, stream, sweep_ctr, 0x52414E44u /* "RAND" */};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        long p = (long)(orc_u01(r[3]) * (double)L->N) + 1;
        if (p > L->N) p = L->N;
        acc += orc_propose(L, spins, p, S, T, kind, sigma, orc_u01(r[0]), orc_u01(r[1]),
                           orc_u01(r[2]), NULL);
    }
    return acc;
}

/* metropolis_adaptive! width update (metropolis.jl:129-131) */
double orc_adapt_sigma(double sigma, double accepted, double attempts)
{
    double a = accepted / attempts;
    double d = 1.0 - a;
    double f = 0.5 / (d > 0.05 ? d : 0.05);
    double s = sigma * f;
    if (s < 0.0) s = 0.0;
    if (s > 100.0) s = 100.0;
    return s;
}

/* get_magnetization (observables.jl:12-18): |sum_i s_i| */
double orc_magnetization(const double *spins, long N, double *m3)
{
    double m0 = 0, m1 = 0, m2 = 0;
    for (long i = 0; i < N; ++i) { m0 += spins[3 * i]; m1 += spins[3 * i + 1]; m2 += spins[3 * i + 2]; }
    if (m3) { m3[0] = m0; m3[1] = m1; m3[2] = m2; }
    return sqrt(m0 * m0 + m1 * m1 + m2 * m2);
}

/* ------------------------------------------------------------------------ */
/* Replica exchange (monte_carlo.jl:308-349), restated over temperature      */
/* slots.  slot r holds temperature temps[r]; energy[r] the energy of the    */
/* configuration currently at slot r.  For exchange number `xc` the pairs    */
/* are (0,1),(2,3).. when xc is even, (1,2),(3,4).. when odd                 */
/* (monte_carlo.jl:311-315, with xc = sweep/swap_rate).  The even member of  */
/* each pair decides: u < min(1, exp((1/T_p - 1/T_r)(E_p - E_r)))            */
/* (monte_carlo.jl:326-331).  On accept the configurations swap: here the    */
/* caller-supplied slot->config map `perm` and `energy` are swapped.         */
/* Returns the number of accepted swaps; accepted[r] counts per slot.        */
/* ------------------------------------------------------------------------ */
int orc_pt_exchange(int R, const double *temps, double *energy, int *perm, uint64_t seed,
                    uint32_t xc, long *accepted)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int n = 0;
    for (int r = (xc % 2 == 0) ? 0 : 1; r + 1 < R; r += 2) {
        int p = r + 1;
        uint32_t ctr[4] = {(uint32_t)r, 0xFFFFFFFFu, xc, ORC_TAG_EXCH};
        uint32_t rr[4];
        orc_philox4x32_10(ctr, key, rr);
        double u = orc_u01(rr[0]);
        double dbeta = (1.0 / temps[p] - 1.0 / temps[r]);
        double dE = energy[p] - energy[r];
        double pacc = exp(dbeta * dE);
        if (pacc > 1.0) pacc = 1.0;
        if (u < pacc) {
            double te = energy[r]; energy[r] = energy[p]; energy[p] = te;
            int tp = perm[r]; perm[r] = perm[p]; perm[p] = tp;
            if (accepted) { accepted[r]++; accepted[p]++; }
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Logarithmic binning with error propagation (observables.jl:5-63 uses     */
/* BinningAnalysis.jl's ErrorPropagator{Float64,32}).  Restated from the    */
/* published method (Ambegaokar & Troyer, Am. J. Phys. 78, 150 (2010)):     */
/* level l holds the running sums of bins that average 2^(l-1) consecutive  */
/* samples; a pending value per level pairs samples up.                     */
/* State layout, K observables (here 2), 32 levels:                         */
/*   count[32] (as double), sum[32*K], sum2[32*K*K], pend[32*K], has[32]    */
/* ------------------------------------------------------------------------ */
#define ORC_LVLS 32
typedef struct {
    int K;
    double count[ORC_LVLS];
    double sum[ORC_LVLS * 2];
    double sum2[ORC_LVLS * 4];
    double pend[ORC_LVLS * 2];
    int has[ORC_LVLS];
} orc_binner;

void orc_binner_init(orc_binner *b, int K)
{
    memset(b, 0, sizeof(*b));
    b->K = K;
}

void orc_binner_push(orc_binner *b, const double *v)
{
    double cur[2] = {v[0], b->K > 1 ? v[1] : 0.0};
    for (int l = 0; l < ORC_LVLS; ++l) {
        int K = b->K;
        b->count[l] += 1.0;
        for (int i = 0; i < K; ++i) {
            b->sum[l * 2 + i] += cur[i];
            for (int j = 0; j < K; ++j) b->sum2[l * 4 + i * 2 + j] += cur[i] * cur[j];
        }
        if (!b->has[l]) {
            b->has[l] = 1;
            b->pend[l * 2] = cur[0];
            b->pend[l * 2 + 1] = cur[1];
            return;
        }
        b->has[l] = 0;
        cur[0] = 0.5 * (b->pend[l * 2] + cur[0]);
        cur[1] = 0.5 * (b->pend[l * 2 + 1] + cur[1]);
    }
}

/* highest level with at least 32 bins (level 0 if none) */
int orc_binner_reliable_level(const orc_binner *b)
{
    int lvl = 0;
    for (int l = 0; l < ORC_LVLS; ++l)
        if (b->count[l] >= 32.0) lvl = l;
    return lvl;
}

/* means at level 0 */
void orc_binner_means(const orc_binner *b, double *m)
{
    for (int i = 0; i < b->K; ++i) m[i] = b->sum[i] / b->count[0];
}

/* gradient-propagated variance of the bin values at level lvl
 * (unbiased covariance of the level's bins contracted with g) */
double orc_binner_var(const orc_binner *b, const double *g, int lvl)
{
    double n = b->count[lvl];
    if (n < 2.0) return 0.0;
    double v = 0.0;
    for (int i = 0; i < b->K; ++i)
        for (int j = 0; j < b->K; ++j) {
            double mi = b->sum[lvl * 2 + i] / n, mj = b->sum[lvl * 2 + j] / n;
            double cij = (b->sum2[lvl * 4 + i * 2 + j] / n - mi * mj) * n / (n - 1.0);
            v += g[i] * cij * g[j];
        }
    return v;
}

/* std_error_tweak (observables.jl:32-34): sqrt(|var/count|) at lvl */
double orc_binner_std_error(const orc_binner *b, const double *g, int lvl)
{
    return sqrt(fabs(orc_binner_var(b, g, lvl) / b->count[lvl]));
}

/* specific_heat (observables.jl:36-49) and susceptibility (:51-63);
 * out = {C, dC, chi, dchi, <E>, dE, <M>, dM} using the level-0 means and the
 * reliable level, errors as observables.jl computes them. */
void orc_observables(const orc_binner *e, const orc_binner *m, double T, long N, double *out)
{
    double me[2], mm[2];
    orc_binner_means(e, me);
    orc_binner_means(m, mm);
    double C = 1.0 / (T * T) * (me[1] - me[0] * me[0]) / (double)N;
    double gC[2] = {-2.0 * 1.0 / (T * T) * me[0] / (double)N, 1.0 / (T * T) / (double)N};
    double dC = orc_binner_std_error(e, gC, orc_binner_reliable_level(e));
    double X = 1.0 / T * (mm[1] - mm[0] * mm[0]) / (double)N;
    double gX[2] = {-2.0 * 1.0 / T * mm[0] / (double)N, 1.0 / T / (double)N};
    double dX = orc_binner_std_error(m, gX, orc_binner_reliable_level(m));
    double g0[2] = {1.0, 0.0};
    out[0] = C; out[1] = dC; out[2] = X; out[3] = dX;
    out[4] = me[0]; out[5] = orc_binner_std_error(e, g0, orc_binner_reliable_level(e));
    out[6] = mm[0]; out[7] = orc_binner_std_error(m, g0, orc_binner_reliable_level(m));
}

size_t orc_binner_size(void) { return sizeof(orc_binner); }

/* ------------------------------------------------------------------------ */
/* The reference's parallel_tempering! sweep loop (monte_carlo.jl:295-349)   */
/* for R replicas held in one process: per sweep, every replica does a      */
/* sequential overrelaxation sweep (monte_carlo.jl:298-300), and every      */
/* dosweep-th sweep a random-site Metropolis sweep + total_energy           */
/* (:302-305) followed by the even/odd exchange (:308-349).  spins holds    */
/* R configurations of N sites; perm[r] is the configuration at slot r.     */
/* Returns the number of spin updates performed (the CPU baseline unit).    */
/* ------------------------------------------------------------------------ */
long orc_pt_reference_run(const orc_lat *L, double *spins, int R, const double *temps, double S,
                          long sweep0, long n_sweeps, long or_rate, long swap_rate, uint64_t seed,
                          double *E, int *perm, long *acc_local, long *acc_exch)
{
    long dosweep = or_rate == 0 ? 1 : or_rate;
    long updates = 0;
    long *order = (long *)malloc(sizeof(long) * (size_t)L->N);
    for (long i = 0; i < L->N; ++i) order[i] = i + 1;
    for (long sw = sweep0; sw < sweep0 + n_sweeps; ++sw) {
        for (int r = 0; r < R; ++r) {
            double *s = spins + (size_t)perm[r] * 3 * L->N;
            if (or_rate != 0) {
                orc_overrelax_order(L, s, order, L->N);
                updates += L->N;
            }
            if (sw % dosweep == 0) {
                long a = orc_metropolis_random(L, s, S, temps[r], 0, 60.0, seed, (uint32_t)perm[r],
                                               (uint32_t)(sw / dosweep));
                if (acc_local) acc_local[r] += a;
                updates += L->N;
                E[r] = orc_total_energy(L, s, NULL);
            }
        }
        if (R > 1 && sw % dosweep == 0 && sw % swap_rate == 0)
            orc_pt_exchange(R, temps, E, perm, seed, (uint32_t)(sw / swap_rate), acc_exch);
    }
    free(order);
    return updates;
}

/* ------------------------------------------------------------------------ */
/* compute_equal_time_correlations (spin_correlations.jl:6-43): for each    */
/* momentum k (column n of the D x n_k matrix ks), s_a(k) = sum_i            */
/* exp(-i k.r_i) s_{i,a} with r_i the site positions (D x N, lattice.jl:12), */
/* then Suv[3a+b][k] = Re(s_a(k) conj(s_b(k))) / N.  out is the reference's   */
/* 9 x n_k column-major matrix: out[9 k + 3 a + b].                          */
/* ------------------------------------------------------------------------ */
void orc_equal_time_correlations(const double *pos, long N, int D, const double *spins,
                                 const double *ks, long nk, double *out)
{
    for (long n = 0; n < nk; ++n) {
        double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
        for (long i = 0; i < N; ++i) {
            double arg = 0.0;
            for (int d = 0; d < D; ++d) arg += ks[n * D + d] * pos[i * D + d];
            const double c = cos(arg), sn = -sin(arg);  /* exp(-i arg) */
            for (int a = 0; a < 3; ++a) {
                re[a] += c * spins[3 * i + a];
                im[a] += sn * spins[3 * i + a];
            }
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                out[9 * n + 3 * a + b] = (re[a] * re[b] + im[a] * im[b]) / (double)N;
    }
}
