/*
 * ls1_oracle.c -- CPU restatement (scalar FP64 C99) of ls1-MarDyn's
 * short-range force path: linked-cell geometry and binning, periodic wrap and
 * halo copies, the C08 cell-pair traversal, the VectorizedCellProcessor
 * site-pair bodies with their masks and macroscopic sums, FullMolecule's SoA
 * set-up and calcFM, and the Leapfrog half steps.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker the tests, the smoke check and
 * the benchmark's cpu_baseline leg compare the HIP path against; nothing in
 * the product links or calls it.
 *
 * Every function names the reference lines whose behaviour it restates
 * (paths relative to the ls1-MarDyn source root). The arithmetic follows the
 * reference's NOVEC double-precision build: IEEE 1/x and sqrt, no fused
 * multiply-add (built with -ffp-contract=off).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *  - pair bodies: closed-form KATs (VectorizedCellProcessorTest.cpp:45-134),
 *    the reference's own potforce.h evaluated through oracle/_ref, and
 *    finite differences F = -dU/dr, torque = -e x dU/de;
 *  - traversal / enumeration: brute-force O(N^2) minimum-image sums, the
 *    reference's Legacy-path output on its Vectorization*.inp fixtures
 *    (through oracle/_ref), and a CellProcessorStub-style coverage count;
 *  - cell index: exhaustive brute-force point-in-box search;
 *  - Leapfrog: FullMolecule::upd_preF/upd_postF through oracle/_ref, plus
 *    energy conservation and a free-rotor invariant;
 *  - component mass/inertia and LJ tables: Component.cpp / Comp2Param.cpp
 *    through oracle/_ref.
 */
#include "ls1_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define OX_MAXCOMP 16
#define OX_MAXSITE 8

typedef struct {
    int nlj, nc, nd, nq;
    double lj[OX_MAXSITE][3], lj_m[OX_MAXSITE], lj_eps[OX_MAXSITE],
        lj_sig[OX_MAXSITE], lj_shift6[OX_MAXSITE];
    double lj_rcs[OX_MAXSITE]; int lj_shifted[OX_MAXSITE]; /* own shift cutoff, flag */
    double c[OX_MAXSITE][3], c_m[OX_MAXSITE], c_q[OX_MAXSITE];
    double d[OX_MAXSITE][3], d_e[OX_MAXSITE][3], d_p[OX_MAXSITE];
    double qd[OX_MAXSITE][3], q_e[OX_MAXSITE][3], q_Q[OX_MAXSITE];
    double m, I[3], invI[3];
    int rotdof;
    int lj_base; /* first LJ lookup id (EnsembleBase.cpp:98-111) */
} ox_comp;

typedef struct {
    int ncomp;
    ox_comp comp[OX_MAXCOMP];
    int nlj_total;
    int nthreads;   /* OpenMP threads of the traversal and the molecule loops (1: serial) */
    double *eps24, *sig2, *shift6; /* nlj_total x nlj_total */
    double epsRF, rc, rcLJ, epsRFInvrc3;
    double bbmin[3], bbmax[3];
    int periodic;   /* any dimension periodic */
    int pdim[3];    /* periodic per dimension */
    int wall[3];    /* reflecting walls per dimension */
    int outflow[3]; /* outflow walls per dimension */
    /* LinkedCells geometry */
    int boxcells[3], cpd[3];
    double cell_len[3], cell_rlen[3], halo_min[3], halo_max[3];
} ox_ctx;

/* ------------------------------------------------------------------ model */

/* Component::calculateLJshift, src/molecules/Component.cpp:118-122 */
static double lj_shift6(double eps, double sigma, double rc) {
    double sigperrc2 = sigma * sigma / (rc * rc);
    double sigperrc6 = sigperrc2 * sigperrc2 * sigperrc2;
    return 24.0 * eps * (sigperrc6 - sigperrc6 * sigperrc6);
}

/* Component::updateMassInertia(Site&), src/molecules/Component.cpp:151-179:
 * only LJ centres and charges carry mass; sites are in the principal frame. */
static void add_mass_site(double *m, double *I6, int *rotdof, const double *r,
                          double sm) {
    *m += sm;
    I6[0] += sm * (r[1] * r[1] + r[2] * r[2]);
    I6[1] += sm * (r[0] * r[0] + r[2] * r[2]);
    I6[2] += sm * (r[0] * r[0] + r[1] * r[1]);
    *rotdof = 3;
    for (int d = 0; d < 3; ++d)
        if (I6[d] == 0.) --(*rotdof);
}

static int parse_components(ox_ctx *x, int ncomp, const double *desc, int len) {
    int p = 0, lj_base = 0;
    if (ncomp < 1 || ncomp > OX_MAXCOMP) return -1;
    x->ncomp = ncomp;
    for (int ci = 0; ci < ncomp; ++ci) {
        ox_comp *c = &x->comp[ci];
        double I6[3] = {0, 0, 0};
        int rotdof = 3;
        memset(c, 0, sizeof(*c));
        if (p + 4 > len) return -2;
        c->nlj = (int)desc[p];
        c->nc = (int)desc[p + 1];
        c->nd = (int)desc[p + 2];
        c->nq = (int)desc[p + 3];
        p += 4;
        if (c->nlj > OX_MAXSITE || c->nc > OX_MAXSITE || c->nd > OX_MAXSITE ||
            c->nq > OX_MAXSITE)
            return -3;
        if (p + 8 * c->nlj + 5 * c->nc + 7 * c->nd + 7 * c->nq + 3 > len) return -2;
        rotdof = 0;
        /* ASCIIReader.cpp:171-177 -> Component::addLJcenter (Component.cpp:103-116) */
        for (int s = 0; s < c->nlj; ++s) {
            const double *v = desc + p;
            for (int d = 0; d < 3; ++d) c->lj[s][d] = v[d];
            c->lj_m[s] = v[3];
            c->lj_eps[s] = v[4];
            c->lj_sig[s] = v[5];
            c->lj_shift6[s] = (v[7] != 0.) ? lj_shift6(v[4], v[5], v[6]) : 0.0;
            c->lj_rcs[s] = v[6];
            c->lj_shifted[s] = v[7] != 0.;
            add_mass_site(&c->m, I6, &rotdof, c->lj[s], c->lj_m[s]);
            p += 8;
        }
        for (int s = 0; s < c->nc; ++s) {
            const double *v = desc + p;
            for (int d = 0; d < 3; ++d) c->c[s][d] = v[d];
            c->c_m[s] = v[3];
            c->c_q[s] = v[4];
            add_mass_site(&c->m, I6, &rotdof, c->c[s], c->c_m[s]);
            p += 5;
        }
        for (int s = 0; s < c->nd; ++s) {
            const double *v = desc + p;
            for (int d = 0; d < 3; ++d) {
                c->d[s][d] = v[d];
                c->d_e[s][d] = v[3 + d];
            }
            c->d_p[s] = v[6];
            p += 7;
        }
        for (int s = 0; s < c->nq; ++s) {
            const double *v = desc + p;
            for (int d = 0; d < 3; ++d) {
                c->qd[s][d] = v[d];
                c->q_e[s][d] = v[3 + d];
            }
            c->q_Q[s] = v[6];
            p += 7;
        }
        /* ASCIIReader.cpp:197-202: positive header moments override */
        for (int d = 0; d < 3; ++d) {
            double Iov = desc[p + d];
            c->I[d] = (Iov > 0.) ? Iov : I6[d];
            /* FullMolecule::updateMassInertia, FullMolecule.h:88-99 */
            c->invI[d] = (c->I[d] != 0.) ? 1. / c->I[d] : 0.;
        }
        p += 3;
        c->rotdof = rotdof;
        c->lj_base = lj_base;
        lj_base += c->nlj;
    }
    x->nlj_total = lj_base;
    return p;
}

/* Comp2Param::initialize LJ part, src/molecules/Comp2Param.cpp:11-110, laid
 * out as VectorizedCellProcessor's _eps_sig/_shift6 tables
 * (VectorizedCellProcessor.cpp:40-80): row/column = LJ lookup id. Same
 * component: site i's own shift6 (Comp2Param.cpp:27-38); different
 * components: always truncated-shifted at rcLJ (Comp2Param.cpp:79-105). */
static void build_lj_tables(ox_ctx *x, const double *mix) {
    int n = x->nlj_total;
    x->eps24 = (double *)calloc((size_t)n * n + 1, sizeof(double));
    x->sig2 = (double *)calloc((size_t)n * n + 1, sizeof(double));
    x->shift6 = (double *)calloc((size_t)n * n + 1, sizeof(double));
    int mixidx = 0;
    double xi_tab[OX_MAXCOMP][OX_MAXCOMP], eta_tab[OX_MAXCOMP][OX_MAXCOMP];
    for (int i = 0; i < x->ncomp; ++i)
        for (int j = i + 1; j < x->ncomp; ++j) {
            xi_tab[i][j] = mix ? mix[2 * mixidx] : 1.0;
            eta_tab[i][j] = mix ? mix[2 * mixidx + 1] : 1.0;
            ++mixidx;
        }
    for (int ci = 0; ci < x->ncomp; ++ci) {
        const ox_comp *a = &x->comp[ci];
        for (int si = 0; si < a->nlj; ++si) {
            for (int sj = 0; sj < a->nlj; ++sj) {
                double e24 = 24. * sqrt(a->lj_eps[si] * a->lj_eps[sj]);
                double s2 = .5 * (a->lj_sig[si] + a->lj_sig[sj]);
                s2 *= s2;
                int k = (a->lj_base + si) * n + (a->lj_base + sj);
                x->eps24[k] = e24;
                x->sig2[k] = s2;
                x->shift6[k] = a->lj_shift6[si];
            }
        }
        for (int cj = ci + 1; cj < x->ncomp; ++cj) {
            const ox_comp *b = &x->comp[cj];
            double xi = xi_tab[ci][cj], eta = eta_tab[ci][cj];
            for (int pass = 0; pass < 2; ++pass) {
                const ox_comp *p1 = pass ? b : a, *p2 = pass ? a : b;
                for (int s1 = 0; s1 < p1->nlj; ++s1)
                    for (int s2i = 0; s2i < p2->nlj; ++s2i) {
                        double e24 = 24. * (xi * sqrt(p1->lj_eps[s1] * p2->lj_eps[s2i]));
                        double s2 = 0.5 * (eta * (p1->lj_sig[s1] + p2->lj_sig[s2i]));
                        s2 *= s2;
                        double sr2 = s2 / (x->rcLJ * x->rcLJ);
                        double sr6 = sr2 * sr2 * sr2;
                        int k = (p1->lj_base + s1) * n + (p2->lj_base + s2i);
                        x->eps24[k] = e24;
                        x->sig2[k] = s2;
                        x->shift6[k] = e24 * (sr6 - sr6 * sr6);
                    }
            }
        }
    }
}

/* LinkedCells::LinkedCells (cell_mode 0, double rc; LinkedCells.cpp:38-106)
 * and LinkedCells::rebuild (cell_mode 1, float rc; LinkedCells.cpp:132-201).
 * Returns -1 where rebuild exits (no cell fits, LinkedCells.cpp:160-164). */
static int build_geometry(ox_ctx *x, int cell_mode) {
    for (int d = 0; d < 3; ++d) {
        double diff = x->bbmax[d] - x->bbmin[d];
        double rcc = cell_mode ? (double)(float)x->rc : x->rc;
        int nb = (int)floor(diff / rcc);
        if (nb == 0 && cell_mode) return -1;
        if (nb == 0) nb = 1;
        x->boxcells[d] = nb;
        x->cell_len[d] = diff / nb;
        x->cell_rlen[d] = nb / diff;
        x->halo_min[d] = x->bbmin[d] - x->cell_len[d];
        x->halo_max[d] = x->bbmax[d] + x->cell_len[d];
        x->cpd[d] = nb + 2;
    }
    return 0;
}

void *ox_create(int ncomp, const double *desc, int desc_len, const double *mix,
                double epsRF, double rc, double rcLJ, const double *box,
                int periodic, int cell_mode) {
    ox_ctx *x = (ox_ctx *)calloc(1, sizeof(ox_ctx));
    if (!x) return NULL;
    if (parse_components(x, ncomp, desc, desc_len) < 0) {
        free(x);
        return NULL;
    }
    x->epsRF = epsRF;
    x->rc = rc;
    x->rcLJ = rcLJ;
    x->nthreads = 1;
    /* VectorizedCellProcessor ctor, VectorizedCellProcessor.cpp:19-23 */
    x->epsRFInvrc3 = 2. * (epsRF - 1.) / ((rc * rc * rc) * (2. * epsRF + 1.));
    for (int d = 0; d < 3; ++d) {
        x->bbmin[d] = 0.0;
        x->bbmax[d] = box[d];
    }
    x->periodic = periodic;
    for (int d = 0; d < 3; ++d) {
        x->pdim[d] = periodic ? 1 : 0;
        x->wall[d] = 0;
        x->outflow[d] = 0;
    }
    build_lj_tables(x, mix);
    if (build_geometry(x, cell_mode) < 0) {
        ox_destroy(x);
        return NULL;
    }
    return x;
}

void ox_destroy(void *h) {
    ox_ctx *x = (ox_ctx *)h;
    if (!x) return;
    free(x->eps24);
    free(x->sig2);
    free(x->shift6);
    free(x);
}

int ox_nlj_total(void *h) { return ((ox_ctx *)h)->nlj_total; }

void ox_lj_tables(void *h, double *eps24, double *sig2, double *shift6) {
    ox_ctx *x = (ox_ctx *)h;
    size_t n = (size_t)x->nlj_total * x->nlj_total;
    memcpy(eps24, x->eps24, n * sizeof(double));
    memcpy(sig2, x->sig2, n * sizeof(double));
    memcpy(shift6, x->shift6, n * sizeof(double));
}

void ox_set_lj_tables(void *h, const double *eps24, const double *sig2, const double *shift6) {
    ox_ctx *x = (ox_ctx *)h;
    size_t n = (size_t)x->nlj_total * x->nlj_total;
    memcpy(x->eps24, eps24, n * sizeof(double));
    memcpy(x->sig2, sig2, n * sizeof(double));
    memcpy(x->shift6, shift6, n * sizeof(double));
}

void ox_component_props(void *h, double *mass, double *inertia, int *rotdof) {
    ox_ctx *x = (ox_ctx *)h;
    for (int c = 0; c < x->ncomp; ++c) {
        mass[c] = x->comp[c].m;
        for (int d = 0; d < 3; ++d) inertia[3 * c + d] = x->comp[c].I[d];
        rotdof[c] = x->comp[c].rotdof;
    }
}

void ox_cell_geometry(void *h, int *cpd, double *cl, double *hmin, double *hmax) {
    ox_ctx *x = (ox_ctx *)h;
    for (int d = 0; d < 3; ++d) {
        cpd[d] = x->cpd[d];
        cl[d] = x->cell_len[d];
        hmin[d] = x->halo_min[d];
        hmax[d] = x->halo_max[d];
    }
}

/* -------------------------------------------------------------- binning */

/* CellBorderAndFlagManager::getBorder, CellBorderAndFlagManager.h:134-152 */
static double border(const ox_ctx *x, int idx, int d) {
    if (idx == 0) return x->halo_min[d];
    if (idx == 1) return x->bbmin[d];
    if (idx == x->cpd[d] - 1) return x->bbmax[d];
    if (idx == x->cpd[d]) return x->halo_max[d];
    return idx * x->cell_len[d] + x->halo_min[d];
}

static long lin3(const ox_ctx *x, long i, long j, long k) {
    /* LinkedCells::cellIndexOf3DIndex, LinkedCells.cpp:925-927 */
    return (k * x->cpd[1] + j) * x->cpd[0] + i;
}

/* LinkedCells::getCellIndexOfPoint, LinkedCells.cpp:870-923 */
long ox_cell_index_of_point(void *h, const double *p) {
    const ox_ctx *x = (const ox_ctx *)h;
    int ci[3];
    double lp[3] = {p[0], p[1], p[2]};
    for (int d = 0; d < 3; ++d) {
        if (lp[d] <= x->halo_min[d])
            lp[d] += x->cell_len[d] * 0.5;
        else if (lp[d] >= x->halo_max[d])
            lp[d] -= x->cell_len[d] * 0.5;
        int v = (int)floor((lp[d] - x->bbmin[d]) * x->cell_rlen[d]) + 1;
        if (v < 0) v = 0;
        if (v > x->cpd[d] - 1) v = x->cpd[d] - 1;
        ci[d] = v;
    }
    int inside = 1;
    for (int d = 0; d < 3; ++d)
        if (!(border(x, ci[d], d) <= lp[d] && lp[d] < border(x, ci[d] + 1, d))) inside = 0;
    if (!inside) {
        for (int d = 0; d < 3; ++d) {
            if (lp[d] < border(x, ci[d], d))
                ci[d]--;
            else if (lp[d] >= border(x, ci[d] + 1, d))
                ci[d]++;
        }
    }
    return lin3(x, ci[0], ci[1], ci[2]);
}

void ox_cell_index_batch(void *h, int n, const double *r, long *out) {
    for (int i = 0; i < n; ++i) out[i] = ox_cell_index_of_point(h, r + 3 * i);
}

static int is_halo_cell(const ox_ctx *x, long c) {
    long i = c % x->cpd[0], j = (c / x->cpd[0]) % x->cpd[1], k = c / ((long)x->cpd[0] * x->cpd[1]);
    return i == 0 || j == 0 || k == 0 || i == x->cpd[0] - 1 || j == x->cpd[1] - 1 ||
           k == x->cpd[2] - 1;
}

/* C08 offsets and base-cell walk: C08BasedTraversals.h:46-149,
 * C08CellPairTraversal.h:45-75,196-228. Emits (c1,c2) as the traversal hands
 * them to the cell processor (halo-halo skipped, non-halo cell first). */
static const int c08_pairs[14][2][3] = {
    {{0, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {0, 1, 0}}, {{0, 1, 0}, {0, 0, 1}},
    {{0, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {0, 1, 1}}, {{1, 0, 0}, {0, 1, 1}},
    {{1, 0, 0}, {0, 1, 0}}, {{1, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {1, 0, 0}},
    {{0, 0, 0}, {1, 1, 0}}, {{1, 1, 0}, {0, 0, 1}}, {{0, 1, 0}, {1, 0, 1}},
    {{0, 0, 0}, {1, 0, 1}}, {{0, 0, 0}, {1, 1, 1}}};

typedef void (*pair_cb)(void *user, long c1, long c2);

static int walk_c08(const ox_ctx *x, pair_cb cb, void *user) {
    int npairs = 0;
    for (int col = 0; col < 8; ++col) {
        int s0 = col & 1, s1 = (col >> 1) & 1, s2 = (col >> 2) & 1;
        for (int k = s2; k < x->cpd[2] - 1; k += 2)
            for (int j = s1; j < x->cpd[1] - 1; j += 2)
                for (int i = s0; i < x->cpd[0] - 1; i += 2) {
                    for (int p = 0; p < 14; ++p) {
                        const int *a = c08_pairs[p][0], *b = c08_pairs[p][1];
                        long c1 = lin3(x, i + a[0], j + a[1], k + a[2]);
                        long c2 = lin3(x, i + b[0], j + b[1], k + b[2]);
                        int h1 = is_halo_cell(x, c1), h2 = is_halo_cell(x, c2);
                        if (h1 && h2) continue;
                        if (c1 != c2 && h1) {
                            long t = c1;
                            c1 = c2;
                            c2 = t;
                        }
                        cb(user, c1, c2);
                        ++npairs;
                    }
                }
    }
    return npairs;
}

typedef struct {
    int *out;
    int max, n;
} pair_sink;

static void sink_pair(void *u, long c1, long c2) {
    pair_sink *s = (pair_sink *)u;
    if (s->n < s->max) {
        s->out[2 * s->n] = (int)c1;
        s->out[2 * s->n + 1] = (int)c2;
    }
    s->n++;
}

int ox_c08_pairs(void *h, int *out, int max_pairs) {
    pair_sink s = {out, max_pairs, 0};
    walk_c08((const ox_ctx *)h, sink_pair, &s);
    return s.n;
}

/* DomainDecompBase::handleDomainLeavingParticles, DomainDecompBase.cpp:199-249:
 * per dimension x,y,z, molecules below the box get +L (clamped below bbmax),
 * molecules at or above bbmax get -L (clamped to bbmin). */
void ox_wrap(void *h, int n, double *r) {
    const ox_ctx *x = (const ox_ctx *)h;
    for (int d = 0; d < 3; ++d) {
        if (!x->pdim[d]) continue;
        double L = x->bbmax[d] - x->bbmin[d];
        for (int i = 0; i < n; ++i) {
            double v = r[3 * i + d];
            if (v < x->bbmin[d]) {
                v = v + L;
                if (v >= x->bbmax[d]) v = nextafter(x->bbmax[d], -INFINITY);
            } else if (v >= x->bbmax[d]) {
                v = v - L;
                if (v <= x->bbmin[d]) v = x->bbmin[d];
            }
            r[3 * i + d] = v;
        }
    }
}

/* ------------------------------------------------------------ pair bodies */

static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Canonical site-pair bodies. Each returns the force on the FIRST site
 * (the second gets -f), torques on oriented sites, and the energy term.
 * kinds: 0 LJ, 1 C-C, 2 C-D, 3 D-D, 4 C-Q, 5 D-Q, 6 Q-Q. */
enum { K_LJ = 0, K_CC, K_CD, K_DD, K_CQ, K_DQ, K_QQ };

typedef struct {
    double f[3], M1[3], M2[3], u, rf;
} pair_out;

/* _loopBodyLJ, VectorizedCellProcessor.cpp:171-224 (u = upot6 term) */
static void body_lj(const double *d, double eps24, double sig2, double shift6,
                    pair_out *o) {
    double r2 = dot3(d, d);
    double r2inv = 1.0 / r2;
    double lj2 = sig2 * r2inv, lj4 = lj2 * lj2, lj6 = lj4 * lj2, lj12 = lj6 * lj6;
    double lj12m6 = lj12 - lj6;
    double scale = (eps24 * r2inv) * (lj12 + lj12m6);
    for (int k = 0; k < 3; ++k) o->f[k] = d[k] * scale;
    o->u = eps24 * lj12m6 + shift6;
}

/* _loopBodyCharge, VectorizedCellProcessor.cpp:227-277 (NOVEC: sqrt of 1/r2) */
static void body_cc(const double *d, double q1, double q2, pair_out *o) {
    double r2inv = 1.0 / dot3(d, d);
    double rinv = sqrt(r2inv);
    double upot = (q1 * q2) * rinv;
    double fac = upot * r2inv;
    for (int k = 0; k < 3; ++k) o->f[k] = d[k] * fac;
    o->u = upot;
}

/* _loopBodyChargeDipole, VectorizedCellProcessor.cpp:279-347: first = charge,
 * second = dipole (axis e, moment p); M2 = torque on the dipole. */
static void body_cd(const double *d, double q, const double *e, double p,
                    pair_out *o) {
    double r2inv = 1.0 / dot3(d, d);
    double rinv = sqrt(r2inv);
    double r3inv = r2inv * rinv;
    double re = dot3(d, e);
    double qpr3 = (q * p) * r3inv;
    double fac = r2inv * 3.0 * re;
    for (int k = 0; k < 3; ++k) o->f[k] = qpr3 * (e[k] - d[k] * fac);
    o->u = -(qpr3 * re);
    o->M2[0] = qpr3 * (e[1] * d[2] - e[2] * d[1]);
    o->M2[1] = qpr3 * (e[2] * d[0] - e[0] * d[2]);
    o->M2[2] = qpr3 * (e[0] * d[1] - e[1] * d[0]);
}

/* _loopBodyDipole, VectorizedCellProcessor.cpp:349-442 (with reaction field) */
static void body_dd(const double *d, const double *e1, double p1, const double *e2,
                    double p2, double epsRFInvrc3, pair_out *o) {
    double r2inv = 1.0 / dot3(d, d);
    double rinv = sqrt(r2inv);
    double r2three = 3.0 * r2inv;
    double p1p2 = p1 * p2;
    double rffac = p1p2 * epsRFInvrc3;
    double p1p2r3 = p1p2 * rinv * r2inv;
    double p1p2r5 = p1p2r3 * r2three;
    double e1e2 = dot3(e1, e2), re1 = dot3(d, e1), re2 = dot3(d, e2);
    double re1t = re1 * r2three, re2t = re2 * r2three;
    double re1re2 = r2inv * re1 * re2;
    double g = e1e2 - 5.0 * re1re2;
    for (int k = 0; k < 3; ++k) o->f[k] = p1p2r5 * (d[k] * g + e1[k] * re2 + e2[k] * re1);
    o->u = p1p2r3 * (e1e2 - 3.0 * re1re2);
    o->rf = rffac * e1e2;
    double c12[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                     e1[0] * e2[1] - e1[1] * e2[0]};
    double e1xd[3] = {e1[1] * d[2] - e1[2] * d[1], e1[2] * d[0] - e1[0] * d[2],
                      e1[0] * d[1] - e1[1] * d[0]};
    double e2xd[3] = {e2[1] * d[2] - e2[2] * d[1], e2[2] * d[0] - e2[0] * d[2],
                      e2[0] * d[1] - e2[1] * d[0]};
    for (int k = 0; k < 3; ++k) {
        o->M1[k] = p1p2r3 * (re2t * e1xd[k] - c12[k]) + rffac * c12[k];
        o->M2[k] = p1p2r3 * (re1t * e2xd[k] + c12[k]) - rffac * c12[k];
    }
}

/* _loopBodyChargeQuadrupole, VectorizedCellProcessor.cpp:444-524: first =
 * charge, second = linear quadrupole (axis e, moment m); M2 on the quadrupole. */
static void body_cq(const double *d, double q, const double *e, double m,
                    pair_out *o) {
    double r2inv = 1.0 / dot3(d, d);
    double rinv = sqrt(r2inv);
    double qQ05 = 0.5 * q * m;
    double costj = dot3(e, d) * rinv;
    double qQr3 = qQ05 * rinv * r2inv;
    double upot = qQr3 * (3.0 * costj * costj - 1.0);
    double mpR = 3.0 * upot * r2inv;
    double pT = 6.0 * costj * qQr3 * rinv;
    double fac = costj * pT * rinv + mpR;
    for (int k = 0; k < 3; ++k) o->f[k] = fac * d[k] - pT * e[k];
    o->u = upot;
    o->M2[0] = pT * (e[2] * d[1] - e[1] * d[2]);
    o->M2[1] = pT * (e[0] * d[2] - e[2] * d[0]);
    o->M2[2] = pT * (e[1] * d[0] - e[0] * d[1]);
}

/* _loopBodyDipoleQuadrupole, VectorizedCellProcessor.cpp:526-648: first =
 * dipole (ei, p), second = quadrupole (ej, m). */
static void body_dq(const double *d, const double *ei, double p, const double *ej,
                    double m, pair_out *o) {
    double r2inv = 1.0 / dot3(d, d);
    double rinv = sqrt(r2inv);
    double myqfac = 1.5 * p * m * r2inv * r2inv;
    double costi = dot3(ei, d) * rinv, costj = dot3(ej, d) * rinv;
    double cos2tj = costj * costj;
    double cosgij = dot3(ei, ej);
    double c5 = 5.0 * cos2tj - 1.0;
    double upot = myqfac * (2.0 * costj * cosgij - costi * c5);
    double myqfacr = myqfac * rinv;
    double mpR = 4.0 * upot * r2inv;
    double mpTi = myqfacr * c5;
    double part1 = 5.0 * (costi * costj) - cosgij;
    double mpTj = myqfacr * 2.0 * part1;
    double pG = myqfac * (2.0 * costj);
    double part3 = costi * mpTi + costj * mpTj;
    double fac = mpR - part3 * rinv;
    for (int k = 0; k < 3; ++k) o->f[k] = fac * d[k] + mpTi * ei[k] + mpTj * ej[k];
    o->u = upot;
    double c12[3] = {ei[1] * ej[2] - ei[2] * ej[1], ei[2] * ej[0] - ei[0] * ej[2],
                     ei[0] * ej[1] - ei[1] * ej[0]};
    double eixd[3] = {ei[1] * d[2] - ei[2] * d[1], ei[2] * d[0] - ei[0] * d[2],
                      ei[0] * d[1] - ei[1] * d[0]};
    double ejxd[3] = {ej[1] * d[2] - ej[2] * d[1], ej[2] * d[0] - ej[0] * d[2],
                      ej[0] * d[1] - ej[1] * d[0]};
    for (int k = 0; k < 3; ++k) {
        o->M1[k] = mpTi * eixd[k] - pG * c12[k];
        o->M2[k] = mpTj * ejxd[k] + pG * c12[k];
    }
}

/* _loopBodyQuadrupole, VectorizedCellProcessor.cpp:650-792 */
static void body_qq(const double *d, const double *ei, double mi, const double *ej,
                    double mj, pair_out *o) {
    double r2inv = 1.0 / dot3(d, d);
    double rinv = sqrt(r2inv);
    double qfac = 0.75 * rinv;
    qfac = qfac * (mi * mj);
    qfac = qfac * (r2inv * r2inv);
    double costi = dot3(ei, d) * rinv, costj = dot3(ej, d) * rinv;
    double c2i = costi * costi, c2j = costj * costj;
    double cosgij = dot3(ei, ej);
    double term = cosgij - 5.0 * (costi * costj);
    double upot = 5.0 * (c2i + c2j) + 15.0 * c2i * c2j;
    upot = (1.0 + 2.0 * term * term) - upot;
    upot = qfac * upot;
    double mpR = 5.0 * upot * r2inv;
    double part1 = qfac * 10.0 * rinv;
    double part2 = 2.0 * term;
    double mpTi = part1 * (costi + (part2 * costj + 3.0 * costi * c2j));
    double mpTj = part1 * (costj + (part2 * costi + 3.0 * costj * c2i));
    double pG = qfac * 4.0 * term;
    double fac = mpR - (mpTi * costi + mpTj * costj) * rinv;
    for (int k = 0; k < 3; ++k) o->f[k] = fac * d[k] + mpTi * ei[k] + mpTj * ej[k];
    o->u = upot;
    double c12[3] = {ei[1] * ej[2] - ei[2] * ej[1], ei[2] * ej[0] - ei[0] * ej[2],
                     ei[0] * ej[1] - ei[1] * ej[0]};
    double eixd[3] = {ei[1] * d[2] - ei[2] * d[1], ei[2] * d[0] - ei[0] * d[2],
                      ei[0] * d[1] - ei[1] * d[0]};
    double ejxd[3] = {ej[1] * d[2] - ej[2] * d[1], ej[2] * d[0] - ej[0] * d[2],
                      ej[0] * d[1] - ej[1] * d[0]};
    for (int k = 0; k < 3; ++k) {
        o->M1[k] = mpTi * eixd[k] - pG * c12[k];
        o->M2[k] = mpTj * ejxd[k] + pG * c12[k];
    }
}

void ox_site_pair(int kind, const double *ri, const double *ei, double pi,
                  const double *rj, const double *ej, double pj, double eps24,
                  double sig2, double shift6, double epsRFInvrc3, double *out) {
    pair_out o;
    memset(&o, 0, sizeof(o));
    double d[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
    switch (kind) {
    case K_LJ: body_lj(d, eps24, sig2, shift6, &o); break;
    case K_CC: body_cc(d, pi, pj, &o); break;
    case K_CD: body_cd(d, pi, ej, pj, &o); break;
    case K_DD: body_dd(d, ei, pi, ej, pj, epsRFInvrc3, &o); break;
    case K_CQ: body_cq(d, pi, ej, pj, &o); break;
    case K_DQ: body_dq(d, ei, pi, ej, pj, &o); break;
    case K_QQ: body_qq(d, ei, pi, ej, pj, &o); break;
    }
    for (int k = 0; k < 3; ++k) {
        out[k] = o.f[k];
        out[3 + k] = o.M1[k];
        out[6 + k] = o.M2[k];
    }
    out[9] = o.u;
    out[10] = o.rf;
}

/* ---------------------------------------------- molecule SoA (per molecule) */

typedef struct {
    double r[3];           /* molecule position */
    int cid;
    int orig;              /* index of the inner molecule this is (a copy of) */
    double q[4];           /* normalised quaternion */
    double s[4][OX_MAXSITE][3]; /* site positions per type LJ,C,D,Q */
    double e[4][OX_MAXSITE][3]; /* orientation of D,Q sites */
    double f[4][OX_MAXSITE][3], V[4][OX_MAXSITE][3], Ms[4][OX_MAXSITE][3];
} ox_mol;

/* Quaternion::rotate, src/molecules/Quaternion.cpp:43-64 */
static void qrot(const double *q, const double *d, double *out) {
    out[0] = out[1] = out[2] = 0.0;
    if (d[0] == 0. && d[1] == 0. && d[2] == 0.) return;
    double ww = q[0] * q[0], xx = q[1] * q[1], yy = q[2] * q[2], zz = q[3] * q[3];
    double wx = q[0] * q[1], wy = q[0] * q[2], wz = q[0] * q[3];
    double xy = q[1] * q[2], xz = q[1] * q[3], yz = q[2] * q[3];
    out[0] = (ww + xx - yy - zz) * d[0] + 2. * (xy - wz) * d[1] + 2. * (wy + xz) * d[2];
    out[1] = 2. * (wz + xy) * d[0] + (ww - xx + yy - zz) * d[1] + 2. * (yz - wx) * d[2];
    out[2] = 2. * (xz - wy) * d[0] + 2. * (wx + yz) * d[1] + (ww - xx - yy + zz) * d[2];
}

/* Quaternion::rotateinv, Quaternion.cpp:66-86 */
static void qrotinv(const double *q, const double *d, double *out) {
    out[0] = out[1] = out[2] = 0.0;
    if (d[0] == 0. && d[1] == 0. && d[2] == 0.) return;
    double ww = q[0] * q[0], xx = q[1] * q[1], yy = q[2] * q[2], zz = q[3] * q[3];
    double wx = q[0] * q[1], wy = q[0] * q[2], wz = q[0] * q[3];
    double xy = q[1] * q[2], xz = q[1] * q[3], yz = q[2] * q[3];
    out[0] = (ww + xx - yy - zz) * d[0] + 2. * (xy + wz) * d[1] + 2. * (xz - wy) * d[2];
    out[1] = 2. * (xy - wz) * d[0] + (ww - xx + yy - zz) * d[1] + 2. * (yz + wx) * d[2];
    out[2] = 2. * (xz + wy) * d[0] + 2. * (yz - wx) * d[1] + (ww - xx - yy + zz) * d[2];
}

/* Quaternion::normalize (Quaternion.h:40-42): scale by 1/sqrt(|q|^2) */
static void qnormalize(double *q) {
    double s = 1. / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] *= s;
}

/* FullMolecule::setupSoACache, FullMolecule.cpp:711-767, and clearFM :476-521 */
static void setup_mol(const ox_ctx *x, ox_mol *m, const double *r, const double *q,
                      int cid, int orig) {
    const ox_comp *c = &x->comp[cid];
    memset(m, 0, sizeof(*m));
    for (int k = 0; k < 3; ++k) m->r[k] = r[k];
    for (int k = 0; k < 4; ++k) m->q[k] = q[k];
    qnormalize(m->q);
    m->cid = cid;
    m->orig = orig;
    double t[3];
    for (int s = 0; s < c->nlj; ++s) {
        qrot(m->q, c->lj[s], t);
        for (int k = 0; k < 3; ++k) m->s[0][s][k] = t[k] + r[k];
    }
    for (int s = 0; s < c->nc; ++s) {
        qrot(m->q, c->c[s], t);
        for (int k = 0; k < 3; ++k) m->s[1][s][k] = t[k] + r[k];
    }
    for (int s = 0; s < c->nd; ++s) {
        qrot(m->q, c->d[s], t);
        for (int k = 0; k < 3; ++k) m->s[2][s][k] = t[k] + r[k];
        qrot(m->q, c->d_e[s], m->e[2][s]);
    }
    for (int s = 0; s < c->nq; ++s) {
        qrot(m->q, c->qd[s], t);
        for (int k = 0; k < 3; ++k) m->s[3][s][k] = t[k] + r[k];
        qrot(m->q, c->q_e[s], m->e[3][s]);
    }
}

/* One site pair of a molecule pair: apply force with Newton 3, virial on both
 * sites, torques to oriented sites and (if macro) the macroscopic sums.
 * first/second are given in the canonical body order. */
static void apply_pair(const ox_ctx *x, ox_mol *A, int ta, int sa, ox_mol *B, int tb,
                       int sb, int kind, double eps24, double sig2, double shift6,
                       int macro, double *mac, long long *cnt) {
    pair_out o;
    memset(&o, 0, sizeof(o));
    const double *ra = A->s[ta][sa], *rb = B->s[tb][sb];
    double d[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
    const ox_comp *ca = &x->comp[A->cid], *cb = &x->comp[B->cid];
    switch (kind) {
    case K_LJ: body_lj(d, eps24, sig2, shift6, &o); break;
    case K_CC: body_cc(d, ca->c_q[sa], cb->c_q[sb], &o); break;
    case K_CD: body_cd(d, ca->c_q[sa], B->e[2][sb], cb->d_p[sb], &o); break;
    case K_DD:
        body_dd(d, A->e[2][sa], ca->d_p[sa], B->e[2][sb], cb->d_p[sb], x->epsRFInvrc3, &o);
        break;
    case K_CQ: body_cq(d, ca->c_q[sa], B->e[3][sb], cb->q_Q[sb], &o); break;
    case K_DQ: body_dq(d, A->e[2][sa], ca->d_p[sa], B->e[3][sb], cb->q_Q[sb], &o); break;
    case K_QQ: body_qq(d, A->e[3][sa], ca->q_Q[sa], B->e[3][sb], cb->q_Q[sb], &o); break;
    }
    double V[3];
    for (int k = 0; k < 3; ++k) {
        V[k] = (A->r[k] - B->r[k]) * o.f[k];
        A->f[ta][sa][k] += o.f[k];
        B->f[tb][sb][k] -= o.f[k];
        A->V[ta][sa][k] += V[k];
        B->V[tb][sb][k] += V[k];
        A->Ms[ta][sa][k] += o.M1[k];
        B->Ms[tb][sb][k] += o.M2[k];
    }
    if (cnt && macro) cnt[1 + kind] += 1;
    if (macro) {
        if (kind == K_LJ)
            mac[OX_UPOT6LJ] += o.u;
        else
            mac[OX_UPOTXPOLES] += o.u;
        mac[OX_VIRIAL] += V[0] + V[1] + V[2];
        if (kind == K_DD) mac[OX_MYRF] -= o.rf;
    }
}

/* One molecule pair. The cutoff is tested on molecule centres
 * (calcDistLookup, SIMD_VectorizedCellProcessorHelpers.h:427-472): LJ sites
 * against rcLJ^2, electrostatic sites against rc^2, strictly less. */
static void mol_pair(const ox_ctx *x, ox_mol *A, ox_mol *B, int macro, double *mac,
                     long long *cnt) {
    double dm[3] = {A->r[0] - B->r[0], A->r[1] - B->r[1], A->r[2] - B->r[2]};
    double r2 = dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2];
    if (cnt && macro) cnt[OX_CNT_MOLPAIRS] += 1;
    const ox_comp *ca = &x->comp[A->cid], *cb = &x->comp[B->cid];
    if (r2 < x->rcLJ * x->rcLJ) {
        int n = x->nlj_total;
        for (int i = 0; i < ca->nlj; ++i)
            for (int j = 0; j < cb->nlj; ++j) {
                int k = (ca->lj_base + i) * n + (cb->lj_base + j);
                apply_pair(x, A, 0, i, B, 0, j, K_LJ, x->eps24[k], x->sig2[k], x->shift6[k],
                           macro, mac, cnt);
            }
    }
    if (!(r2 < x->rc * x->rc)) return;
    /* charge, dipole and quadrupole loops of _calculatePairs
     * (VectorizedCellProcessor.cpp:1181-2720); every ordered type pair once,
     * in canonical body order with roles swapped where the reference swaps */
    for (int i = 0; i < ca->nc; ++i) {
        for (int j = 0; j < cb->nc; ++j) apply_pair(x, A, 1, i, B, 1, j, K_CC, 0, 0, 0, macro, mac, cnt);
        for (int j = 0; j < cb->nd; ++j) apply_pair(x, A, 1, i, B, 2, j, K_CD, 0, 0, 0, macro, mac, cnt);
        for (int j = 0; j < cb->nq; ++j) apply_pair(x, A, 1, i, B, 3, j, K_CQ, 0, 0, 0, macro, mac, cnt);
    }
    for (int i = 0; i < ca->nd; ++i) {
        for (int j = 0; j < cb->nc; ++j) apply_pair(x, B, 1, j, A, 2, i, K_CD, 0, 0, 0, macro, mac, cnt);
        for (int j = 0; j < cb->nd; ++j) apply_pair(x, A, 2, i, B, 2, j, K_DD, 0, 0, 0, macro, mac, cnt);
        for (int j = 0; j < cb->nq; ++j) apply_pair(x, A, 2, i, B, 3, j, K_DQ, 0, 0, 0, macro, mac, cnt);
    }
    for (int i = 0; i < ca->nq; ++i) {
        for (int j = 0; j < cb->nc; ++j) apply_pair(x, B, 1, j, A, 3, i, K_CQ, 0, 0, 0, macro, mac, cnt);
        for (int j = 0; j < cb->nd; ++j) apply_pair(x, B, 2, j, A, 3, i, K_DQ, 0, 0, 0, macro, mac, cnt);
        for (int j = 0; j < cb->nq; ++j) apply_pair(x, A, 3, i, B, 3, j, K_QQ, 0, 0, 0, macro, mac, cnt);
    }
}

/* FullMolecule::calcFM, FullMolecule.cpp:548-631 (site offsets from
 * computeXX_d = rotate(q, body offset), FullMolecule.h:213-239) */
static void calc_fm(const ox_ctx *x, const ox_mol *m, double *F, double *M, double *Vi) {
    const ox_comp *c = &x->comp[m->cid];
    double tM[3] = {0, 0, 0}, tV[3] = {0, 0, 0};
    F[0] = F[1] = F[2] = 0;
    M[0] = M[1] = M[2] = 0;
    int ns[4] = {c->nlj, c->nc, c->nd, c->nq};
    for (int t = 0; t < 4; ++t) {
        for (int s = 0; s < ns[t]; ++s) {
            const double *body = t == 0 ? c->lj[s] : t == 1 ? c->c[s] : t == 2 ? c->d[s] : c->qd[s];
            double ds[3];
            qrot(m->q, body, ds);
            const double *fs = m->f[t][s];
            for (int k = 0; k < 3; ++k) F[k] += fs[k];
            M[0] += ds[1] * fs[2] - ds[2] * fs[1];
            M[1] += ds[2] * fs[0] - ds[0] * fs[2];
            M[2] += ds[0] * fs[1] - ds[1] * fs[0];
            for (int k = 0; k < 3; ++k) tV[k] += m->V[t][s][k];
            if (t >= 2)
                for (int k = 0; k < 3; ++k) tM[k] += m->Ms[t][s][k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        Vi[k] = 0.5 * tV[k];
        M[k] += tM[k];
    }
}

/* ---------------------------------------------------------------- traversal */

typedef struct {
    ox_ctx *x;
    ox_mol *mols;
    int *cell_start, *cell_list;
    double *mac;
    long long *cnt;
} trav_state;

/* VectorizedCellProcessor::processCell / processCellPair,
 * VectorizedCellProcessor.cpp:2732-2819 (sumAll = false path) */
static void trav_cb(void *u, long c1, long c2) {
    trav_state *t = (trav_state *)u;
    int b1 = t->cell_start[c1], e1 = t->cell_start[c1 + 1];
    if (c1 == c2) {
        if (is_halo_cell(t->x, c1) || e1 - b1 < 2) return;
        for (int a = b1; a < e1; ++a)
            for (int b = a + 1; b < e1; ++b) {
                ox_mol *A = &t->mols[t->cell_list[a]], *B = &t->mols[t->cell_list[b]];
                double dm[3] = {A->r[0] - B->r[0], A->r[1] - B->r[1], A->r[2] - B->r[2]};
                /* SingleCellPolicy_::GetForceMask: m_r2 != 0 */
                if (dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2] == 0.0) {
                    if (t->cnt) t->cnt[OX_CNT_MOLPAIRS] += 1;
                    continue;
                }
                mol_pair(t->x, A, B, 1, t->mac, t->cnt);
            }
        return;
    }
    int b2 = t->cell_start[c2], e2 = t->cell_start[c2 + 1];
    if (e1 == b1 || e2 == b2) return;
    int h1 = is_halo_cell(t->x, c1), h2 = is_halo_cell(t->x, c2);
    if (h1 && h2) return;
    int macro = (!h1 && !h2) || (c1 < c2);
    for (int a = b1; a < e1; ++a)
        for (int b = b2; b < e2; ++b)
            mol_pair(t->x, &t->mols[t->cell_list[a]], &t->mols[t->cell_list[b]], macro, t->mac,
                     t->cnt);
}

/* DomainDecompBase::populateHaloLayerWithCopies, DomainDecompBase.cpp:317-376:
 * for x, then y, then z, copy every molecule (inner and earlier copies)
 * within rc of a face to the opposite side. */
static int make_halo(const ox_ctx *x, ox_mol **pm, int n, int *cap) {
    int total = n;
    for (int d = 0; d < 3; ++d) {
        /* non-periodic dimensions keep no halo copies
         * (BoundaryHandler::removeNonPeriodicHalos, BoundaryHandler.cpp:136-170) */
        if (!x->pdim[d]) continue;
        double L = x->bbmax[d] - x->bbmin[d];
        for (int dir = 0; dir < 2; ++dir) {
            double shift = dir == 0 ? L : -L;
            double lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = x->bbmin[k] - x->rc;
                hi[k] = x->bbmax[k] + x->rc;
            }
            if (dir == 0) {
                lo[d] = x->bbmin[d];
                hi[d] = x->bbmin[d] + x->rc;
            } else {
                lo[d] = x->bbmax[d] - x->rc;
                hi[d] = x->bbmax[d];
            }
            int cur = total;
            for (int i = 0; i < cur; ++i) {
                ox_mol *m = &(*pm)[i];
                int in = 1;
                for (int k = 0; k < 3; ++k)
                    if (!(lo[k] <= m->r[k] && m->r[k] < hi[k])) in = 0;
                if (!in) continue;
                if (total >= *cap) {
                    *cap *= 2;
                    *pm = (ox_mol *)realloc(*pm, sizeof(ox_mol) * (size_t)(*cap));
                    m = &(*pm)[i];
                }
                double r[3] = {m->r[0], m->r[1], m->r[2]};
                r[d] = r[d] + shift;
                if (shift < 0) {
                    if (r[d] >= x->bbmin[d]) r[d] = nextafter(x->bbmin[d], -INFINITY);
                } else {
                    if (r[d] < x->bbmax[d]) r[d] = nextafter(x->bbmax[d], INFINITY);
                }
                ox_mol tmp;
                setup_mol(x, &tmp, r, m->q, m->cid, m->orig);
                (*pm)[total++] = tmp;
            }
        }
    }
    return total;
}

/* The C08 walk with the colours in order and the base cells of one colour
 * shared among OpenMP threads, as ls1's C08CellPairTraversal does
 * (C08CellPairTraversal.h:45-75): cells of one colour's base blocks are
 * disjoint, so every molecule's sums keep their order; the macroscopic sums
 * and counters are per thread, added in thread order at the end. */
static void walk_c08_threads(const ox_ctx *x, trav_state *ts) {
    const int nt = x->nthreads;
    double *mac = (double *)calloc((size_t)nt * 4, sizeof(double));
    long long *cnt = (long long *)calloc((size_t)nt * OX_NCNT, sizeof(long long));
    for (int col = 0; col < 8; ++col) {
        const int s0 = col & 1, s1 = (col >> 1) & 1, s2 = (col >> 2) & 1;
        const int nx = (x->cpd[0] - 1 - s0 + 1) / 2, ny = (x->cpd[1] - 1 - s1 + 1) / 2,
                  nz = (x->cpd[2] - 1 - s2 + 1) / 2;
        const long nb = (long)nx * ny * nz;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
        for (long b = 0; b < nb; ++b) {
            const int tid = omp_get_thread_num();
            trav_state t = *ts;
            t.mac = mac + 4 * tid;
            t.cnt = ts->cnt ? cnt + OX_NCNT * tid : NULL;
            const int i = s0 + 2 * (int)(b % nx), j = s1 + 2 * (int)((b / nx) % ny),
                      k = s2 + 2 * (int)(b / ((long)nx * ny));
            for (int p = 0; p < 14; ++p) {
                const int *a = c08_pairs[p][0], *bb = c08_pairs[p][1];
                long c1 = lin3(x, i + a[0], j + a[1], k + a[2]);
                long c2 = lin3(x, i + bb[0], j + bb[1], k + bb[2]);
                int h1 = is_halo_cell(x, c1), h2 = is_halo_cell(x, c2);
                if (h1 && h2) continue;
                if (c1 != c2 && h1) {
                    long tt = c1;
                    c1 = c2;
                    c2 = tt;
                }
                trav_cb(&t, c1, c2);
            }
        }
    }
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 4; ++k) ts->mac[k] += mac[4 * t + k];
        if (ts->cnt)
            for (int k = 0; k < OX_NCNT; ++k) ts->cnt[k] += cnt[OX_NCNT * t + k];
    }
    free(mac);
    free(cnt);
}

void ox_set_threads(void *h, int nthreads) {
    ((ox_ctx *)h)->nthreads = nthreads > 1 ? nthreads : 1;
}

int ox_forces(void *h, int n, const double *r, double *q, const int *cid, double *F,
              double *M, double *Vi, double *macro, long long *counters) {
    ox_ctx *x = (ox_ctx *)h;
    int cap = n + 16 + n / 2;
    ox_mol *mols = (ox_mol *)malloc(sizeof(ox_mol) * (size_t)cap);
    if (!mols) return -1;
#pragma omp parallel for num_threads(x->nthreads) schedule(static) if (x->nthreads > 1)
    for (int i = 0; i < n; ++i) {
        setup_mol(x, &mols[i], r + 3 * i, q + 4 * i, cid[i], i);
        for (int k = 0; k < 4; ++k) q[4 * i + k] = mols[i].q[k];
    }
    int total = n;
    if (x->periodic) total = make_halo(x, &mols, n, &cap);
    long ncell = (long)x->cpd[0] * x->cpd[1] * x->cpd[2];
    int *cs = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *cl = (int *)malloc(sizeof(int) * (size_t)(total + 1));
    long *ci = (long *)malloc(sizeof(long) * (size_t)(total + 1));
    for (int i = 0; i < total; ++i) {
        ci[i] = ox_cell_index_of_point(x, mols[i].r);
        cs[ci[i] + 1]++;
    }
    for (long c = 0; c < ncell; ++c) cs[c + 1] += cs[c];
    int *fill = (int *)calloc((size_t)ncell, sizeof(int));
    for (int i = 0; i < total; ++i) cl[cs[ci[i]] + fill[ci[i]]++] = i;
    double mac[4] = {0, 0, 0, 0};
    long long cnt[OX_NCNT];
    memset(cnt, 0, sizeof(cnt));
    trav_state ts = {x, mols, cs, cl, mac, cnt};
    if (x->nthreads > 1)
        walk_c08_threads(x, &ts);
    else
        walk_c08(x, trav_cb, &ts);
#pragma omp parallel for num_threads(x->nthreads) schedule(static) if (x->nthreads > 1)
    for (int i = 0; i < n; ++i) calc_fm(x, &mols[i], F + 3 * i, M + 3 * i, Vi + 3 * i);
    for (int k = 0; k < 4; ++k) macro[k] = mac[k];
    if (counters)
        for (int k = 0; k < OX_NCNT; ++k) counters[k] = cnt[k];
    free(fill);
    free(ci);
    free(cl);
    free(cs);
    free(mols);
    return 0;
}

/* Molecule-centre distance checks of the reference's traversal
 * (FlopCounter::processCell / processCellPair, FlopCounter.cpp:206-305: n (n-1)
 * / 2 per cell, n1 n2 per cell pair, over the C08 walk with the halo copies):
 * out[0] = every check the reference makes, including the second visit of a
 * boundary-halo pair from the other periodic side (processed without
 * macroscopic sums), out[1] = the checks of the pairs processed with them
 * (each unordered molecule pair once, the basis of the flop count here). No
 * forces are evaluated. */
typedef struct {
    const ox_ctx *x;
    const int *cell_start;
    long long all, macro;
} count_state;

static void count_cb(void *u, long c1, long c2) {
    count_state *t = (count_state *)u;
    const long long n1 = t->cell_start[c1 + 1] - t->cell_start[c1];
    if (c1 == c2) {
        if (is_halo_cell(t->x, c1) || n1 < 2) return;
        t->all += n1 * (n1 - 1) / 2;
        t->macro += n1 * (n1 - 1) / 2;
        return;
    }
    const long long n2 = t->cell_start[c2 + 1] - t->cell_start[c2];
    if (n1 == 0 || n2 == 0) return;
    const int h1 = is_halo_cell(t->x, c1), h2 = is_halo_cell(t->x, c2);
    if (h1 && h2) return;
    t->all += n1 * n2;
    if ((!h1 && !h2) || (c1 < c2)) t->macro += n1 * n2;
}

int ox_count_checks(void *h, int n, const double *r, const double *q, const int *cid,
                    long long *out) {
    ox_ctx *x = (ox_ctx *)h;
    int cap = n + 16 + n / 2;
    ox_mol *mols = (ox_mol *)malloc(sizeof(ox_mol) * (size_t)cap);
    if (!mols) return -1;
    for (int i = 0; i < n; ++i) setup_mol(x, &mols[i], r + 3 * i, q + 4 * i, cid[i], i);
    int total = n;
    if (x->periodic) total = make_halo(x, &mols, n, &cap);
    long ncell = (long)x->cpd[0] * x->cpd[1] * x->cpd[2];
    int *cs = (int *)calloc((size_t)ncell + 1, sizeof(int));
    for (int i = 0; i < total; ++i) cs[ox_cell_index_of_point(x, mols[i].r) + 1]++;
    for (long c = 0; c < ncell; ++c) cs[c + 1] += cs[c];
    count_state st = {x, cs, 0, 0};
    walk_c08(x, count_cb, &st);
    out[0] = st.all;
    out[1] = st.macro;
    free(cs);
    free(mols);
    return 0;
}

/* Independent check of the enumeration: every unordered molecule pair once,
 * minimum image per dimension (requires box > 2 rc in periodic dimensions). */
int ox_forces_brute(void *h, int n, const double *r, double *q, const int *cid, double *F,
                    double *M, double *Vi, double *macro) {
    ox_ctx *x = (ox_ctx *)h;
    ox_mol *mols = (ox_mol *)malloc(sizeof(ox_mol) * (size_t)(n + 1));
    ox_mol img;
    double mac[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        setup_mol(x, &mols[i], r + 3 * i, q + 4 * i, cid[i], i);
        for (int k = 0; k < 4; ++k) q[4 * i + k] = mols[i].q[k];
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double s[3] = {0, 0, 0};
            if (x->periodic)
                for (int d = 0; d < 3; ++d) {
                    if (!x->pdim[d]) continue;
                    double L = x->bbmax[d] - x->bbmin[d];
                    double dd = mols[i].r[d] - mols[j].r[d];
                    if (dd > 0.5 * L) s[d] = L;
                    else if (dd < -0.5 * L) s[d] = -L;
                }
            if (s[0] == 0 && s[1] == 0 && s[2] == 0) {
                mol_pair(x, &mols[i], &mols[j], 1, mac, NULL);
            } else {
                double rr[3] = {mols[j].r[0] + s[0], mols[j].r[1] + s[1], mols[j].r[2] + s[2]};
                setup_mol(x, &img, rr, mols[j].q, mols[j].cid, j);
                mol_pair(x, &mols[i], &img, 1, mac, NULL);
                for (int t = 0; t < 4; ++t)
                    for (int si = 0; si < OX_MAXSITE; ++si)
                        for (int k = 0; k < 3; ++k) {
                            mols[j].f[t][si][k] += img.f[t][si][k];
                            mols[j].V[t][si][k] += img.V[t][si][k];
                            mols[j].Ms[t][si][k] += img.Ms[t][si][k];
                        }
            }
        }
    for (int i = 0; i < n; ++i) calc_fm(x, &mols[i], F + 3 * i, M + 3 * i, Vi + 3 * i);
    for (int k = 0; k < 4; ++k) macro[k] = mac[k];
    free(mols);
    return 0;
}

/* ----------------------------------------------------------------- Leapfrog */

/* Quaternion::differentiate, Quaternion.cpp:98-103 */
static void qdiff(const double *q, const double *w, double *dq) {
    dq[0] = .5 * (-q[1] * w[0] - q[2] * w[1] - q[3] * w[2]);
    dq[1] = .5 * (q[0] * w[0] - q[3] * w[1] + q[2] * w[2]);
    dq[2] = .5 * (q[3] * w[0] + q[0] * w[1] - q[1] * w[2]);
    dq[3] = .5 * (-q[2] * w[0] + q[1] * w[1] + q[0] * w[2]);
}

/* FullMolecule::upd_preF, FullMolecule.cpp:331-361 (Leapfrog::transition1to2,
 * Leapfrog.cpp:45-61) */
void ox_leapfrog_pre(void *h, int n, double dt, double *r, double *v, double *q, double *L,
                     const double *F, const double *M, const int *cid) {
    ox_ctx *x = (ox_ctx *)h;
#pragma omp parallel for num_threads(x->nthreads) schedule(static) if (x->nthreads > 1)
    for (int i = 0; i < n; ++i) {
        const ox_comp *c = &x->comp[cid[i]];
        double dt_halve = .5 * dt;
        double dtInv2m = dt_halve / c->m;
        for (int d = 0; d < 3; ++d) {
            v[3 * i + d] += dtInv2m * F[3 * i + d];
            r[3 * i + d] += dt * v[3 * i + d];
        }
        double *qi = q + 4 * i, *Li = L + 3 * i;
        double w[3], qh[4], qi2[4];
        qrotinv(qi, Li, w);
        for (int d = 0; d < 3; ++d) w[d] *= c->invI[d];
        qdiff(qi, w, qh);
        for (int k = 0; k < 4; ++k) qh[k] *= dt_halve;
        for (int k = 0; k < 4; ++k) qh[k] += qi[k];
        double qcorr = 1. / sqrt(qh[0] * qh[0] + qh[1] * qh[1] + qh[2] * qh[2] + qh[3] * qh[3]);
        for (int k = 0; k < 4; ++k) qh[k] *= qcorr;
        for (int d = 0; d < 3; ++d) Li[d] += dt_halve * M[3 * i + d];
        qrotinv(qh, Li, w);
        for (int d = 0; d < 3; ++d) w[d] *= c->invI[d];
        qdiff(qh, w, qi2);
        for (int k = 0; k < 4; ++k) qi2[k] *= dt;
        for (int k = 0; k < 4; ++k) qi[k] += qi2[k];
        qcorr = 1. / sqrt(qi[0] * qi[0] + qi[1] * qi[1] + qi[2] * qi[2] + qi[3] * qi[3]);
        for (int k = 0; k < 4; ++k) qi[k] *= qcorr;
    }
}

/* FullMolecule::upd_postF, FullMolecule.cpp:363-387 (Leapfrog::transition2to3,
 * Leapfrog.cpp:63-147, single thermostat): sums[0] = sum m v^2,
 * sums[1] = sum I w^2, sums[2] = N, sums[3] = rotational DOF. */
void ox_leapfrog_post(void *h, int n, double dt, double *v, double *L, const double *q,
                      const double *F, const double *M, const int *cid, double *sums) {
    ox_ctx *x = (ox_ctx *)h;
    double summv2 = 0, sumIw2 = 0, rot = 0;
    double dt_halve = 0.5 * dt;
#pragma omp parallel for num_threads(x->nthreads) schedule(static) if (x->nthreads > 1) \
    reduction(+ : summv2, sumIw2, rot)
    for (int i = 0; i < n; ++i) {
        const ox_comp *c = &x->comp[cid[i]];
        double dtInv2m = dt_halve / c->m;
        double v2 = 0.;
        for (int d = 0; d < 3; ++d) {
            v[3 * i + d] += dtInv2m * F[3 * i + d];
            v2 += v[3 * i + d] * v[3 * i + d];
            L[3 * i + d] += dt_halve * M[3 * i + d];
        }
        summv2 += c->m * v2;
        double w[3], Iw2 = 0.;
        qrotinv(q + 4 * i, L + 3 * i, w);
        for (int d = 0; d < 3; ++d) {
            w[d] *= c->invI[d];
            Iw2 += c->I[d] * w[d] * w[d];
        }
        sumIw2 += Iw2;
        rot += c->rotdof;
    }
    sums[0] = summv2;
    sums[1] = sumIw2;
    sums[2] = n;
    sums[3] = rot;
}

/* <boundaries> per dimension (Simulation.cpp:429-444), in
 * BoundaryUtils::BoundaryType order (BoundaryUtils.h:29): 0 periodic,
 * 1 outflow, 2 reflecting. */
int ox_set_boundaries(void *h, const int *types) {
    ox_ctx *x = (ox_ctx *)h;
    int any = 0;
    for (int d = 0; d < 3; ++d) {
        if (types[d] < 0 || types[d] > 2) return -1;
        x->pdim[d] = types[d] == 0;
        x->wall[d] = types[d] == 2;
        x->outflow[d] = types[d] == 1;
        any |= x->pdim[d];
    }
    x->periodic = any;
    return 0;
}

/* BoundaryHandler::processGlobalWallLeavingParticles (BoundaryHandler.cpp:
 * 82-134), walls in the map order of DimensionType (+x, -x, +y, -y, +z, -z,
 * DimensionUtils.h:26): a molecule inside the wall's slab of width rc
 * (RegionUtils::getInnerRegionSlab, RegionUtils.cpp:11-45; inBox l <= r < u)
 * whose predicted position r + dt (v + dt/2m F) reaches the wall
 * (RegionUtils::isMoleculeLeaving, RegionUtils.cpp:47-56) gets
 * v = -v - 2 (dt/2m) F at a reflecting wall and is deleted at an outflow
 * wall (dead[i] = 1; a deleted molecule is not seen by later walls). dead may
 * be NULL when no wall is an outflow wall; molecules already dead on entry
 * are skipped. */
void ox_reflect(void *h, int n, double dt, const double *r, double *v, const double *F,
                const int *cid, int *dead) {
    const ox_ctx *x = (const ox_ctx *)h;
    for (int wall = 0; wall < 6; ++wall) {
        const int d = wall / 2, pos = (wall % 2) == 0;
        if (!x->wall[d] && !(x->outflow[d] && dead)) continue;
        double lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = x->bbmin[k];
            hi[k] = x->bbmax[k];
        }
        if (pos)
            lo[d] = fmax(x->bbmax[d] - x->rc, x->bbmin[d]);
        else
            hi[d] = fmin(x->bbmin[d] + x->rc, x->bbmax[d]);
        for (int i = 0; i < n; ++i) {
            if (dead && dead[i]) continue;
            const double *ri = r + 3 * i;
            int in = 1;
            for (int k = 0; k < 3; ++k) in &= (ri[k] >= lo[k] && ri[k] < hi[k]);
            if (!in) continue;
            const double half = .5 * dt;
            const double hm = half / x->comp[cid[i]].m;
            const double adj = hm * F[3 * i + d];
            const double np = ri[d] + (dt * (v[3 * i + d] + adj));
            if ((pos && np >= x->bbmax[d]) || (!pos && np <= x->bbmin[d])) {
                if (x->wall[d])
                    v[3 * i + d] = -v[3 * i + d] - adj - adj;
                else
                    dead[i] = 1;
            }
        }
    }
}

static int any_wall(const ox_ctx *x) {
    return x->wall[0] | x->wall[1] | x->wall[2] | x->outflow[0] | x->outflow[1] | x->outflow[2];
}

static int any_outflow(const ox_ctx *x) { return x->outflow[0] | x->outflow[1] | x->outflow[2]; }

/* Removes the molecules flagged dead, keeping the order of the rest (the
 * container's deleteMolecule, ParticleContainer.h:230); orig (may be NULL)
 * moves along. Returns the new count. */
static int compact_dead(int n, const int *dead, double *r, double *v, double *q, double *L,
                        int *cid, double *F, double *M, double *Vi, int *orig) {
    int k = 0;
    for (int i = 0; i < n; ++i) {
        if (dead[i]) continue;
        if (k != i) {
            for (int d = 0; d < 3; ++d) {
                r[3 * k + d] = r[3 * i + d];
                v[3 * k + d] = v[3 * i + d];
                L[3 * k + d] = L[3 * i + d];
                F[3 * k + d] = F[3 * i + d];
                M[3 * k + d] = M[3 * i + d];
                Vi[3 * k + d] = Vi[3 * i + d];
            }
            for (int d = 0; d < 4; ++d) q[4 * k + d] = q[4 * i + d];
            cid[k] = cid[i];
            if (orig) orig[k] = orig[i];
        }
        ++k;
    }
    return k;
}

/* Simulation::simulateOneTimestep step order (Simulation.cpp:1076-1307)
 * without thermostat, long-range correction or plugins: upd_preF, periodic
 * wrap, re-bin + halo + SoA, traversal, calcFM, upd_postF. F/M must hold the
 * forces of the initial configuration on entry (Simulation.cpp:909-925).
 * per_step[s*6 + ...] = Upot, virial, summv2, sumIw2, upot6lj/6, upotXpoles+myRF */
static int simulate_impl(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                         double *L, int *cid, double *F, double *M, double *Vi, int *orig,
                         double *per_step, int width) {
    double mac[4], sums[4];
    int *dead = any_outflow((ox_ctx *)h) ? (int *)calloc((size_t)n + 1, sizeof(int)) : NULL;
    for (int s = 0; s < nsteps; ++s) {
        if (any_wall((ox_ctx *)h)) ox_reflect(h, n, dt, r, v, F, cid, dead);
        if (dead) {  /* deleted before their upd_preF (Simulation.cpp:1114-1117) */
            n = compact_dead(n, dead, r, v, q, L, cid, F, M, Vi, orig);
            memset(dead, 0, sizeof(int) * (size_t)n);
        }
        ox_leapfrog_pre(h, n, dt, r, v, q, L, F, M, cid);
        if (((ox_ctx *)h)->periodic) ox_wrap(h, n, r);
        if (ox_forces(h, n, r, q, cid, F, M, Vi, mac, NULL)) {
            free(dead);
            return -1;
        }
        ox_leapfrog_post(h, n, dt, v, L, q, F, M, cid, sums);
        if (per_step) {
            double *o = per_step + width * s;
            o[0] = mac[OX_UPOT6LJ] / 6.0 + mac[OX_UPOTXPOLES] + mac[OX_MYRF];
            o[1] = mac[OX_VIRIAL] + 3.0 * mac[OX_MYRF];
            o[2] = sums[0];
            o[3] = sums[1];
            o[4] = mac[OX_UPOT6LJ] / 6.0;
            o[5] = mac[OX_UPOTXPOLES] + mac[OX_MYRF];
            if (width > 6) o[6] = sums[2];
        }
    }
    free(dead);
    return n;
}

int ox_simulate(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                double *L, const int *cid, double *F, double *M, double *Vi, double *per_step) {
    if (any_outflow((ox_ctx *)h)) return -1; /* the count would change: ox_simulate_open */
    return simulate_impl(h, n, nsteps, dt, r, v, q, L, (int *)cid, F, M, Vi, NULL, per_step, 6) < 0
               ? -1
               : 0;
}

/* ox_simulate with outflow walls: molecules leaving through them are
 * deleted, the arrays are compacted in place (survivors in their order, orig
 * carries each survivor's index on entry) and the count is returned;
 * per_step rows get a 7th column, N. */
int ox_simulate_open(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                     double *L, int *cid, double *F, double *M, double *Vi, int *orig,
                     double *per_step) {
    return simulate_impl(h, n, nsteps, dt, r, v, q, L, cid, F, M, Vi, orig, per_step, 7);
}

/* ------------------------------------------- global values, LRC, thermostat */

/* Homogeneous::_TICCu/_TICSu/_TISSu/_TICCv/_TICSv/_TISSv
 * (src/longRange/Homogeneous.cpp:158-196): radial integrals of the r^(2n)
 * terms beyond rc for centre-centre, centre-site and site-site distances
 * tau (orientation averaged). */
static double ticc_u(int n, double rc, double s2) {
    return -pow(rc, 2 * n + 3) / (pow(s2, n) * (2 * n + 3));
}
static double tics_u(int n, double rc, double s2, double tau) {
    return -(pow(rc + tau, 2 * n + 3) - pow(rc - tau, 2 * n + 3)) * rc /
               (4 * pow(s2, n) * tau * (n + 1) * (2 * n + 3)) +
           (pow(rc + tau, 2 * n + 4) - pow(rc - tau, 2 * n + 4)) /
               (4 * pow(s2, n) * tau * (n + 1) * (2 * n + 3) * (2 * n + 4));
}
static double tiss_u(int n, double rc, double s2, double t1, double t2) {
    double tp = t1 + t2, tm = t1 - t2;
    return -(pow(rc + tp, 2 * n + 4) - pow(rc + tm, 2 * n + 4) - pow(rc - tm, 2 * n + 4) +
             pow(rc - tp, 2 * n + 4)) *
               rc / (8 * pow(s2, n) * t1 * t2 * (n + 1) * (2 * n + 3) * (2 * n + 4)) +
           (pow(rc + tp, 2 * n + 5) - pow(rc + tm, 2 * n + 5) - pow(rc - tm, 2 * n + 5) +
            pow(rc - tp, 2 * n + 5)) /
               (8 * pow(s2, n) * t1 * t2 * (n + 1) * (2 * n + 3) * (2 * n + 4) * (2 * n + 5));
}
static double ticc_v(int n, double rc, double s2) { return 2 * n * ticc_u(n, rc, s2); }
static double tics_v(int n, double rc, double s2, double tau) {
    return -(pow(rc + tau, 2 * n + 2) - pow(rc - tau, 2 * n + 2)) * rc * rc /
               (4 * pow(s2, n) * tau * (n + 1)) -
           3 * tics_u(n, rc, s2, tau);
}
static double tiss_v(int n, double rc, double s2, double t1, double t2) {
    double tp = t1 + t2, tm = t1 - t2;
    return -(pow(rc + tp, 2 * n + 3) - pow(rc + tm, 2 * n + 3) - pow(rc - tm, 2 * n + 3) +
             pow(rc - tp, 2 * n + 3)) *
               rc * rc / (8 * pow(s2, n) * t1 * t2 * (n + 1) * (2 * n + 3)) -
           3 * tiss_u(n, rc, s2, t1, t2);
}

/* Homogeneous::init + calculateLongRange (Homogeneous.cpp:21-146) for nmol[i]
 * molecules of component i in a box of volume `volume`: out[0] = UpotCorr,
 * out[1] = VirialCorr, out[2] = the per-molecule virial diagonal addition
 * VirialCorr / (3 N). Only LJ table entries with shift6 == 0 contribute (the
 * reference reads uLJshift6 from Comp2Param, so truncated-shifted and all
 * cross-component pairs add nothing); MySelbstTerm is the reaction-field self
 * term of the molecules' dipole (charges' first moment + point dipoles).
 * Returns -1 where the reference exits (tau1 + tau2 >= rcLJ). */
int ox_lrc_homogeneous(void *h, const double *nmol, double volume, double *out) {
    ox_ctx *x = (ox_ctx *)h;
    double upot = 0., vir = 0., self = 0., N = 0.;
    const int nl = x->nlj_total;
    for (int i = 0; i < x->ncomp; ++i) {
        const ox_comp *ci = &x->comp[i];
        N += nmol[i];
        double cb[3] = {0, 0, 0};
        for (int s = 0; s < ci->nc; ++s)
            for (int d = 0; d < 3; ++d) cb[d] += ci->c_q[s] * ci->c[s][d];
        for (int s = 0; s < ci->nd; ++s) {
            double ev = 0.;
            for (int d = 0; d < 3; ++d) ev += ci->d_e[s][d] * ci->d_e[s][d];
            double norm = 1.0 / sqrt(ev);
            for (int d = 0; d < 3; ++d) cb[d] += ci->d_p[s] * ci->d_e[s][d] * norm;
        }
        double my2 = 0.;
        for (int d = 0; d < 3; ++d) my2 += cb[d] * cb[d];
        self += my2 * nmol[i];
        for (int j = 0; j < x->ncomp; ++j) {
            const ox_comp *cj = &x->comp[j];
            for (int si = 0; si < ci->nlj; ++si) {
                const double *a = ci->lj[si];
                double t1 = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
                for (int sj = 0; sj < cj->nlj; ++sj) {
                    const double *b = cj->lj[sj];
                    double t2 = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
                    if (t1 + t2 >= x->rcLJ) return -1;
                    const int k = (ci->lj_base + si) * nl + cj->lj_base + sj;
                    const double eps24 = x->eps24[k], sig2 = x->sig2[k];
                    if (x->shift6[k] != 0.0) continue;
                    double fac = nmol[i] * nmol[j] * eps24;
                    const double rc = x->rcLJ;
                    if (t1 == 0. && t2 == 0.) {
                        upot += fac * (ticc_u(-6, rc, sig2) - ticc_u(-3, rc, sig2));
                        vir += fac * (ticc_v(-6, rc, sig2) - ticc_v(-3, rc, sig2));
                    } else if (t1 != 0. && t2 != 0.) {
                        upot += fac * (tiss_u(-6, rc, sig2, t1, t2) - tiss_u(-3, rc, sig2, t1, t2));
                        vir += fac * (tiss_v(-6, rc, sig2, t1, t2) - tiss_v(-3, rc, sig2, t1, t2));
                    } else {
                        double t = (t2 == 0.) ? t1 : t2;
                        upot += fac * (tics_u(-6, rc, sig2, t) - tics_u(-3, rc, sig2, t));
                        vir += fac * (tics_v(-6, rc, sig2, t) - tics_v(-3, rc, sig2, t));
                    }
                }
            }
        }
    }
    const double rho = N / volume;
    const double fac = M_PI * rho / (3. * N);
    const double uLJ = fac * upot, wLJ = -fac * vir;
    const double epsRFInvrc3 =
        2. * (x->epsRF - 1.) / ((x->rc * x->rc * x->rc) * (2. * x->epsRF + 1.));
    const double myself = -0.5 * epsRFInvrc3 * self;
    out[0] = uLJ + myself;
    out[1] = wLJ + 3. * myself;
    out[2] = out[1] / (3. * N);
    return 0;
}

/* Domain::calculateGlobalValues for the one global thermostat
 * (Domain.cpp:157-312) with the default temperature schedule Tfactor = 1
 * (Simulation.cpp:104-105, 1489-1501). in = {summv2, sumIw2, N, rotDOF} from
 * upd_postF. out = {T, beta_trans, beta_rot, explode}. The explosion
 * heuristic (Domain.cpp:261-313) fires when a beta is below MIN_BETA = 0.9;
 * its error counter can never become positive in the reference's bookkeeping
 * (the counter is only ever set to 4000), so it stays armed. */
void ox_thermostat_params(const double *in, double Ttarget, int heuristics, double *out) {
    const double summv2 = in[0], N = in[2], rot = in[3];
    const double sumIw2 = (rot > 0) ? in[1] : 0.0;
    out[0] = (N > 0) ? (summv2 + sumIw2) / (3 * N + rot) : Ttarget;
    double bt = 1.0, br = 1.0;
    if (Ttarget > 0.0 && N > 0) {
        bt = pow(3.0 * N * Ttarget / summv2, 0.4);
        br = (sumIw2 == 0.0) ? 1.0 : pow(rot * Ttarget / sumIw2, 0.4);
    }
    int explode = 0;
    if ((bt < 0.9 || br < 0.9) && heuristics) {
        explode = 1;
        bt = 1.0;
        br = pow(br, 0.0091);
    }
    out[1] = bt;
    out[2] = br;
    out[3] = explode;
}

/* The explosion capping (Domain.cpp:263-297: kinetic energies above
 * 10 Ti translational / 3 rotDOF Ti rotational are scaled down together with
 * F and M) when tp[3] != 0, then VelocityScalingThermostat::apply without a
 * directed velocity (VelocityScalingThermostat.cpp:70-90): v *= beta_trans,
 * D *= beta_rot. tp = ox_thermostat_params output. */
void ox_velocity_scaling(void *h, int n, double *v, double *L, double *F, double *M,
                         const double *q, const int *cid, const double *tp, double Ttarget) {
    ox_ctx *x = (ox_ctx *)h;
    for (int i = 0; i < n; ++i) {
        const ox_comp *c = &x->comp[cid[i]];
        double *vi = v + 3 * i, *Li = L + 3 * i;
        if (tp[3] != 0.) {
            const double limit = 10.0 * Ttarget;
            const double ut = 0.5 * c->m * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
            if (ut > limit) {
                const double vc = sqrt(limit / ut);
                for (int d = 0; d < 3; ++d) {
                    vi[d] *= vc;
                    F[3 * i + d] *= vc;
                }
            }
            if (c->rotdof > 0) {
                const double lr = 3.0 * c->rotdof * Ttarget;
                double w[3], Iw2 = 0.;
                qrotinv(q + 4 * i, Li, w);
                for (int d = 0; d < 3; ++d) {
                    w[d] *= c->invI[d];
                    Iw2 += c->I[d] * w[d] * w[d];
                }
                const double ur = 0.5 * Iw2;
                if (ur > lr) {
                    const double dc = sqrt(lr / ur);
                    for (int d = 0; d < 3; ++d) {
                        Li[d] *= dc;
                        M[3 * i + d] *= dc;
                    }
                }
            }
        }
        for (int d = 0; d < 3; ++d) {
            vi[d] *= tp[1];
            Li[d] *= tp[2];
        }
    }
}

/* ox_simulate with the NVT tail of Simulation::simulateOneTimestep
 * (Simulation.cpp:1161-1266): the homogeneous long-range correction adds
 * lrc[2] to every molecule's virial diagonal after the traversal and lrc[0],
 * lrc[1] to the global Upot / virial (Domain.cpp:187-188); after upd_postF the
 * thermostat parameters are computed and the velocities scaled. lrc may be
 * NULL (no correction); Ttarget <= 0 disables the thermostat (betas 1).
 * per_step[s*9 + ...] = the six ox_simulate values (Upot and virial including
 * the correction), then T, beta_trans, beta_rot. */
int ox_simulate_nvt(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                    double *L, const int *cid, double *F, double *M, double *Vi,
                    const double *lrc, double Ttarget, int heuristics, double *per_step) {
    double mac[4], sums[4], tp[4];
    if (any_outflow((ox_ctx *)h)) return -1; /* NVT runs keep their molecule count */
    for (int s = 0; s < nsteps; ++s) {
        if (any_wall((ox_ctx *)h)) ox_reflect(h, n, dt, r, v, F, cid, NULL);
        ox_leapfrog_pre(h, n, dt, r, v, q, L, F, M, cid);
        if (((ox_ctx *)h)->periodic) ox_wrap(h, n, r);
        if (ox_forces(h, n, r, q, cid, F, M, Vi, mac, NULL)) return -1;
        if (lrc)
            for (int i = 0; i < 3 * n; ++i) Vi[i] = lrc[2] + Vi[i];
        ox_leapfrog_post(h, n, dt, v, L, q, F, M, cid, sums);
        ox_thermostat_params(sums, Ttarget, heuristics, tp);
        if (Ttarget > 0.0) ox_velocity_scaling(h, n, v, L, F, M, q, cid, tp, Ttarget);
        if (per_step) {
            double *o = per_step + 9 * s;
            o[0] = mac[OX_UPOT6LJ] / 6.0 + mac[OX_UPOTXPOLES] + mac[OX_MYRF];
            o[1] = mac[OX_VIRIAL] + 3.0 * mac[OX_MYRF];
            if (lrc) {
                o[0] = o[0] + lrc[0];
                o[1] = o[1] + lrc[1];
            }
            o[2] = sums[0];
            o[3] = sums[1];
            o[4] = mac[OX_UPOT6LJ] / 6.0;
            o[5] = mac[OX_UPOTXPOLES] + mac[OX_MYRF];
            o[6] = tp[0];
            o[7] = tp[1];
            o[8] = tp[2];
        }
    }
    return 0;
}

/* Reduced-memory one-centre mode: VCP1CLJRMM with LeapfrogRMM
 * (VCP1CLJRMM.cpp:300-362, LeapfrogRMM.cpp:42-107, Simulation.cpp:902-937).
 * Velocities live at half steps: the initial traversal adds (dt/m)/2 F(r0)
 * to v; each step moves r += dt v (PositionCellProcessorRMM), wraps, and the
 * force traversal adds dt/m F to v; VelocityCellProcessorRMM then sums v^2
 * (no mass factor). One component of one LJ centre. per_step rows as
 * ox_simulate: Upot, virial, sum v^2, 0, Upot_LJ, Upot_rest. */
int ox_simulate_rmm(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                    const int *cid, double *per_step) {
    ox_ctx *x = (ox_ctx *)h;
    if (x->ncomp != 1 || x->comp[0].nlj != 1 || x->comp[0].nc || x->comp[0].nd ||
        x->comp[0].nq)
        return -2;
    double *F = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *M = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *Vi = (double *)calloc(3 * (size_t)n, sizeof(double));
    double mac[4];
    int rc = 0;
    const double dtInvm = dt / x->comp[0].m;
    if (ox_forces(h, n, r, q, cid, F, M, Vi, mac, NULL)) rc = -1;
    const double half = dtInvm * 0.5;
    for (int i = 0; !rc && i < 3 * n; ++i) v[i] = v[i] + F[i] * half;
    for (int s = 0; !rc && s < nsteps; ++s) {
        for (int i = 0; i < 3 * n; ++i) r[i] = r[i] + dt * v[i];
        if (x->periodic) ox_wrap(h, n, r);
        if (ox_forces(h, n, r, q, cid, F, M, Vi, mac, NULL)) {
            rc = -1;
            break;
        }
        double summv2 = 0.;
        for (int i = 0; i < n; ++i) {
            double v2 = 0.;
            for (int d = 0; d < 3; ++d) {
                v[3 * i + d] = v[3 * i + d] + F[3 * i + d] * dtInvm;
                v2 += v[3 * i + d] * v[3 * i + d];
            }
            summv2 += v2;
        }
        if (per_step) {
            double *o = per_step + 6 * s;
            o[0] = mac[OX_UPOT6LJ] / 6.0 + mac[OX_UPOTXPOLES] + mac[OX_MYRF];
            o[1] = mac[OX_VIRIAL] + 3.0 * mac[OX_MYRF];
            o[2] = summv2;
            o[3] = 0.0;
            o[4] = mac[OX_UPOT6LJ] / 6.0;
            o[5] = mac[OX_UPOTXPOLES] + mac[OX_MYRF];
        }
    }
    free(F);
    free(M);
    free(Vi);
    return rc;
}

/* ------------------------------------------------------------------ adjoint */

/* How the LJ table entry (a,b) depends on its own (eps_ab, sigma_ab):
 * shift6 = 24 eps (s6 - s6^2), s6 = (sigma / rcs)^6, when the entry is
 * shifted with a cutoff that belongs to the pair: different components
 * (always shifted at rcLJ, Comp2Param.cpp:79-105) or the same site of one
 * component with do_shift (its own rc_shift, Component.cpp:103-116). Returns
 * the shift cutoff, or 0 when shift6 does not depend on (eps_ab, sigma_ab)
 * (unshifted, or two different sites of one component, whose entry carries
 * site a's own shift: Comp2Param.cpp:27-38). */
static double pair_shift_rc(const ox_ctx *x, int a, int b) {
    int ca = -1, cb = -1, sa = 0, sb = 0;
    for (int c = 0; c < x->ncomp; ++c) {
        const ox_comp *k = &x->comp[c];
        if (a >= k->lj_base && a < k->lj_base + k->nlj) { ca = c; sa = a - k->lj_base; }
        if (b >= k->lj_base && b < k->lj_base + k->nlj) { cb = c; sb = b - k->lj_base; }
    }
    if (ca != cb) return x->rcLJ;
    if (sa == sb && x->comp[ca].lj_shifted[sa]) return x->comp[ca].lj_rcs[sa];
    return 0.0;
}

void ox_lj_shift_rc(void *h, double *out) {
    ox_ctx *x = (ox_ctx *)h;
    int n = x->nlj_total;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) out[a * n + b] = pair_shift_rc(x, a, b);
}

/* Brute-force FP64 adjoint over unordered molecule pairs (minimum image when
 * periodic), LJ and charge-charge site pairs:
 *   dU[a*n+b] (and dW) = derivative with respect to the symmetric pair
 *   parameter eps_ab / sigma_ab, U = sum upot6/6 + sum upotXpoles,
 *   W = sum (m_i - m_j) . f_site;
 *   dUq[c*8+s], dWq = derivative with respect to charge s of component c;
 *   vjp[3i] = sum_j H_ij (cot_i - cot_j), H = d f_site / d r_i (molecule
 *   translations, orientations held fixed). */
int ox_adjoint(void *h, int n, const double *r, double *q, const int *cid, const double *cot,
               double *dUe, double *dUs, double *dWe, double *dWs, double *dUq, double *dWq,
               double *vjp) {
    ox_ctx *x = (ox_ctx *)h;
    const int nl = x->nlj_total;
    for (int k = 0; k < nl * nl; ++k) dUe[k] = dUs[k] = dWe[k] = dWs[k] = 0;
    for (int k = 0; k < x->ncomp * OX_MAXSITE; ++k) dUq[k] = dWq[k] = 0;
    for (int k = 0; k < 3 * n; ++k) vjp[k] = 0;
    ox_mol *mols = (ox_mol *)malloc(sizeof(ox_mol) * (size_t)(n + 1));
    if (!mols) return -1;
    for (int i = 0; i < n; ++i) {
        setup_mol(x, &mols[i], r + 3 * i, q + 4 * i, cid[i], i);
        for (int k = 0; k < 4; ++k) q[4 * i + k] = mols[i].q[k];
    }
    ox_mol img;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            ox_mol *B = &mols[j];
            double sft[3] = {0, 0, 0};
            if (x->periodic)
                for (int d = 0; d < 3; ++d) {
                    if (!x->pdim[d]) continue;
                    double L = x->bbmax[d] - x->bbmin[d], dd = mols[i].r[d] - mols[j].r[d];
                    if (dd > 0.5 * L) sft[d] = L;
                    else if (dd < -0.5 * L) sft[d] = -L;
                }
            if (sft[0] != 0 || sft[1] != 0 || sft[2] != 0) {
                double rr[3] = {B->r[0] + sft[0], B->r[1] + sft[1], B->r[2] + sft[2]};
                setup_mol(x, &img, rr, B->q, B->cid, j);
                B = &img;
            }
            const ox_mol *A = &mols[i];
            double dm[3] = {A->r[0] - B->r[0], A->r[1] - B->r[1], A->r[2] - B->r[2]};
            double r2m = dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2];
            const ox_comp *ca = &x->comp[A->cid], *cb = &x->comp[B->cid];
            double dc[3] = {cot[3 * i] - cot[3 * j], cot[3 * i + 1] - cot[3 * j + 1],
                            cot[3 * i + 2] - cot[3 * j + 2]};
            if (r2m < x->rcLJ * x->rcLJ) {
                for (int s = 0; s < ca->nlj; ++s)
                    for (int t = 0; t < cb->nlj; ++t) {
                        int a = ca->lj_base + s, b = cb->lj_base + t, k = a * nl + b;
                        const double *ra = A->s[0][s], *rb = B->s[0][t];
                        double d[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
                        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r2i = 1.0 / r2;
                        double e24 = x->eps24[k], s2 = x->sig2[k], eps = e24 / 24., sig = sqrt(s2);
                        double l6 = s2 * r2i; l6 = l6 * l6 * l6;
                        double l12 = l6 * l6;
                        double sc = e24 * r2i * (2 * l12 - l6);
                        double sp = e24 * r2i * r2i * (4 * l6 - 14 * l12);
                        double dmd = dm[0] * d[0] + dm[1] * d[1] + dm[2] * d[2];
                        double gU_e = 4 * (l12 - l6), gU_s = 4 * eps * (12 * l12 - 6 * l6) / sig;
                        double rcs = pair_shift_rc(x, a, b);
                        if (rcs > 0) {
                            double s6 = s2 / (rcs * rcs); s6 = s6 * s6 * s6;
                            gU_e += 4 * (s6 - s6 * s6);
                            gU_s += 4 * eps * (6 * s6 - 12 * s6 * s6) / sig;
                        }
                        double gW_e = dmd * 24 * r2i * (2 * l12 - l6);
                        double gW_s = dmd * e24 * r2i * (24 * l12 - 6 * l6) / sig;
                        int kk[2] = {a * nl + b, b * nl + a};
                        for (int u = 0; u < (a == b ? 1 : 2); ++u) {
                            dUe[kk[u]] += gU_e; dUs[kk[u]] += gU_s;
                            dWe[kk[u]] += gW_e; dWs[kk[u]] += gW_s;
                        }
                        double ddc = d[0] * dc[0] + d[1] * dc[1] + d[2] * dc[2];
                        for (int c = 0; c < 3; ++c) {
                            double g = sc * dc[c] + 2 * sp * d[c] * ddc;
                            vjp[3 * i + c] += g;
                            vjp[3 * j + c] -= g;
                        }
                    }
            }
            if (!(r2m < x->rc * x->rc)) continue;
            for (int s = 0; s < ca->nc; ++s)
                for (int t = 0; t < cb->nc; ++t) {
                    const double *ra = A->s[1][s], *rb = B->s[1][t];
                    double d[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r2i = 1.0 / r2;
                    double ri = sqrt(r2i), r3i = r2i * ri;
                    double qa = ca->c_q[s], qb = cb->c_q[t];
                    double dmd = dm[0] * d[0] + dm[1] * d[1] + dm[2] * d[2];
                    dUq[A->cid * OX_MAXSITE + s] += qb * ri;
                    dUq[B->cid * OX_MAXSITE + t] += qa * ri;
                    dWq[A->cid * OX_MAXSITE + s] += qb * dmd * r3i;
                    dWq[B->cid * OX_MAXSITE + t] += qa * dmd * r3i;
                    double sc = qa * qb * r3i, sp = -1.5 * qa * qb * r3i * r2i;
                    double ddc = d[0] * dc[0] + d[1] * dc[1] + d[2] * dc[2];
                    for (int c = 0; c < 3; ++c) {
                        double g = sc * dc[c] + 2 * sp * d[c] * ddc;
                        vjp[3 * i + c] += g;
                        vjp[3 * j + c] -= g;
                    }
                }
        }
    free(mols);
    return 0;
}
