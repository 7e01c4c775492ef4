/*
 * ls1_oracle.h -- CPU restatement (scalar FP64 C) of ls1-MarDyn's short-range
 * force path. TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * benchmark's cpu_baseline leg call it as the checker. The HIP product never
 * links, loads or calls anything in this directory.
 *
 * Component records mirror the ls1 ASCII phase-space header
 * (src/io/ASCIIReader.cpp:152-232), flattened to doubles, per component:
 *   nlj nc nd nq
 *   nlj x (x y z m eps sigma rc_shift do_shift)
 *   nc  x (x y z m q)
 *   nd  x (x y z ex ey ez absMy)
 *   nq  x (x y z ex ey ez absQ)
 *   I11 I22 I33            (values > 0 override the computed moments)
 * Mixing coefficients: (xi, eta) for every component pair i<j in the order
 * (0,1),(0,2),..,(1,2),.. exactly as ASCIIReader.cpp:213-232 reads them.
 */
#ifndef LS1_ORACLE_H
#define LS1_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* macro[] layout used by every force routine */
enum { OX_UPOT6LJ = 0, OX_UPOTXPOLES = 1, OX_VIRIAL = 2, OX_MYRF = 3 };

/* counters[] layout: molecule pairs whose centre distance was tested,
 * site pairs inside the cutoff per kernel type; pairs the reference evaluates
 * twice across a periodic boundary (without macroscopic sums the second
 * time) are counted once */
enum { OX_CNT_MOLPAIRS = 0, OX_CNT_LJ = 1, OX_CNT_CC = 2, OX_CNT_CD = 3,
       OX_CNT_DD = 4, OX_CNT_CQ = 5, OX_CNT_DQ = 6, OX_CNT_QQ = 7, OX_NCNT = 8 };

void *ox_create(int ncomp, const double *desc, int desc_len, const double *mix,
                double epsRF, double rc, double rcLJ, const double *box,
                int periodic, int cell_mode);
void ox_destroy(void *h);
/* OpenMP threads of the force traversal and the molecule loops (default 1,
 * the serial path every test uses; more only for timing the CPU baseline). */
void ox_set_threads(void *h, int nthreads);
/* Molecule-centre distance checks of the reference's C08 traversal: out[0]
 * all (halo pairs of periodic boundaries twice), out[1] those with
 * macroscopic sums (each unordered pair once). */
int ox_count_checks(void *h, int n, const double *r, const double *q, const int *cid,
                    long long *out);

int ox_nlj_total(void *h);
void ox_lj_tables(void *h, double *eps24, double *sig2, double *shift6);
void ox_set_lj_tables(void *h, const double *eps24, const double *sig2, const double *shift6);
void ox_component_props(void *h, double *mass, double *inertia, int *rotdof);
void ox_cell_geometry(void *h, int *cells_per_dim, double *cell_length,
                      double *halo_min, double *halo_max);

long ox_cell_index_of_point(void *h, const double *p);
void ox_cell_index_batch(void *h, int n, const double *r, long *out);
int ox_c08_pairs(void *h, int *out, int max_pairs);

void ox_wrap(void *h, int n, double *r);
int ox_forces(void *h, int n, const double *r, double *q, const int *cid,
              double *F, double *M, double *Vi, double *macro,
              long long *counters);
int ox_forces_brute(void *h, int n, const double *r, double *q, const int *cid,
                    double *F, double *M, double *Vi, double *macro);
void ox_leapfrog_pre(void *h, int n, double dt, double *r, double *v,
                     double *q, double *L, const double *F, const double *M,
                     const int *cid);
void ox_leapfrog_post(void *h, int n, double dt, double *v, double *L,
                      const double *q, const double *F, const double *M,
                      const int *cid, double *sums);
int ox_simulate(void *h, int n, int nsteps, double dt, double *r, double *v,
                double *q, double *L, const int *cid, double *F, double *M,
                double *Vi, double *per_step);

/* per-dimension boundaries (0 periodic, 2 reflecting) and the reflecting
 * walls' velocity reversal; see ls1_oracle.c */
int ox_set_boundaries(void *h, const int *types);
void ox_reflect(void *h, int n, double dt, const double *r, double *v, const double *F,
                const int *cid, int *dead);
int ox_simulate_open(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                     double *L, int *cid, double *F, double *M, double *Vi, int *orig,
                     double *per_step);

/* global values, homogeneous long-range correction, velocity-scaling
 * thermostat; see ls1_oracle.c */
int ox_lrc_homogeneous(void *h, const double *nmol, double volume, double *out);
void ox_thermostat_params(const double *in, double Ttarget, int heuristics, double *out);
void ox_velocity_scaling(void *h, int n, double *v, double *L, double *F, double *M,
                         const double *q, const int *cid, const double *tp, double Ttarget);
int ox_simulate_nvt(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                    double *L, const int *cid, double *F, double *M, double *Vi,
                    const double *lrc, double Ttarget, int heuristics, double *per_step);

/* reduced-memory one-centre mode (half-step velocities, kick in the force
 * traversal); see ls1_oracle.c */
int ox_simulate_rmm(void *h, int n, int nsteps, double dt, double *r, double *v, double *q,
                    const int *cid, double *per_step);

/* FP64 adjoint (brute force, LJ and charge-charge): parameter gradients of
 * Upot and virial and the force vector-Jacobian product; see ls1_oracle.c */
void ox_lj_shift_rc(void *h, double *out);
int ox_adjoint(void *h, int n, const double *r, double *q, const int *cid, const double *cot,
               double *dUe, double *dUs, double *dWe, double *dWs, double *dUq, double *dWq,
               double *vjp);

/* single site pair, canonical argument order; out = f[3], M1[3], M2[3], u, rf */
void ox_site_pair(int kind, const double *ri, const double *ei, double pi,
                  const double *rj, const double *ej, double pj,
                  double eps24, double sig2, double shift6, double epsRFInvrc3,
                  double *out);

#ifdef __cplusplus
}
#endif
#endif
