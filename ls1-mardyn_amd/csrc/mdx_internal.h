// mdx_internal.h -- data layout shared by the kernels and the C ABI.
//
// HBM layout (all arrays structure-of-arrays, one entry per molecule or site):
//  * molecule state, double buffered for the counting-sort permutation:
//    r, v, L (3 x FP64 each, SoA), q (4 x FP64), cid (int32), orig (int32 =
//    upload index). After every re-bin the molecules are stored cell by cell
//    (cells x-fastest, ls1's linearisation), so every later pass is coalesced.
//  * forces F, M, Vi (3 x FP64) per molecule, in the current order.
//  * per site type t in {LJ, charge, dipole, quadrupole}: contiguous site
//    arrays in molecule order; cell c owns sites [cstart_t[c], cstart_t[c+1]).
//    Each site carries its position and its molecule's position (the mask is
//    on molecule centres), FP64 absolute and FP32 relative to its cell's lower
//    corner, its orientation (dipoles, quadrupoles) and its scalar parameter
//    (LJ lookup id, charge, |mu|, |Q|), plus FP64 force / virial / torque
//    outputs of the pair kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdx {

constexpr int kMaxComp = 16;
constexpr int kMaxSite = 8;
constexpr int kTypes = 4;  // LJ, C, D, Q
enum SiteType { T_LJ = 0, T_C = 1, T_D = 2, T_Q = 3 };

struct Component {
    int n[kTypes];          // nlj, nc, nd, nq
    int lj_base;            // first LJ lookup id (EnsembleBase.cpp:98-111)
    int rotdof;
    int point;              // all sites at the centre, no D/Q sites, invI = 0: the
                            // orientation never changes and carries no force
    double m, I[3], invI[3];
    double body[kTypes][kMaxSite][3];   // site offsets, principal frame
    double e[kTypes][kMaxSite][3];      // D, Q axes (body frame, as given)
    double param[kTypes][kMaxSite];     // C: q, D: |mu|, Q: |Q|
};

struct Model {
    int ncomp;
    int nlj;
    Component comp[kMaxComp];
};

struct Geom {
    double bbmin[3], bbmax[3], L[3];
    double cl[3], rcl[3], halo_min[3], halo_max[3];
    int n[3];          // inner cells per dimension
    int cpd[3];        // ls1 cells per dimension (incl. halo)
    int pdim[3];       // periodic per dimension (images across the box)
    int reflect[3];    // reflecting walls per dimension (BoundaryType::REFLECTING)
    int outflow[3];    // outflow walls per dimension (BoundaryType::OUTFLOW): leaving
                       // molecules are deleted (their position becomes NaN and the
                       // next re-bin moves them behind the last cell)
    double wall_rc;    // width of the wall slab checked for leaving molecules (rc)
    double rc2, rclj2, epsRFInvrc3;
};

// Site arrays of one type (device pointers).
struct SiteArrays {
    int *mol;                  // molecule index (current order)
    double *x, *y, *z;         // site position, FP64 absolute
    double *mx, *my, *mz;      // molecule position, FP64 absolute
    float *fx, *fy, *fz;       // site position relative to the cell corner
    float *fmx, *fmy, *fmz;    // molecule position relative to the cell corner
    double *ex, *ey, *ez;      // orientation (D, Q)
    float *fex, *fey, *fez;
    double *p;                 // C: q, D: |mu|, Q: |Q|
    float *fp;
    int *ljid;                 // LJ lookup id
    int *nhead;                // charges: the molecule's charge count at its first
                               // charge site, 0 at the others (per-molecule tests)
    double *Fx, *Fy, *Fz;      // outputs: force
    double *Vx, *Vy, *Vz;      // virial
    double *Mx, *My, *Mz;      // torque (D, Q)
    int *cstart;               // per cell start (ncell + 1)
    int *mstart;               // per molecule start (n + 1)
};

// Up to kTypes independent exclusive scans of n ints each, array t on
// blockIdx.y = t; aux: the block sums (k_scan_block) or their exclusive scan
// (k_scan_add).
struct ScanArgs {
    const int *in[kTypes];
    int *out[kTypes];
    int *aux[kTypes];
    int n;
};

struct MolArrays {
    double *rx, *ry, *rz, *vx, *vy, *vz, *qw, *qx, *qy, *qz, *Lx, *Ly, *Lz;
    int *cid, *orig;
};

struct PairParams {
    const double *eps24, *sig2, *shift6;  // nlj x nlj
    const float *feps24, *fsig2, *fshift6;
    // the same tables interleaved, one 4-value record {eps24, sig2, shift6, 0}
    // per pair of LJ ids: a pair kernel fetches a partner's parameters with
    // one (FP32) or two (FP64) 16-byte loads at a 32-bit offset
    const double *dpar;
    const float *fpar;
    int nlj;
};


// Replica table: independent systems (same components and cutoffs, own box
// and cells) stored back to back. Replica r owns the global cells
// [cell_off[r], cell_off[r+1]) and, because the counting sort orders
// molecules by global cell, always the same molecule index range. A single
// system is the one-replica case (mol_rep / cell_rep null).
struct Reps {
    const Geom *geom;      // [nrep]
    const int *cell_off;   // [nrep + 1]
    const int *mol_rep;    // [n] replica of each molecule, or null
    const int *cell_rep;   // [ncell] replica of each global cell, or null
    int nrep;
};

__device__ __forceinline__ int mol_replica(const Reps &R, int i) {
    return R.mol_rep ? R.mol_rep[i] : 0;
}
__device__ __forceinline__ int cell_replica(const Reps &R, int c) {
    return R.cell_rep ? R.cell_rep[c] : 0;
}

// Per-replica record of the NVT tail (global thermostat, homogeneous LRC):
// target T, Upot / virial corrections, per-molecule virial addition, then the
// step's T, beta_trans, beta_rot and explosion flag.
enum {
    kThTarget = 0, kThUcorr = 1, kThWcorr = 2, kThVi = 3,
    kThT = 4, kThBetaT = 5, kThBetaR = 6, kThExplode = 7, kThermoW = 8
};

struct PairArgs {
    Reps R;
    double rc2, rclj2, epsRFInvrc3;   // shared by all replicas
    SiteArrays s[kTypes];
    PairParams pp;
    int ncell;          // global cells
    int cpb;            // cells per work item (consecutive cells, one block)
    int items_per_xcd;  // work items of each XCD's contiguous range
    int present[kTypes];
    double *macro_part; // [ncell][4] sums {u6 (LJ), uXpoles, virial, rf}: each work item
                        // stores its cells' sums in the row of its last non-empty cell
                        // (per replica) and zeros in the others
    int accumulate;     // add to macro_part rows (a later i-site type) instead of writing
    double vscale;      // factor on the written site virial: 1, or 0.5 where the site
                        // arrays are the molecule arrays (calcFM's Vi = 0.5 sum V_site)
    const double *thermo;  // [nrep][kThermoW] when the written virial also takes the
                           // long-range correction's per-molecule addition, else null
    const double *mrx, *mry, *mrz;  // molecule positions (the FP32 form's FP64 re-check)
    double *kvx, *kvy, *kvz;  // reduced-memory mode: the molecule velocities that take
    double kick;              // v += kick * F instead of storing F and Vi (else null)
    unsigned long long *counters;  // [8] or null
};

typedef void (*PairKernel)(PairArgs);
PairKernel pair_kernel(int ti, bool fp32, bool count, bool molsite, bool multi);

}  // namespace mdx
