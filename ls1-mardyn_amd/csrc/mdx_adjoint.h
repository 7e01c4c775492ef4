// mdx_adjoint.h -- arguments of the FP64 adjoint kernels (LJ and charge sites).
#pragma once

#include "mdx_internal.h"

namespace mdx {

constexpr int kAdjMaxLJ = 16;   // LJ lookup ids supported by the adjoint
constexpr int kAdjBlocks = 4096; // upper bound on adjoint blocks (partials rows)
constexpr int kAdjPar = 6;       // per-pair constants of the LJ adjoint (LDS)

struct AdjArgs {
    Geom g;
    SiteArrays s;              // i and j sites (one type: LJ or charge)
    PairParams pp;             // FP64 tables (LJ)
    const double *shift_rc;    // nlj x nlj: radius of the shift tied to (eps, sigma), 0 = none
    const int *mcid;           // molecule component ids, current order
    const double *cx, *cy, *cz;  // cotangent on molecule forces, current order (null: no VJP)
    double *gx, *gy, *gz;      // per-site VJP terms (null: not written)
    double *part;              // [blocks][K * W] parameter partials
    int nlj, ncell;
    int K, W;                  // keys per block table, values per key
    int site_is_mol;           // site j belongs to molecule j (one-centre systems)
};

typedef void (*AdjKernel)(AdjArgs);
typedef void (*AdjMolKernel)(int, const int *, const int *, const double *, const double *,
                             const double *, const double *, const double *, const double *,
                             double *);
typedef void (*GatherKernel)(int, const int *, const double *, double *);

AdjKernel adjoint_kernel(int ti, int nlj, bool sites_at_centre);
bool adjoint_columns_in_tile(int ti, int nlj);
AdjMolKernel adjoint_mol_kernel();
GatherKernel gather_orig3_kernel();

}  // namespace mdx
