// mdx_pair_common.h -- pieces shared by the pair kernels (mdx_pairs.hip) and
// the adjoint kernel (mdx_adjoint.hip): vector helpers, the seven site-pair
// bodies, site accessors per precision, the neighbour-image table and the LDS
// tile of staged neighbour sites.
#pragma once

#include "mdx_internal.h"

namespace mdx {


template <typename R> struct V3 { R x, y, z; };

template <typename R> __device__ __forceinline__ R dot3(V3<R> a, V3<R> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename R> __device__ __forceinline__ V3<R> cross3(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <typename R> struct Math;
template <> struct Math<double> {
    // 1/x: v_rcp_f64 refined by two Newton steps (error <= 1 ulp; the
    // reference's AVX2 build likewise refines rcp estimates instead of
    // dividing, RealVecDouble.h:414-467); sqrt(1/r^2) as in the NOVEC build.
    static __device__ __forceinline__ double rcp(double x) {
        double y = __builtin_amdgcn_rcp(x);
        double e = __builtin_fma(-x, y, 1.0);
        y = __builtin_fma(y, e, y);
        e = __builtin_fma(-x, y, 1.0);
        return __builtin_fma(y, e, y);
    }
    static __device__ __forceinline__ double rinv(double r2, double r2inv) { return sqrt(r2inv); }
    // 1/r^2 and 1/r of one distance
    static __device__ __forceinline__ void inv2(double r2, double &r2inv, double &rinv) {
        r2inv = rcp(r2);
        rinv = sqrt(r2inv);
    }
};
template <> struct Math<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rinv(float r2, float) { return __builtin_amdgcn_rsqf(r2); }
    // two hardware approximations (1 ulp each); 1/r^2 as the square of 1/r
    // would save one but doubles the bias of v_rsq_f32 in every force, which
    // the sums of neutral molecules (charges of both signs) amplify
    static __device__ __forceinline__ void inv2(float r2, float &r2inv, float &rinv) {
        r2inv = __builtin_amdgcn_rcpf(r2);
        rinv = __builtin_amdgcn_rsqf(r2);
    }
};

template <typename R> struct PO {
    V3<R> f, M1, M2;
    R u, rf;
};

// _loopBodyCharge (VectorizedCellProcessor.cpp:227-277)
template <typename R> __device__ __forceinline__ void body_cc(V3<R> d, R q1, R q2, PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv, rinv;
    Math<R>::inv2(r2, r2inv, rinv);
    R upot = q1 * q2 * rinv;
    R fac = upot * r2inv;
    o.f = {d.x * fac, d.y * fac, d.z * fac};
    o.u = upot;
}

// _loopBodyChargeDipole (VectorizedCellProcessor.cpp:279-347): first charge,
// second dipole; M2 = torque on the dipole.
template <typename R>
__device__ __forceinline__ void body_cd(V3<R> d, R q, V3<R> e, R p, PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv, rinv;
    Math<R>::inv2(r2, r2inv, rinv);
    R re = dot3(d, e);
    R qpr3 = q * p * (r2inv * rinv);
    R fac = r2inv * R(3) * re;
    o.f = {qpr3 * (e.x - d.x * fac), qpr3 * (e.y - d.y * fac), qpr3 * (e.z - d.z * fac)};
    o.u = -(qpr3 * re);
    V3<R> c = cross3(e, d);
    o.M2 = {qpr3 * c.x, qpr3 * c.y, qpr3 * c.z};
}

// _loopBodyDipole (VectorizedCellProcessor.cpp:349-442) incl. reaction field
template <typename R>
__device__ __forceinline__ void body_dd(V3<R> d, V3<R> e1, R p1, V3<R> e2, R p2, R epsRFInvrc3,
                                        PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv, rinv;
    Math<R>::inv2(r2, r2inv, rinv);
    R r2three = R(3) * r2inv;
    R p1p2 = p1 * p2;
    R rffac = p1p2 * epsRFInvrc3;
    R p1p2r3 = p1p2 * rinv * r2inv;
    R p1p2r5 = p1p2r3 * r2three;
    R e1e2 = dot3(e1, e2), re1 = dot3(d, e1), re2 = dot3(d, e2);
    R re1t = re1 * r2three, re2t = re2 * r2three;
    R re1re2 = r2inv * re1 * re2;
    R gg = e1e2 - R(5) * re1re2;
    o.f = {p1p2r5 * (d.x * gg + e1.x * re2 + e2.x * re1),
           p1p2r5 * (d.y * gg + e1.y * re2 + e2.y * re1),
           p1p2r5 * (d.z * gg + e1.z * re2 + e2.z * re1)};
    o.u = p1p2r3 * (e1e2 - R(3) * re1re2);
    o.rf = rffac * e1e2;
    V3<R> c12 = cross3(e1, e2), e1xd = cross3(e1, d), e2xd = cross3(e2, d);
    o.M1 = {p1p2r3 * (re2t * e1xd.x - c12.x) + rffac * c12.x,
            p1p2r3 * (re2t * e1xd.y - c12.y) + rffac * c12.y,
            p1p2r3 * (re2t * e1xd.z - c12.z) + rffac * c12.z};
    o.M2 = {p1p2r3 * (re1t * e2xd.x + c12.x) - rffac * c12.x,
            p1p2r3 * (re1t * e2xd.y + c12.y) - rffac * c12.y,
            p1p2r3 * (re1t * e2xd.z + c12.z) - rffac * c12.z};
}

// _loopBodyChargeQuadrupole (VectorizedCellProcessor.cpp:444-524): first
// charge, second quadrupole; M2 = torque on the quadrupole.
template <typename R>
__device__ __forceinline__ void body_cq(V3<R> d, R q, V3<R> e, R m, PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv, rinv;
    Math<R>::inv2(r2, r2inv, rinv);
    R costj = dot3(e, d) * rinv;
    R qQr3 = R(0.5) * q * m * rinv * r2inv;
    R upot = qQr3 * (R(3) * costj * costj - R(1));
    R mpR = R(3) * upot * r2inv;
    R pT = R(6) * costj * qQr3 * rinv;
    R fac = costj * pT * rinv + mpR;
    o.f = {fac * d.x - pT * e.x, fac * d.y - pT * e.y, fac * d.z - pT * e.z};
    o.u = upot;
    V3<R> c = cross3(d, e);
    o.M2 = {pT * c.x, pT * c.y, pT * c.z};
}

// _loopBodyDipoleQuadrupole (VectorizedCellProcessor.cpp:526-648): first
// dipole (ei, p), second quadrupole (ej, m).
template <typename R>
__device__ __forceinline__ void body_dq(V3<R> d, V3<R> ei, R p, V3<R> ej, R m, PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv, rinv;
    Math<R>::inv2(r2, r2inv, rinv);
    R myqfac = R(1.5) * p * m * r2inv * r2inv;
    R costi = dot3(ei, d) * rinv, costj = dot3(ej, d) * rinv;
    R cosgij = dot3(ei, ej);
    R c5 = R(5) * costj * costj - R(1);
    R upot = myqfac * (R(2) * costj * cosgij - costi * c5);
    R myqfacr = myqfac * rinv;
    R mpR = R(4) * upot * r2inv;
    R mpTi = myqfacr * c5;
    R mpTj = myqfacr * R(2) * (R(5) * costi * costj - cosgij);
    R pG = myqfac * R(2) * costj;
    R fac = mpR - (costi * mpTi + costj * mpTj) * rinv;
    o.f = {fac * d.x + mpTi * ei.x + mpTj * ej.x, fac * d.y + mpTi * ei.y + mpTj * ej.y,
           fac * d.z + mpTi * ei.z + mpTj * ej.z};
    o.u = upot;
    V3<R> c12 = cross3(ei, ej), eixd = cross3(ei, d), ejxd = cross3(ej, d);
    o.M1 = {mpTi * eixd.x - pG * c12.x, mpTi * eixd.y - pG * c12.y, mpTi * eixd.z - pG * c12.z};
    o.M2 = {mpTj * ejxd.x + pG * c12.x, mpTj * ejxd.y + pG * c12.y, mpTj * ejxd.z + pG * c12.z};
}

// _loopBodyQuadrupole (VectorizedCellProcessor.cpp:650-792)
template <typename R>
__device__ __forceinline__ void body_qq(V3<R> d, V3<R> ei, R mi, V3<R> ej, R mj, PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv, rinv;
    Math<R>::inv2(r2, r2inv, rinv);
    R qfac = R(0.75) * rinv * (mi * mj) * (r2inv * r2inv);
    R costi = dot3(ei, d) * rinv, costj = dot3(ej, d) * rinv;
    R c2i = costi * costi, c2j = costj * costj;
    R cosgij = dot3(ei, ej);
    R term = cosgij - R(5) * (costi * costj);
    R upot = qfac * ((R(1) + R(2) * term * term) - (R(5) * (c2i + c2j) + R(15) * c2i * c2j));
    R mpR = R(5) * upot * r2inv;
    R part1 = qfac * R(10) * rinv;
    R part2 = R(2) * term;
    R mpTi = part1 * (costi + (part2 * costj + R(3) * costi * c2j));
    R mpTj = part1 * (costj + (part2 * costi + R(3) * costj * c2i));
    R pG = qfac * R(4) * term;
    R fac = mpR - (mpTi * costi + mpTj * costj) * rinv;
    o.f = {fac * d.x + mpTi * ei.x + mpTj * ej.x, fac * d.y + mpTi * ei.y + mpTj * ej.y,
           fac * d.z + mpTi * ei.z + mpTj * ej.z};
    o.u = upot;
    V3<R> c12 = cross3(ei, ej), eixd = cross3(ei, d), ejxd = cross3(ej, d);
    o.M1 = {mpTi * eixd.x - pG * c12.x, mpTi * eixd.y - pG * c12.y, mpTi * eixd.z - pG * c12.z};
    o.M2 = {mpTj * ejxd.x + pG * c12.x, mpTj * ejxd.y + pG * c12.y, mpTj * ejxd.z + pG * c12.z};
}

// ------------------------------------------------------------- site access

template <typename R> struct SiteView;

template <> struct SiteView<double> {
    static __device__ __forceinline__ V3<double> pos(const SiteArrays &A, int j) {
        return {A.x[j], A.y[j], A.z[j]};
    }
    static __device__ __forceinline__ V3<double> mol(const SiteArrays &A, int j) {
        return {A.mx[j], A.my[j], A.mz[j]};
    }
    static __device__ __forceinline__ V3<double> ori(const SiteArrays &A, int j) {
        return {A.ex[j], A.ey[j], A.ez[j]};
    }
    static __device__ __forceinline__ double par(const SiteArrays &A, int j) { return A.p[j]; }
};

template <> struct SiteView<float> {
    static __device__ __forceinline__ V3<float> pos(const SiteArrays &A, int j) {
        return {A.fx[j], A.fy[j], A.fz[j]};
    }
    static __device__ __forceinline__ V3<float> mol(const SiteArrays &A, int j) {
        return {A.fmx[j], A.fmy[j], A.fmz[j]};
    }
    static __device__ __forceinline__ V3<float> ori(const SiteArrays &A, int j) {
        return {A.fex[j], A.fey[j], A.fez[j]};
    }
    static __device__ __forceinline__ float par(const SiteArrays &A, int j) { return A.fp[j]; }
};

// Neighbour k (0..26, dz-major) of cell (cx,cy,cz): index of the cell that
// holds the image, the FP64 shift of that image (0 or +-L) and the unwrapped
// cell offset. ok = false for open boundaries outside the domain.
struct Nbr {
    int cell;
    bool ok;
    double s[3];
    int dd[3];
};

// nxs = 4 (two cells side by side in x): x offsets -1 .. 2 from the first.
__device__ __forceinline__ Nbr neighbour(const Geom &g, int cx, int cy, int cz, int k,
                                         int nxs = 3) {
    Nbr b;
    b.dd[0] = k % nxs - 1;
    b.dd[1] = (k / nxs) % 3 - 1;
    b.dd[2] = k / (3 * nxs) - 1;
    int c3[3] = {cx + b.dd[0], cy + b.dd[1], cz + b.dd[2]};
    b.ok = true;
    for (int d = 0; d < 3; ++d) {
        b.s[d] = 0.0;
        if (c3[d] < 0) {
            if (!g.pdim[d]) b.ok = false;
            c3[d] += g.n[d];
            b.s[d] = -g.L[d];
        } else if (c3[d] >= g.n[d]) {
            if (!g.pdim[d]) b.ok = false;
            c3[d] -= g.n[d];
            b.s[d] = g.L[d];
        }
    }
    b.cell = (c3[2] * g.n[1] + c3[1]) * g.n[0] + c3[0];
    return b;
}

// Neighbour index of tile entry idx (concatenated neighbour order): the
// largest k < nb with pre[k] <= idx.
__device__ __forceinline__ int tile_k(const int *pre, int idx, int nb = 27) {
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= idx) lo = mid; else hi = mid;
    }
    return lo;
}

template <typename R>
__device__ __forceinline__ V3<R> image_offset(const Geom &g, const Nbr &b);

template <> __device__ __forceinline__ V3<double> image_offset<double>(const Geom &, const Nbr &b) {
    return {b.s[0], b.s[1], b.s[2]};
}
template <> __device__ __forceinline__ V3<float> image_offset<float>(const Geom &g, const Nbr &b) {
    return {(float)(b.dd[0] * g.cl[0]), (float)(b.dd[1] * g.cl[1]), (float)(b.dd[2] * g.cl[2])};
}

struct Acc {
    double f[3], v[3], m[3];
    double u, vir, rf;
};

// i-site data held in registers for the whole cell
template <typename R> struct ISite {
    V3<R> pos, mol, e;
    R par;
    int ljid;
    int ljrow;          // ljid * nlj * 4: the row of the interleaved LJ table
    int idx;            // site index (FP64 molecule position for the cutoff re-check)
};

template <int TI, int TJ> struct Kind;  // 0 LJ,1 CC,2 CD,3 DD,4 CQ,5 DQ,6 QQ
template <> struct Kind<T_LJ, T_LJ> { static constexpr int k = 0; };
template <> struct Kind<T_C, T_C> { static constexpr int k = 1; };
template <> struct Kind<T_C, T_D> { static constexpr int k = 2; };
template <> struct Kind<T_D, T_C> { static constexpr int k = 2; };
template <> struct Kind<T_D, T_D> { static constexpr int k = 3; };
template <> struct Kind<T_C, T_Q> { static constexpr int k = 4; };
template <> struct Kind<T_Q, T_C> { static constexpr int k = 4; };
template <> struct Kind<T_D, T_Q> { static constexpr int k = 5; };
template <> struct Kind<T_Q, T_D> { static constexpr int k = 5; };
template <> struct Kind<T_Q, T_Q> { static constexpr int k = 6; };

// LDS tile: the j sites of the 27 neighbour images of one cell, positions
// already moved into the i cell's frame (FP64: + image shift, exactly the
// reference's halo copy r + L; FP32: + offset of the neighbour's corner).
// One 16-byte (FP32) or 32-byte (FP64) record per site and field, so a site
// is fetched with one or two ds_read_b128.

template <typename R> struct alignas(4 * sizeof(R)) Q4 {
    R x, y, z, w;
};

template <int TI, typename R, bool MOLSITE, bool MULTI> struct TileCfg {
    static constexpr bool ORI = TI != T_LJ;  // partners may be dipoles / quadrupoles
    static constexpr bool MOL = !MOLSITE;    // molecule positions differ from sites
    static constexpr bool LJID = TI == T_LJ && MULTI;  // several LJ centre types
    // Per-lane hit list entries (LIST) and sites per LDS chunk (CAP). The
    // kernels are bound by VALU issue, so the tile is sized for occupancy
    // first (160 KB of LDS per CU shared by the wavefronts of its 4 SIMDs,
    // next to the register limit of each instantiation) and chunk count
    // second; values measured best on the 4 M-molecule LJ slab (one-centre)
    // and the 1 M-molecule multi-site mixture (2CLJQ + water). The FP64
    // tiles that carry molecule positions or orientations keep 16-entry
    // lists for their larger records.
    static constexpr int LIST = (sizeof(R) == 8 && !(TI == T_LJ && MOLSITE)) ? 16 : 24;
    static constexpr int CAP =
        TI == T_LJ ? (MOLSITE ? (sizeof(R) == 4 ? 352 : 192) : (sizeof(R) == 4 ? 192 : 112))
        : TI == T_C ? (sizeof(R) == 4 ? 112 : 72)
        : TI == T_Q ? (sizeof(R) == 4 ? 114 : 104)
                    : (sizeof(R) == 4 ? 192 : 160);
    // Quadrupole i sites are few per cell (one per 2CLJQ molecule) against
    // a neighbourhood of many partner sites, so the kernel takes two cells
    // side by side in x at a time (JOINT): their neighbourhood of 4 x 3 x 3
    // cells is staged once for both, two thirds of what two 27-cell tiles
    // load, at the price of tests against the 9 cells that are beyond the
    // cutoff of one of the two (NB neighbour ranges at most).
    static constexpr bool JOINT = TI == T_Q;
    static constexpr int NB = JOINT ? 36 : 27;
    // Minimum waves per SIMD the kernel's registers are allocated for (1: no
    // bound, the compiler's choice). Tiles of at most 10 KB of LDS (lists
    // included) let 16 blocks share a CU, 4 waves per SIMD, when the
    // registers fit 128: the FP32 quadrupole kernel (114-site tile) and the
    // FP64 charge kernel (72 sites) get the bound, which makes the compiler
    // fit them with a few spilled values outside the pair loops; the FP64
    // multi-site LJ kernel (112 sites) fits without it, and the FP64
    // quadrupole kernel's 104-site tile (13 KB) reaches 3 waves per SIMD
    // with its 166 registers. Measured on the mixture against the larger
    // tiles at fewer waves: FP32 quadrupole kernel -6 %; FP64 LJ -3 %,
    // charge -4 %, quadrupole -9 %.
    static constexpr int WAVES =
        (TI == T_Q && sizeof(R) == 4) || (TI == T_C && sizeof(R) == 8) ? 4 : 1;
};

template <int TI, typename R, bool MOLSITE, bool MULTI> struct Tile {
    using C = TileCfg<TI, R, MOLSITE, MULTI>;
    // CAP slots of the largest record set (site, molecule, orientation,
    // head index and site count); each partner pass lays out only the records
    // its partner type needs (PassTile), so a pass with smaller records holds
    // more sites in the same bytes
    static constexpr int SLOT =
        (int)sizeof(Q4<R>) * (1 + C::MOL + C::ORI) + (C::MOL && TI != T_LJ ? 3 : 0);
    alignas(32) unsigned char raw[C::CAP * SLOT];
    unsigned short lst[C::LIST][64];  // per-lane hit lists (tile indices)
    int j0[C::NB], pre[C::NB + 1];
    double shift[C::NB][3];
};

// The records of one partner pass (partner type TJ) in a tile's slot bytes:
// p site position (w = charge / moment, LJ table column, or |x|^2 in FP32
// one-centre tiles), m molecule position (w = |m|^2 in FP32 tiles), e
// orientation (dipole and quadrupole partners), h first-site (head) indices
// and ns site counts at heads (charge partners, tested per molecule).
template <int TI, int TJ, typename R, bool MOLSITE, bool MULTI> struct PassTile {
    using C = TileCfg<TI, R, MOLSITE, MULTI>;
    using TT = Tile<TI, R, MOLSITE, MULTI>;
    static constexpr bool HEADS = C::MOL && TJ == T_C && TI != T_LJ;
    static constexpr bool ORIJ = TI != T_LJ && TJ >= T_D;
    static constexpr int REC = 1 + C::MOL + ORIJ;
    static constexpr int CAP =
        (C::CAP * TT::SLOT) / ((int)sizeof(Q4<R>) * REC + (HEADS ? 3 : 0));
    static_assert(CAP >= C::CAP && CAP < 65536, "pass capacity");
    Q4<R> *p, *m, *e;
    unsigned short *h;
    unsigned char *ns;
    __device__ __forceinline__ explicit PassTile(TT &t) {
        p = reinterpret_cast<Q4<R> *>(t.raw);
        m = p + (C::MOL ? CAP : 0);
        e = p + CAP * (1 + C::MOL);
        h = reinterpret_cast<unsigned short *>(p + CAP * REC);
        ns = reinterpret_cast<unsigned char *>(h + (HEADS ? CAP : 0));
    }
};

template <typename R>
__device__ __forceinline__ V3<R> tile_offset(const Geom &g, const double (*shift)[3], int k,
                                             int nxs);
template <>
__device__ __forceinline__ V3<double> tile_offset<double>(const Geom &, const double (*shift)[3],
                                                          int k, int) {
    return {shift[k][0], shift[k][1], shift[k][2]};
}
template <>
__device__ __forceinline__ V3<float> tile_offset<float>(const Geom &g, const double (*)[3], int k,
                                                        int nxs) {
    return {(float)((k % nxs - 1) * g.cl[0]), (float)((k / nxs % 3 - 1) * g.cl[1]),
            (float)((k / (3 * nxs) - 1) * g.cl[2])};
}

}  // namespace mdx
