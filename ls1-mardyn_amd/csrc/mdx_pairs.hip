// mdx_pairs.hip -- site-pair traversal kernels (gfx950, wave64, no MFMA).
//
// The ten site-pair loops of VectorizedCellProcessor::_calculatePairs
// (VectorizedCellProcessor.cpp:794-2730) with the seven bodies _loopBodyLJ ..
// _loopBodyQuadrupole (VectorizedCellProcessor.cpp:171-792), evaluated owner-
// computes: kernel k_pairs<TI> runs one wavefront per cell for the sites of
// type TI and accumulates force, virial and torque of those sites only.
// The cutoff is tested on molecule centres (calcDistLookup,
// SIMD_VectorizedCellProcessorHelpers.h:427-472): LJ against rcLJ^2, the
// electrostatic types against rc^2, strictly less; inside the own cell a pair
// at molecule distance 0 (the same molecule) is skipped (SingleCellPolicy_).
//
// Precision forms (template R):
//   double : all arithmetic FP64, absolute positions, j images shifted by +-L
//            exactly as the reference's halo copies are (r + L).
//   float  : pair arithmetic FP32 on positions relative to the owning cell's
//            corner (|x| <= ~cell length, so FP32 keeps ~1e-7 absolute
//            accuracy at any box size); the cutoff decision is re-made in FP64
//            when the FP32 r^2 lies within 1e-5 relative of rc^2 so the pair
//            set equals the FP64 one; every per-pair result is converted to
//            FP64 before it is summed (RealAccumVec in the reference's MIXED
//            build, RealAccumVecSPDP.h).
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
    // NOVEC semantics: IEEE 1/x and sqrt(1/r^2) (RealVecDouble.h:414-467)
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double rinv(double r2, double r2inv) { return sqrt(r2inv); }
};
template <> struct Math<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rinv(float r2, float) { return __builtin_amdgcn_rsqf(r2); }
};

template <typename R> struct PO {
    V3<R> f, M1, M2;
    R u, rf;
};

// _loopBodyLJ (VectorizedCellProcessor.cpp:171-224); u = 6 * U_LJ + shift6
template <typename R>
__device__ __forceinline__ void body_lj(V3<R> d, R eps24, R sig2, R shift6, PO<R> &o) {
    R r2inv = Math<R>::rcp(dot3(d, d));
    R lj2 = sig2 * r2inv, lj6 = lj2 * lj2 * lj2, lj12 = lj6 * lj6;
    R lj12m6 = lj12 - lj6;
    R scale = eps24 * r2inv * (lj12 + lj12m6);
    o.f = {d.x * scale, d.y * scale, d.z * scale};
    o.u = eps24 * lj12m6 + shift6;
}

// _loopBodyCharge (VectorizedCellProcessor.cpp:227-277)
template <typename R> __device__ __forceinline__ void body_cc(V3<R> d, R q1, R q2, PO<R> &o) {
    R r2 = dot3(d, d);
    R r2inv = Math<R>::rcp(r2);
    R rinv = Math<R>::rinv(r2, r2inv);
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
    R r2inv = Math<R>::rcp(r2);
    R rinv = Math<R>::rinv(r2, r2inv);
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
    R r2inv = Math<R>::rcp(r2);
    R rinv = Math<R>::rinv(r2, r2inv);
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
    R r2inv = Math<R>::rcp(r2);
    R rinv = Math<R>::rinv(r2, r2inv);
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
    R r2inv = Math<R>::rcp(r2);
    R rinv = Math<R>::rinv(r2, r2inv);
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
    R r2inv = Math<R>::rcp(r2);
    R rinv = Math<R>::rinv(r2, r2inv);
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

__device__ __forceinline__ Nbr neighbour(const Geom &g, int cx, int cy, int cz, int k) {
    Nbr b;
    b.dd[0] = k % 3 - 1;
    b.dd[1] = (k / 3) % 3 - 1;
    b.dd[2] = k / 9 - 1;
    int c3[3] = {cx + b.dd[0], cy + b.dd[1], cz + b.dd[2]};
    b.ok = true;
    for (int d = 0; d < 3; ++d) {
        b.s[d] = 0.0;
        if (c3[d] < 0) {
            if (!g.periodic) b.ok = false;
            c3[d] += g.n[d];
            b.s[d] = -g.L[d];
        } else if (c3[d] >= g.n[d]) {
            if (!g.periodic) b.ok = false;
            c3[d] -= g.n[d];
            b.s[d] = g.L[d];
        }
    }
    b.cell = (c3[2] * g.n[1] + c3[1]) * g.n[0] + c3[0];
    return b;
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
    double mx, my, mz;  // FP64 molecule position for the cutoff re-check
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
constexpr int kList = 32;  // per-lane hit list entries

template <int TI, typename R, bool MOLSITE> struct TileCfg {
    static constexpr bool ORI = TI != T_LJ;   // partners may be dipoles / quadrupoles
    static constexpr bool MOL = !MOLSITE;     // molecule positions differ from sites
    static constexpr int CAP = (TI == T_LJ) ? 384 : 256;
};

template <int TI, typename R, bool MOLSITE> struct Tile {
    using C = TileCfg<TI, R, MOLSITE>;
    R px[C::CAP], py[C::CAP], pz[C::CAP];
    R mx[C::MOL ? C::CAP : 1], my[C::MOL ? C::CAP : 1], mz[C::MOL ? C::CAP : 1];
    R ex[C::ORI ? C::CAP : 1], ey[C::ORI ? C::CAP : 1], ez[C::ORI ? C::CAP : 1];
    R par[C::ORI ? C::CAP : 1];
    int ljid[TI == T_LJ ? C::CAP : 1];
    unsigned gj[sizeof(R) == 4 ? C::CAP : 1];  // j | k << 27 for the FP64 re-check
    unsigned short lst[kList][64];
    int j0[27], pre[28];
    double shift[27][3];
};

template <typename R>
__device__ __forceinline__ V3<R> tile_offset(const Geom &g, const double (*shift)[3], int k);
template <>
__device__ __forceinline__ V3<double> tile_offset<double>(const Geom &, const double (*shift)[3],
                                                          int k) {
    return {shift[k][0], shift[k][1], shift[k][2]};
}
template <>
__device__ __forceinline__ V3<float> tile_offset<float>(const Geom &g, const double (*)[3], int k) {
    return {(float)((k % 3 - 1) * g.cl[0]), (float)((k / 3 % 3 - 1) * g.cl[1]),
            (float)((k / 9 - 1) * g.cl[2])};
}

// One site pair seen from site i (tile entry s): force on i, virial
// (m_i - m_j) * f_i, torque on i; roles swapped where the reference swaps.
template <int TI, int TJ, typename R, bool MOLSITE>
__device__ __forceinline__ void eval_pair(const PairArgs &a, const ISite<R> &I,
                                          const Tile<TI, R, MOLSITE> &T, int s, R eps24,
                                          R sig2, R shift6, Acc &acc) {
    V3<R> d = {I.pos.x - T.px[s], I.pos.y - T.py[s], I.pos.z - T.pz[s]};
    V3<R> dm = d;
    if constexpr (!MOLSITE) dm = {I.mol.x - T.mx[s], I.mol.y - T.my[s], I.mol.z - T.mz[s]};
    PO<R> o;
    o.rf = R(0);
    R sign = R(1);
    V3<R> tq = {R(0), R(0), R(0)};
    if constexpr (TI == T_LJ) {
        if (a.pp.nlj > 1) {
            int k = I.ljid * a.pp.nlj + T.ljid[s];
            if constexpr (sizeof(R) == 8) {
                eps24 = a.pp.eps24[k]; sig2 = a.pp.sig2[k]; shift6 = a.pp.shift6[k];
            } else {
                eps24 = a.pp.feps24[k]; sig2 = a.pp.fsig2[k]; shift6 = a.pp.fshift6[k];
            }
        }
        body_lj<R>(d, eps24, sig2, shift6, o);
    }
    if constexpr (TI != T_LJ) {
        V3<R> ej = {R(0), R(0), R(0)};
        if constexpr (TJ >= T_D) ej = {T.ex[s], T.ey[s], T.ez[s]};
        R pj = T.par[s];
        if constexpr (TI == T_C && TJ == T_C) {
            body_cc<R>(d, I.par, pj, o);
        } else if constexpr (TI == T_C && TJ == T_D) {
            body_cd<R>(d, I.par, ej, pj, o);
        } else if constexpr (TI == T_C && TJ == T_Q) {
            body_cq<R>(d, I.par, ej, pj, o);
        } else if constexpr (TI == T_D && TJ == T_C) {
            V3<R> nd = {-d.x, -d.y, -d.z};
            body_cd<R>(nd, pj, I.e, I.par, o);
            sign = R(-1);
            tq = o.M2;
        } else if constexpr (TI == T_D && TJ == T_D) {
            body_dd<R>(d, I.e, I.par, ej, pj, (R)a.g.epsRFInvrc3, o);
            tq = o.M1;
        } else if constexpr (TI == T_D && TJ == T_Q) {
            body_dq<R>(d, I.e, I.par, ej, pj, o);
            tq = o.M1;
        } else if constexpr (TI == T_Q && TJ == T_C) {
            V3<R> nd = {-d.x, -d.y, -d.z};
            body_cq<R>(nd, pj, I.e, I.par, o);
            sign = R(-1);
            tq = o.M2;
        } else if constexpr (TI == T_Q && TJ == T_D) {
            V3<R> nd = {-d.x, -d.y, -d.z};
            body_dq<R>(nd, ej, pj, I.e, I.par, o);
            sign = R(-1);
            tq = o.M2;
        } else if constexpr (TI == T_Q && TJ == T_Q) {
            body_qq<R>(d, I.e, I.par, ej, pj, o);
            tq = o.M1;
        }
    }
    V3<R> f = {sign * o.f.x, sign * o.f.y, sign * o.f.z};
    acc.f[0] += (double)f.x;
    acc.f[1] += (double)f.y;
    acc.f[2] += (double)f.z;
    double vx = (double)(dm.x * f.x), vy = (double)(dm.y * f.y), vz = (double)(dm.z * f.z);
    acc.v[0] += vx;
    acc.v[1] += vy;
    acc.v[2] += vz;
    acc.vir += vx + vy + vz;
    acc.u += (double)o.u;
    if constexpr (TI == T_D && TJ == T_D) acc.rf += (double)o.rf;
    if constexpr (TI >= T_D) {
        acc.m[0] += (double)tq.x;
        acc.m[1] += (double)tq.y;
        acc.m[2] += (double)tq.z;
    }
}

// All pairs of the lane's i site with the j sites of type TJ in the 27
// neighbour images: (1) lanes 0..26 fetch the neighbour ranges in one round,
// (2) the neighbourhood is staged into LDS in chunks of CAP sites with
// coalesced, independent loads, (3) stage A tests the molecule-centre cutoff
// from LDS and keeps per-lane hit lists, (4) stage B evaluates the bodies.
template <int TI, int TJ, typename R, bool MOLSITE, bool COUNT>
__device__ __forceinline__ void partner_pass(const PairArgs &a, const ISite<R> &I, bool valid,
                                             int cx, int cy, int cz, int ph, int P, int lane,
                                             Tile<TI, R, MOLSITE> &T, Acc &acc, R eps24, R sig2,
                                             R shift6, unsigned long long &ntest) {
    using C = TileCfg<TI, R, MOLSITE>;
    const SiteArrays &J = a.s[TJ];
    const double rc2 = (TI == T_LJ) ? a.g.rclj2 : a.g.rc2;
    const R rc2r = (R)rc2;
    unsigned long long nhit = 0;
    int nk = 0;
    if (lane < 27) {
        Nbr b = neighbour(a.g, cx, cy, cz, lane);
        int j0 = 0;
        if (b.ok) {
            j0 = J.cstart[b.cell];
            nk = J.cstart[b.cell + 1] - j0;
        }
        T.j0[lane] = j0;
        T.shift[lane][0] = b.s[0];
        T.shift[lane][1] = b.s[1];
        T.shift[lane][2] = b.s[2];
    }
    int incl = nk;
    for (int o = 1; o < 32; o <<= 1) {
        int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane < 27) T.pre[lane + 1] = incl;
    if (lane == 0) T.pre[0] = 0;
    const int total = __shfl(incl, 26, 64);
    __syncthreads();
    const int own_b = T.pre[13], own_e = T.pre[14];
    for (int base = 0; base < total; base += C::CAP) {
        const int nst = min(C::CAP, total - base);
        for (int s = lane; s < nst; s += 64) {
            const int idx = base + s;
            int lo = 0, hi = 27;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (T.pre[mid] <= idx) lo = mid; else hi = mid;
            }
            const int k = lo;
            const int j = T.j0[k] + (idx - T.pre[k]);
            const V3<R> off = tile_offset<R>(a.g, T.shift, k);
            V3<R> p = SiteView<R>::pos(J, j);
            T.px[s] = p.x + off.x;
            T.py[s] = p.y + off.y;
            T.pz[s] = p.z + off.z;
            if constexpr (C::MOL) {
                V3<R> m = SiteView<R>::mol(J, j);
                T.mx[s] = m.x + off.x;
                T.my[s] = m.y + off.y;
                T.mz[s] = m.z + off.z;
            }
            if constexpr (TI == T_LJ) {
                T.ljid[s] = J.ljid[j];
            } else {
                T.par[s] = SiteView<R>::par(J, j);
                if constexpr (TJ >= T_D) {
                    V3<R> e = SiteView<R>::ori(J, j);
                    T.ex[s] = e.x;
                    T.ey[s] = e.y;
                    T.ez[s] = e.z;
                }
            }
            if constexpr (sizeof(R) == 4) T.gj[s] = (unsigned)j | ((unsigned)k << 27);
        }
        __syncthreads();
        // own cell (same molecule exclusion) inside this chunk
        const int sb = own_b - base, se = own_e - base;
        int cnt = 0, ub = 0;
        auto drain = [&]() {
            for (int e = 0;; ++e) {
                bool act = e < cnt;
                if (!__any(act)) break;
                if (act) {
                    eval_pair<TI, TJ, R, MOLSITE>(a, I, T, (int)T.lst[e][lane], eps24, sig2,
                                                  shift6, acc);
                    if (COUNT) ++nhit;
                }
            }
            cnt = 0;
            ub = 0;
        };
        const int per = (nst + P - 1) / P;
        for (int m = 0; m < per; ++m) {
            if (ub == kList) drain();
            ++ub;
            const int s = ph + m * P;
            if (valid && s < nst) {
                R dx, dy, dz;
                if constexpr (C::MOL) {
                    dx = I.mol.x - T.mx[s]; dy = I.mol.y - T.my[s]; dz = I.mol.z - T.mz[s];
                } else {
                    dx = I.pos.x - T.px[s]; dy = I.pos.y - T.py[s]; dz = I.pos.z - T.pz[s];
                }
                R r2 = dx * dx + dy * dy + dz * dz;
                bool hit = r2 < rc2r;
                bool own = s >= sb && s < se;
                if constexpr (sizeof(R) == 4) {
                    if (fabsf(r2 - rc2r) <= 1e-5f * rc2r) {
                        unsigned gj = T.gj[s];
                        int j = (int)(gj & 0x07FFFFFFu), k = (int)(gj >> 27);
                        double ex = I.mx - (J.mx[j] + T.shift[k][0]);
                        double ey = I.my - (J.my[j] + T.shift[k][1]);
                        double ez = I.mz - (J.mz[j] + T.shift[k][2]);
                        double r2d = ex * ex + ey * ey + ez * ez;
                        hit = r2d < rc2;
                        if (own && r2d == 0.0) hit = false;
                    }
                }
                if (own && r2 == R(0)) hit = false;
                if (COUNT) ++ntest;
                if (hit) {
                    T.lst[cnt][lane] = (unsigned short)s;
                    ++cnt;
                }
            }
        }
        drain();
        __syncthreads();
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) nhit += __shfl_xor(nhit, o, 64);
        if (lane == 0) atomicAdd(&a.counters[1 + Kind<TI, TJ>::k], nhit);
    }
}

__device__ __forceinline__ double shfl_xor_d(double v, int o) { return __shfl_xor(v, o, 64); }

// One 64-lane block (one wavefront) per cell.
template <int TI, typename R, bool COUNT, bool MOLSITE>
__global__ __launch_bounds__(64) void k_pairs(PairArgs a) {
    __shared__ Tile<TI, R, MOLSITE> T;
    const int lane = threadIdx.x;
    // XCD-aware order: hardware block b runs on XCD b % 8; each XCD gets a
    // contiguous range of cells so neighbour data is shared in its L2.
    const int hb = blockIdx.x;
    const int cell = (hb & 7) * a.nblk_per_xcd + (hb >> 3);
    double wmac[3] = {0, 0, 0};
    unsigned long long ntest = 0;
    if (cell < a.ncell) {
        const SiteArrays &S = a.s[TI];
        const int i_begin = S.cstart[cell], i_end = S.cstart[cell + 1];
        const int cx = cell % a.g.n[0], cy = (cell / a.g.n[0]) % a.g.n[1],
                  cz = cell / (a.g.n[0] * a.g.n[1]);
        R eps24 = R(0), sig2 = R(0), shift6 = R(0);
        if constexpr (TI == T_LJ) {
            if constexpr (sizeof(R) == 8) {
                eps24 = a.pp.eps24[0]; sig2 = a.pp.sig2[0]; shift6 = a.pp.shift6[0];
            } else {
                eps24 = a.pp.feps24[0]; sig2 = a.pp.fsig2[0]; shift6 = a.pp.fshift6[0];
            }
        }
        for (int chunk = i_begin; chunk < i_end; chunk += 64) {
            const int nI = min(64, i_end - chunk);
            int lg = 0;
            while ((1 << lg) < nI) ++lg;
            const int Ipad = 1 << lg, P = 64 >> lg;
            const int il = lane & (Ipad - 1), ph = lane >> lg;
            const bool valid = il < nI;
            const int i = chunk + (valid ? il : 0);
            ISite<R> I;
            I.pos = SiteView<R>::pos(S, i);
            I.mol = MOLSITE ? I.pos : SiteView<R>::mol(S, i);
            I.mx = S.mx[i];
            I.my = S.my[i];
            I.mz = S.mz[i];
            if constexpr (TI >= T_D) I.e = SiteView<R>::ori(S, i);
            if constexpr (TI == T_LJ)
                I.ljid = S.ljid[i];
            else
                I.par = SiteView<R>::par(S, i);
            Acc acc;
            for (int k = 0; k < 3; ++k) acc.f[k] = acc.v[k] = acc.m[k] = 0.0;
            acc.u = acc.vir = acc.rf = 0.0;
            if constexpr (TI == T_LJ) {
                partner_pass<T_LJ, T_LJ, R, MOLSITE, COUNT>(a, I, valid, cx, cy, cz, ph, P, lane, T,
                                                            acc, eps24, sig2, shift6, ntest);
            } else {
                if (a.present[T_C])
                    partner_pass<TI, T_C, R, MOLSITE, COUNT>(a, I, valid, cx, cy, cz, ph, P, lane, T,
                                                             acc, eps24, sig2, shift6, ntest);
                if (a.present[T_D])
                    partner_pass<TI, T_D, R, MOLSITE, COUNT>(a, I, valid, cx, cy, cz, ph, P, lane, T,
                                                             acc, eps24, sig2, shift6, ntest);
                if (a.present[T_Q])
                    partner_pass<TI, T_Q, R, MOLSITE, COUNT>(a, I, valid, cx, cy, cz, ph, P, lane, T,
                                                             acc, eps24, sig2, shift6, ntest);
            }
            for (int o = Ipad; o < 64; o <<= 1) {
                for (int k = 0; k < 3; ++k) {
                    acc.f[k] += shfl_xor_d(acc.f[k], o);
                    acc.v[k] += shfl_xor_d(acc.v[k], o);
                    if constexpr (TI >= T_D) acc.m[k] += shfl_xor_d(acc.m[k], o);
                }
            }
            if (valid && ph == 0) {
                S.Fx[i] = acc.f[0]; S.Fy[i] = acc.f[1]; S.Fz[i] = acc.f[2];
                S.Vx[i] = acc.v[0]; S.Vy[i] = acc.v[1]; S.Vz[i] = acc.v[2];
                if constexpr (TI >= T_D) {
                    S.Mx[i] = acc.m[0]; S.My[i] = acc.m[1]; S.Mz[i] = acc.m[2];
                }
            }
            wmac[0] += acc.u;
            wmac[1] += acc.vir;
            wmac[2] += acc.rf;
        }
    }
    for (int k = 0; k < 3; ++k)
        for (int o = 32; o > 0; o >>= 1) wmac[k] += shfl_xor_d(wmac[k], o);
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) ntest += __shfl_xor(ntest, o, 64);
        if (lane == 0) atomicAdd(&a.counters[0], ntest);
    }
    if (lane == 0) {
        double *o = a.macro_part + 4 * hb;
        o[0] = (TI == T_LJ) ? wmac[0] : 0.0;
        o[1] = (TI == T_LJ) ? 0.0 : wmac[0];
        o[2] = wmac[1];
        o[3] = wmac[2];
    }
}


// The cell pairs of the traversal on ls1's halo-extended grid: thread
// (cell, k) emits (ls1 index of cell, ls1 index of its neighbour k), i.e. the
// halo cell when the neighbour is a periodic image; -1 when the neighbour
// does not exist (open boundaries).
__global__ void k_enum_pairs(Geom g, int ncell, int *out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncell * 27) return;
    int c = (int)(t / 27), k = (int)(t % 27);
    int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
    Nbr b = neighbour(g, cx, cy, cz, k);
    int e1 = ((cz + 1) * g.cpd[1] + (cy + 1)) * g.cpd[0] + (cx + 1);
    int e2 = ((cz + 1 + b.dd[2]) * g.cpd[1] + (cy + 1 + b.dd[1])) * g.cpd[0] + (cx + 1 + b.dd[0]);
    out[2 * t] = b.ok ? e1 : -1;
    out[2 * t + 1] = b.ok ? e2 : -1;
}

// Host-side launcher table.
PairKernel pair_kernel(int ti, bool fp32, bool count, bool molsite) {
#define MDX_SEL2(T, R, C)                                            \
    return molsite ? k_pairs<T, R, C, true> : k_pairs<T, R, C, false>;
#define MDX_SEL(T)                                                   \
    if (ti == T) {                                                   \
        if (fp32) {                                                  \
            if (count) { MDX_SEL2(T, float, true) }                  \
            MDX_SEL2(T, float, false)                                \
        }                                                            \
        if (count) { MDX_SEL2(T, double, true) }                     \
        MDX_SEL2(T, double, false)                                   \
    }
    MDX_SEL(T_LJ)
    MDX_SEL(T_C)
    MDX_SEL(T_D)
    MDX_SEL(T_Q)
#undef MDX_SEL
#undef MDX_SEL2
    return nullptr;
}

}  // namespace mdx
