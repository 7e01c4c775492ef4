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

constexpr int kCap = 48;   // per-lane hit list entries in LDS
constexpr int kWaves = kWavesPerBlock;

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

// One site pair seen from site i: force on i, virial (m_i - m_j) x f_i,
// torque on i (roles swapped where the reference swaps them).
template <int TI, int TJ, typename R>
__device__ __forceinline__ void eval_pair(const PairArgs &a, const ISite<R> &I, int j,
                                          V3<R> off, Acc &acc) {
    const SiteArrays &J = a.s[TJ];
    V3<R> pj = SiteView<R>::pos(J, j), mj = SiteView<R>::mol(J, j);
    pj = {pj.x + off.x, pj.y + off.y, pj.z + off.z};
    mj = {mj.x + off.x, mj.y + off.y, mj.z + off.z};
    V3<R> d = {I.pos.x - pj.x, I.pos.y - pj.y, I.pos.z - pj.z};
    V3<R> dm = {I.mol.x - mj.x, I.mol.y - mj.y, I.mol.z - mj.z};
    PO<R> o;
    o.rf = R(0);
    R sign = R(1);
    V3<R> tq = {R(0), R(0), R(0)};
    if constexpr (TI == T_LJ) {
        int k = I.ljid * a.pp.nlj + J.ljid[j];
        if constexpr (sizeof(R) == 8)
            body_lj<R>(d, a.pp.eps24[k], a.pp.sig2[k], a.pp.shift6[k], o);
        else
            body_lj<R>(d, a.pp.feps24[k], a.pp.fsig2[k], a.pp.fshift6[k], o);
    } else if constexpr (TI == T_C && TJ == T_C) {
        body_cc<R>(d, I.par, SiteView<R>::par(J, j), o);
    } else if constexpr (TI == T_C && TJ == T_D) {
        body_cd<R>(d, I.par, SiteView<R>::ori(J, j), SiteView<R>::par(J, j), o);
    } else if constexpr (TI == T_C && TJ == T_Q) {
        body_cq<R>(d, I.par, SiteView<R>::ori(J, j), SiteView<R>::par(J, j), o);
    } else if constexpr (TI == T_D && TJ == T_C) {
        V3<R> nd = {-d.x, -d.y, -d.z};
        body_cd<R>(nd, SiteView<R>::par(J, j), I.e, I.par, o);
        sign = R(-1);
        tq = o.M2;
    } else if constexpr (TI == T_D && TJ == T_D) {
        body_dd<R>(d, I.e, I.par, SiteView<R>::ori(J, j), SiteView<R>::par(J, j),
                   (R)a.g.epsRFInvrc3, o);
        tq = o.M1;
    } else if constexpr (TI == T_D && TJ == T_Q) {
        body_dq<R>(d, I.e, I.par, SiteView<R>::ori(J, j), SiteView<R>::par(J, j), o);
        tq = o.M1;
    } else if constexpr (TI == T_Q && TJ == T_C) {
        V3<R> nd = {-d.x, -d.y, -d.z};
        body_cq<R>(nd, SiteView<R>::par(J, j), I.e, I.par, o);
        sign = R(-1);
        tq = o.M2;
    } else if constexpr (TI == T_Q && TJ == T_D) {
        V3<R> nd = {-d.x, -d.y, -d.z};
        body_dq<R>(nd, SiteView<R>::ori(J, j), SiteView<R>::par(J, j), I.e, I.par, o);
        sign = R(-1);
        tq = o.M2;
    } else if constexpr (TI == T_Q && TJ == T_Q) {
        body_qq<R>(d, I.e, I.par, SiteView<R>::ori(J, j), SiteView<R>::par(J, j), o);
        tq = o.M1;
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

template <int TI, int TJ, typename R, bool COUNT>
__device__ __forceinline__ void partner_pass(const PairArgs &a, const ISite<R> &I, bool valid,
                                             int il_cell, int cx, int cy, int cz, int ph, int P,
                                             int lane, unsigned (*lst)[64], Acc &acc,
                                             unsigned long long &ntest) {
    unsigned long long nhit = 0;
    const SiteArrays &J = a.s[TJ];
    const double rc2 = (TI == T_LJ) ? a.g.rclj2 : a.g.rc2;
    const R rc2r = (R)rc2;
    int cnt = 0, ub = 0;
    auto drain = [&]() {
        for (int e = 0;; ++e) {
            bool act = e < cnt;
            if (!__any(act)) break;
            if (act) {
                unsigned ent = lst[e][lane];
                int j = (int)(ent & 0x07FFFFFFu);
                int k = (int)(ent >> 27);
                Nbr b = neighbour(a.g, cx, cy, cz, k);
                eval_pair<TI, TJ, R>(a, I, j, image_offset<R>(a.g, b), acc);
                if (COUNT) ++nhit;
            }
        }
        cnt = 0;
        ub = 0;
    };
    for (int k = 0; k < 27; ++k) {
        Nbr b = neighbour(a.g, cx, cy, cz, k);
        if (!b.ok) continue;
        int j0 = J.cstart[b.cell];
        int nB = J.cstart[b.cell + 1] - j0;
        if (nB == 0) continue;
        V3<R> off = image_offset<R>(a.g, b);
        int per = (nB + P - 1) / P;
        for (int m = 0; m < per; ++m) {
            if (ub == kCap) drain();
            ++ub;
            int jj = ph + m * P;
            bool hit = false;
            if (valid && jj < nB) {
                int j = j0 + jj;
                V3<R> mj = SiteView<R>::mol(J, j);
                R dx = I.mol.x - (mj.x + off.x), dy = I.mol.y - (mj.y + off.y),
                  dz = I.mol.z - (mj.z + off.z);
                R r2 = dx * dx + dy * dy + dz * dz;
                hit = r2 < rc2r;
                if constexpr (sizeof(R) == 4) {
                    if (fabsf(r2 - rc2r) <= 1e-5f * rc2r) {
                        double ex = I.mx - (J.mx[j] + b.s[0]);
                        double ey = I.my - (J.my[j] + b.s[1]);
                        double ez = I.mz - (J.mz[j] + b.s[2]);
                        double r2d = ex * ex + ey * ey + ez * ez;
                        hit = r2d < rc2;
                        if (k == 13 && r2d == 0.0) hit = false;
                    }
                }
                if (k == 13 && r2 == R(0)) hit = false;
                if (COUNT) ++ntest;
                if (hit) {
                    lst[cnt][lane] = (unsigned)j | ((unsigned)k << 27);
                    ++cnt;
                }
            }
        }
    }
    drain();
    (void)il_cell;
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) nhit += __shfl_xor(nhit, o, 64);
        if (lane == 0) atomicAdd(&a.counters[1 + Kind<TI, TJ>::k], nhit);
    }
}

__device__ __forceinline__ double shfl_xor_d(double v, int o) { return __shfl_xor(v, o, 64); }

template <int TI, typename R, bool COUNT>
__global__ __launch_bounds__(256) void k_pairs(PairArgs a) {
    __shared__ unsigned lst[kWaves][kCap][64];
    __shared__ double red[kWaves][4];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    // XCD-aware block order: hardware block b runs on XCD b % 8; give each XCD
    // a contiguous range of cells so neighbour data is shared in its L2.
    const int hb = blockIdx.x;
    const int lb = (hb & 7) * a.nblk_per_xcd + (hb >> 3);
    const int cell = lb * kWaves + wv;
    double wmac[3] = {0, 0, 0};
    unsigned long long ntest = 0;
    if (lb < a.nblocks && cell < a.ncell) {
        const SiteArrays &S = a.s[TI];
        const int i_begin = S.cstart[cell], i_end = S.cstart[cell + 1];
        const int cx = cell % a.g.n[0], cy = (cell / a.g.n[0]) % a.g.n[1],
                  cz = cell / (a.g.n[0] * a.g.n[1]);
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
            I.mol = SiteView<R>::mol(S, i);
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
            unsigned (*L)[64] = lst[wv];
            if constexpr (TI == T_LJ) {
                partner_pass<T_LJ, T_LJ, R, COUNT>(a, I, valid, il, cx, cy, cz, ph, P, lane, L, acc,
                                                  ntest);
            } else {
                if (a.present[T_C])
                    partner_pass<TI, T_C, R, COUNT>(a, I, valid, il, cx, cy, cz, ph, P, lane, L,
                                                    acc, ntest);
                if (a.present[T_D])
                    partner_pass<TI, T_D, R, COUNT>(a, I, valid, il, cx, cy, cz, ph, P, lane, L,
                                                    acc, ntest);
                if (a.present[T_Q])
                    partner_pass<TI, T_Q, R, COUNT>(a, I, valid, il, cx, cy, cz, ph, P, lane, L,
                                                    acc, ntest);
            }
            // fold the j phases of each i site
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
    for (int k = 0; k < 3; ++k) {
        for (int o = 32; o > 0; o >>= 1) wmac[k] += shfl_xor_d(wmac[k], o);
    }
    if (lane == 0) {
        red[wv][0] = wmac[0];
        red[wv][1] = wmac[1];
        red[wv][2] = wmac[2];
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) ntest += __shfl_xor(ntest, o, 64);
        if (lane == 0) atomicAdd(&a.counters[0], ntest);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double u = 0, vir = 0, rf = 0;
        for (int k = 0; k < kWaves; ++k) {
            u += red[k][0];
            vir += red[k][1];
            rf += red[k][2];
        }
        double *o = a.macro_part + 4 * hb;
        o[0] = (TI == T_LJ) ? u : 0.0;
        o[1] = (TI == T_LJ) ? 0.0 : u;
        o[2] = vir;
        o[3] = rf;
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
PairKernel pair_kernel(int ti, bool fp32, bool count) {
#define MDX_SEL(T)                                                            \
    if (ti == T) {                                                            \
        if (fp32) return count ? k_pairs<T, float, true> : k_pairs<T, float, false>; \
        return count ? k_pairs<T, double, true> : k_pairs<T, double, false>;  \
    }
    MDX_SEL(T_LJ)
    MDX_SEL(T_C)
    MDX_SEL(T_D)
    MDX_SEL(T_Q)
#undef MDX_SEL
    return nullptr;
}

}  // namespace mdx
