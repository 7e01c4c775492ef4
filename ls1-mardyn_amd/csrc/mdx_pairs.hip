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
//            set equals the FP64 one; the pair terms of one drained hit list
//            (at most a list length of partners per lane) are summed in FP32
//            and added to the site's FP64 accumulators once per drain, so
//            every per-site and per-molecule sum across drains, cells and
//            partner types is FP64 (the reference's MIXED build accumulates
//            in RealAccumVec, RealAccumVecSPDP.h).
#include "mdx_pair_common.h"

namespace mdx {

template <typename R> __device__ __forceinline__ V3<R> v3(const Q4<R> &q) { return {q.x, q.y, q.z}; }

// FP32 form: partial sums of one drained hit list, folded into the FP64
// accumulators by flush().
struct Part32 {
    float f[3], v[3], m[3], u, rf;
    __device__ __forceinline__ void clear() {
        for (int k = 0; k < 3; ++k) f[k] = v[k] = m[k] = 0.f;
        u = rf = 0.f;
    }
    __device__ __forceinline__ void flush(Acc &a) const {
        for (int k = 0; k < 3; ++k) {
            a.f[k] += (double)f[k];
            a.v[k] += (double)v[k];
            a.m[k] += (double)m[k];
        }
        a.u += (double)u;
        a.rf += (double)rf;
    }
};

// LJ pair seen from site i, evaluated on every lane of the wave: a lane
// whose entry is not accepted (ok false) runs the body at a huge r^2 (FP32:
// +inf; FP64: 1e300, which keeps the Newton steps of 1/r^2 finite), where
// force and virial come out exactly 0, and adds no energy; u = 6 U_LJ +
// shift6 (_loopBodyLJ, VectorizedCellProcessor.cpp:171-224). FP64 form: into
// acc; FP32 form: into the drain's FP32 partial sums pa.
template <typename R, bool MOLSITE, bool MULTI>
__device__ __forceinline__ void eval_lj(const PairArgs &a, const ISite<R> &I, const Q4<R> &pj,
                                        const Q4<R> &mj, bool ok, R eps24, R sig2, R shift6,
                                        Acc &acc, Part32 &pa) {
    const V3<R> d = {I.pos.x - pj.x, I.pos.y - pj.y, I.pos.z - pj.z};
    V3<R> dm = d;
    if constexpr (!MOLSITE) dm = {I.mol.x - mj.x, I.mol.y - mj.y, I.mol.z - mj.z};
    if constexpr (MULTI) {
        // pj.w carries the partner's column offset (4 * its LJ id) as integer
        // bits in the low word
        if constexpr (sizeof(R) == 8) {
            const double *t = a.pp.dpar + (unsigned)(I.ljrow + __double2loint(pj.w));
            const double2 es = *reinterpret_cast<const double2 *>(t);
            eps24 = es.x;
            sig2 = es.y;
            shift6 = t[2];
        } else {
            const float4 t =
                *reinterpret_cast<const float4 *>(a.pp.fpar + (unsigned)(I.ljrow + __float_as_int(pj.w)));
            eps24 = t.x;
            sig2 = t.y;
            shift6 = t.z;
        }
    }
    const R far = sizeof(R) == 4 ? R(__builtin_inff()) : R(1e300);
    const R r2inv = Math<R>::rcp(ok ? dot3(d, d) : far);
    const R lj2 = sig2 * r2inv, lj6 = lj2 * lj2 * lj2, lj12 = lj6 * lj6;
    const R lj12m6 = lj12 - lj6;
    const R s = eps24 * r2inv * (lj12 + lj12m6);
    const R fx = d.x * s, fy = d.y * s, fz = d.z * s;
    const R u = ok ? eps24 * lj12m6 + shift6 : R(0);
    if constexpr (sizeof(R) == 4) {
        pa.f[0] += fx;
        pa.f[1] += fy;
        pa.f[2] += fz;
        pa.v[0] += dm.x * fx;
        pa.v[1] += dm.y * fy;
        pa.v[2] += dm.z * fz;
        pa.u += u;
    } else {
        acc.f[0] += fx;
        acc.f[1] += fy;
        acc.f[2] += fz;
        acc.v[0] += dm.x * fx;
        acc.v[1] += dm.y * fy;
        acc.v[2] += dm.z * fz;
        acc.u += u;
    }
}

// One electrostatic site pair seen from site i: force on i, virial
// (m_i - m_j) * f_i, torque on i; roles swapped where the reference swaps.
// pj, mj, ej: the partner's site, molecule and orientation records (the
// caller reads them from its pass's tile records; ej only for dipole and
// quadrupole partners).
template <int TI, int TJ, typename R, bool MOLSITE, bool MULTI>
__device__ __forceinline__ void eval_pair(const PairArgs &a, const ISite<R> &I,
                                          const Q4<R> &eq, const Q4<R> &pj, const Q4<R> &mj,
                                          R eps24, R sig2, R shift6, Acc &acc, Part32 &pa) {
    V3<R> d = {I.pos.x - pj.x, I.pos.y - pj.y, I.pos.z - pj.z};
    V3<R> dm = d;
    if constexpr (!MOLSITE) dm = {I.mol.x - mj.x, I.mol.y - mj.y, I.mol.z - mj.z};
    PO<R> o;
    o.rf = R(0);
    R sign = R(1);
    V3<R> tq = {R(0), R(0), R(0)};
    static_assert(TI != T_LJ, "LJ pairs go through eval_lj");
    V3<R> ej = {R(0), R(0), R(0)};
    if constexpr (TJ >= T_D) ej = v3(eq);
    const R qj = pj.w;
    if constexpr (TI == T_C && TJ == T_C) {
        body_cc<R>(d, I.par, qj, o);
    } else if constexpr (TI == T_C && TJ == T_D) {
        body_cd<R>(d, I.par, ej, qj, o);
    } else if constexpr (TI == T_C && TJ == T_Q) {
        body_cq<R>(d, I.par, ej, qj, o);
    } else if constexpr (TI == T_D && TJ == T_C) {
        V3<R> nd = {-d.x, -d.y, -d.z};
        body_cd<R>(nd, qj, I.e, I.par, o);
        sign = R(-1);
        tq = o.M2;
    } else if constexpr (TI == T_D && TJ == T_D) {
        body_dd<R>(d, I.e, I.par, ej, qj, (R)a.epsRFInvrc3, o);
        tq = o.M1;
    } else if constexpr (TI == T_D && TJ == T_Q) {
        body_dq<R>(d, I.e, I.par, ej, qj, o);
        tq = o.M1;
    } else if constexpr (TI == T_Q && TJ == T_C) {
        V3<R> nd = {-d.x, -d.y, -d.z};
        body_cq<R>(nd, qj, I.e, I.par, o);
        sign = R(-1);
        tq = o.M2;
    } else if constexpr (TI == T_Q && TJ == T_D) {
        V3<R> nd = {-d.x, -d.y, -d.z};
        body_dq<R>(nd, ej, qj, I.e, I.par, o);
        sign = R(-1);
        tq = o.M2;
    } else if constexpr (TI == T_Q && TJ == T_Q) {
        body_qq<R>(d, I.e, I.par, ej, qj, o);
        tq = o.M1;
    }
    V3<R> f = {sign * o.f.x, sign * o.f.y, sign * o.f.z};
    if constexpr (sizeof(R) == 4) {
        pa.f[0] += f.x;
        pa.f[1] += f.y;
        pa.f[2] += f.z;
        pa.v[0] += dm.x * f.x;
        pa.v[1] += dm.y * f.y;
        pa.v[2] += dm.z * f.z;
        pa.u += o.u;
        if constexpr (TI == T_D && TJ == T_D) pa.rf += o.rf;
        if constexpr (TI >= T_D) {
            pa.m[0] += tq.x;
            pa.m[1] += tq.y;
            pa.m[2] += tq.z;
        }
    } else {
        acc.f[0] += f.x;
        acc.f[1] += f.y;
        acc.f[2] += f.z;
        acc.v[0] += dm.x * f.x;
        acc.v[1] += dm.y * f.y;
        acc.v[2] += dm.z * f.z;
        acc.u += o.u;
        if constexpr (TI == T_D && TJ == T_D) acc.rf += o.rf;
        if constexpr (TI >= T_D) {
            acc.m[0] += tq.x;
            acc.m[1] += tq.y;
            acc.m[2] += tq.z;
        }
    }
}

constexpr int kUnroll = 4;  // stage-A candidates in flight per lane

// All pairs of the lane's i site with the j sites of type TJ in the 27
// neighbour images: (1) lanes 0..26 fetch the neighbour ranges in one round,
// (2) the neighbourhood is staged into LDS in chunks of CAP sites with
// coalesced, independent loads, (3) stage A tests the molecule-centre cutoff
// from LDS (kUnroll candidates per lane in flight) and keeps per-lane hit
// lists, (4) stage B evaluates the bodies for the hits.
template <int TI, int TJ, typename R, bool MOLSITE, bool MULTI, bool COUNT>
__device__ __forceinline__ void partner_pass(const PairArgs &a, const Geom &g, int cbase,
                                             const ISite<R> &I, bool valid,
                                             int cx, int cy, int cz, int nxs, int ph, int P,
                                             int lane,
                                             Tile<TI, R, MOLSITE, MULTI> &T, Acc &acc, R eps24,
                                             R sig2, R shift6, unsigned long long &ntest) {
    using C = TileCfg<TI, R, MOLSITE, MULTI>;
    const SiteArrays &J = a.s[TJ];
    const double rc2 = (TI == T_LJ) ? a.rclj2 : a.rc2;
    const R rc2r = (R)rc2;
    // FP32 tiles: the test record's w holds |x|^2 and stage A computes
    // r^2 = (|i|^2 + |j|^2) - 2 i.j (4 instead of 6 operations); its rounding
    // stays below 1.5e-5 (cl / rc)^2 rc^2 (coordinates within two cell edges
    // of the i cell's corner), so stage A keeps candidates up to that much
    // above rc^2 and stage B re-decides the band in FP64 as before
    constexpr bool NORM = sizeof(R) == 4 && (C::MOL || (TI == T_LJ && !C::LJID));
    // idle lanes test from far away: r^2 overflows or exceeds any cutoff
    const V3<R> far = {R(1e30), R(1e30), R(1e30)};
    const V3<R> Imol = valid ? I.mol : far, Ipos = valid ? I.pos : far;
    // FP32 re-check band around rc^2 (relative 1e-5, far above FP32 rounding
    // of cell-relative coordinates)
    const R lo = rc2r * R(1 - 1e-5), hi = rc2r * R(1 + 1e-5);
    // stage-A bound: FP64 rc^2; FP32 the top of the re-check band, widened
    // by the rounding bound of the expanded r^2 where that is used
    R cut = sizeof(R) == 4 ? hi : rc2r;
    if constexpr (NORM) {
        // (the bound grows with the square of the frame's extent: nxs - 1
        // cell edges in x when two cells share the tile)
        const double clm = fmax(g.cl[0], fmax(g.cl[1], g.cl[2]));
        const double ext = 0.5 * (nxs - 1);
        const double band = fmax(1e-5, 1.5e-5 * ext * ext * clm * clm / rc2);
        cut = (R)(rc2 * (1.0 + band));
    }
    unsigned long long nhit = 0;
    int nk = 0;
    const int nb = 9 * nxs;  // neighbour ranges: 27, or 36 for two cells
    if (lane < nb) {
        Nbr b = neighbour(g, cx, cy, cz, lane, nxs);
        int j0 = 0;
        if (b.ok) {
            j0 = J.cstart[cbase + b.cell];
            nk = J.cstart[cbase + b.cell + 1] - j0;
        }
        T.j0[lane] = j0;
        T.shift[lane][0] = b.s[0];
        T.shift[lane][1] = b.s[1];
        T.shift[lane][2] = b.s[2];
    }
    int incl = nk;
    for (int o = 1; o < (C::NB > 32 ? 64 : 32); o <<= 1) {
        int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane < nb) T.pre[lane + 1] = incl;
    if (lane == 0) T.pre[0] = 0;
    const int total = __shfl(incl, nb - 1, 64);
    __syncthreads();
    // Charge partners are tested once per molecule (neutral molecules carry
    // two or more charges): the tile keeps a molecule's sites contiguous, the
    // head list names each molecule's first site, stage A runs over heads and
    // stage B over a hit's sites; chunks then end at molecule boundaries. For
    // LJ, dipole and quadrupole partners (typically one or two per molecule)
    // the indirection costs more than it saves (measured) and every site is a
    // candidate.
    constexpr bool heads = C::MOL && TJ == T_C;
    using PT = PassTile<TI, TJ, R, MOLSITE, MULTI>;
    static_assert(PT::HEADS == heads, "head layout");
    const PT V(T);
    // (a molecule's charges are contiguous, and its first one carries the
    // count: one load per site, no chain through the molecule index)
    for (int base = 0; base < total;) {
        int nst = min(PT::CAP, total - base);
        // heads: a chunk that ends inside a molecule ends at that molecule's
        // first site instead (the last head of the chunk); whether it does is
        // the head count of the site after the chunk, loaded beside the
        // chunk's own sites
        int next_head = 1;
        if constexpr (heads) {
            if (base + nst < total) {
                const int idx = base + nst;
                const int k = tile_k(T.pre, idx, nb);
                next_head = J.nhead[T.j0[k] + (idx - T.pre[k])];
            }
        }
        int nh = nst;
        int hb = 0, last_head = 0;
        for (int s0 = 0; s0 < nst; s0 += 64) {
            const int s = s0 + lane;
            bool head = false;
            if (s < nst) {
            const int idx = base + s;
            const int k = tile_k(T.pre, idx, nb);
            const int j = T.j0[k] + (idx - T.pre[k]);
            const V3<R> off = tile_offset<R>(g, T.shift, k, nxs);
            const V3<R> p = SiteView<R>::pos(J, j);
            Q4<R> q;
            q.x = p.x + off.x;
            q.y = p.y + off.y;
            q.z = p.z + off.z;
            q.w = R(0);
            if constexpr (C::LJID) {  // column offset of the interleaved table, as bits
                if constexpr (sizeof(R) == 8)
                    q.w = __hiloint2double(0, 4 * J.ljid[j]);
                else
                    q.w = __int_as_float(4 * J.ljid[j]);
            }
            if constexpr (TI != T_LJ) q.w = SiteView<R>::par(J, j);
            if constexpr (NORM && !C::MOL) q.w = q.x * q.x + q.y * q.y + q.z * q.z;
            V.p[s] = q;
            if constexpr (C::MOL) {
                const V3<R> m = SiteView<R>::mol(J, j);
                if constexpr (heads) {
                    const int w = J.nhead[j];
                    head = w != 0;
                    V.ns[s] = (unsigned char)w;
                }
                Q4<R> mq{m.x + off.x, m.y + off.y, m.z + off.z, R(0)};
                if constexpr (NORM) mq.w = mq.x * mq.x + mq.y * mq.y + mq.z * mq.z;
                V.m[s] = mq;
            }
            if constexpr (TI != T_LJ && TJ >= T_D) {
                const V3<R> e = SiteView<R>::ori(J, j);
                V.e[s] = Q4<R>{e.x, e.y, e.z, R(0)};
            }
            }
            if constexpr (heads) {  // compact the heads in tile order
                const unsigned long long hm = __ballot(head);
                if (head) V.h[hb + __popcll(hm & ((1ull << lane) - 1ull))] = (unsigned short)s;
                hb += __popcll(hm);
                if (hm) last_head = s0 + 63 - __clzll((long long)hm);
            }
        }
        if constexpr (heads) {
            nh = hb;
            // the split molecule's sites run from the last head to the chunk
            // end: drop them (and that head), the next chunk starts there
            // (a chunk holds many molecules of at most kMaxSite sites, so
            // last_head > 0)
            if (next_head == 0 && last_head > 0) {
                nst = last_head;
                nh = hb - 1;
            }
        }
        __syncthreads();
        int cnt = 0;
        // Stage B. A listed candidate is final unless its molecule distance
        // is 0 (the molecule itself, SingleCellPolicy_) or, in FP32, lies in
        // the band [lo, hi] around rc^2, where the decision is re-made in FP64
        // from the unrounded positions so the pair set equals the FP64 one.
        auto recheck = [&](int s) -> bool {  // the FP64 decision for tile entry s
            const int idx = base + s;
            const int k = tile_k(T.pre, idx, nb);
            const int j = T.j0[k] + (idx - T.pre[k]);
            // molecule positions through the sites' molecule indices (the
            // FP32 form builds no FP64 per-site copies)
            const int mi = a.s[TI].mol[I.idx], mj = J.mol[j];
            const double ex = a.mrx[mi] - (a.mrx[mj] + T.shift[k][0]);
            const double ey = a.mry[mi] - (a.mry[mj] + T.shift[k][1]);
            const double ez = a.mrz[mi] - (a.mrz[mj] + T.shift[k][2]);
            const double r2d = ex * ex + ey * ey + ez * ez;
            return r2d < rc2 && r2d != 0.0;
        };
        auto accept = [&](int s, const Q4<R> &q) -> bool {
            const V3<R> mi = C::MOL ? I.mol : I.pos;
            const R dx = mi.x - q.x, dy = mi.y - q.y, dz = mi.z - q.z;
            const R r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == R(0)) return false;
            if constexpr (sizeof(R) == 4) {
                if (r2 >= lo) return recheck(s);
            }
            return true;
        };
        auto drain = [&]() {
            Part32 pa;
            pa.clear();
            if constexpr (TI == T_LJ) {
                // LJ: every lane runs every entry up to the longest list
                // (exhausted lanes on entry 0 with the pair masked off), so
                // the loop has no branch. FP32: entries in the re-check band
                // around rc^2 are masked off here and marked in band (bit e),
                // and re-decided in FP64 after the loop (rare)
                unsigned band = 0u;
                // the longest list, as a wave-uniform (scalar) trip count
                int emax = cnt;
                for (int o = 32; o > 0; o >>= 1) emax = max(emax, __shfl_xor(emax, o, 64));
                emax = __builtin_amdgcn_readfirstlane(emax);
                for (int e = 0; e < emax; ++e) {
                    const bool act = e < cnt;
                    const int raw = (int)T.lst[e][lane];  // e < LIST: in bounds
                    const int sj = act ? raw : 0;
                    const Q4<R> qj = C::MOL ? V.m[sj] : V.p[sj];
                    const V3<R> mi = C::MOL ? I.mol : I.pos;
                    const R dx = mi.x - qj.x, dy = mi.y - qj.y, dz = mi.z - qj.z;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    bool ok = act && r2 != R(0);
                    if constexpr (sizeof(R) == 4) {
                        const bool b = ok && r2 >= lo;
                        band |= b ? (1u << e) : 0u;
                        ok = ok && !b;
                    }
                    const Q4<R> pj = C::MOL ? V.p[sj] : qj;
                    eval_lj<R, MOLSITE, MULTI>(a, I, pj, qj, ok, eps24, sig2, shift6, acc, pa);
                    if (COUNT) nhit += ok;
                }
                if constexpr (sizeof(R) == 4) {
                    while (__any(band != 0u)) {
                        const bool has = band != 0u;
                        const int e = has ? __builtin_ctz(band) : 0;
                        band &= has ? band - 1u : 0u;
                        const int sj = has ? (int)T.lst[e][lane] : 0;
                        const bool ok = has && recheck(sj);
                        const Q4<R> qj = C::MOL ? V.m[sj] : V.p[sj];
                        const Q4<R> pj = C::MOL ? V.p[sj] : qj;
                        eval_lj<R, MOLSITE, MULTI>(a, I, pj, qj, ok, eps24, sig2, shift6, acc, pa);
                        if (COUNT) nhit += ok;
                    }
                }
            } else {
                // the next entry's index and cutoff record are read one
                // iteration ahead, so their LDS latency overlaps the current body
                int sn = 0;
                Q4<R> qn{};
                if (cnt > 0) {
                    sn = (int)T.lst[0][lane];
                    qn = C::MOL ? V.m[sn] : V.p[sn];
                }
                int emax = cnt;  // the longest list: a scalar trip count
                for (int o = 32; o > 0; o >>= 1) emax = max(emax, __shfl_xor(emax, o, 64));
                emax = __builtin_amdgcn_readfirstlane(emax);
                for (int e = 0; e < emax; ++e) {
                    const bool act = e < cnt;
                    const int sj = sn;
                    const Q4<R> qj = qn;
                    if (e + 1 < cnt) {
                        sn = (int)T.lst[e + 1][lane];
                        qn = C::MOL ? V.m[sn] : V.p[sn];
                    }
                    if (act && accept(sj, qj)) {
                        if constexpr (heads) {
                            const int ns = (int)V.ns[sj];  // the molecule's sites sj .. sj + ns - 1
                            for (int t = 0; t < ns; ++t) {
                                const Q4<R> pj = V.p[sj + t];
                                eval_pair<TI, TJ, R, MOLSITE, MULTI>(a, I, Q4<R>{}, pj, qj,
                                                                     eps24, sig2, shift6, acc, pa);
                            }
                            if (COUNT) nhit += ns;
                        } else {
                            const Q4<R> pj = C::MOL ? V.p[sj] : qj;
                            Q4<R> eq{};
                            if constexpr (PT::ORIJ) eq = V.e[sj];
                            eval_pair<TI, TJ, R, MOLSITE, MULTI>(a, I, eq, pj, qj, eps24, sig2,
                                                                 shift6, acc, pa);
                            if (COUNT) ++nhit;
                        }
                    }
                }
            }
            if constexpr (sizeof(R) == 4) pa.flush(acc);
            cnt = 0;
        };
        // Stage A: keep candidates below the cutoff (FP32: up to the top of
        // the re-check band). Row m holds candidates m * P + ph; rows below
        // full are inside [0, nst) for every phase and need no bounds test,
        // and idle lanes carry a far-away position, so the bulk loop is a
        // read, a distance and a compare per candidate.
        // wave-uniform trip counts (scalar loop control; nh and P are
        // wave-uniform, so the division runs on the scalar unit)
        const int nhs = __builtin_amdgcn_readfirstlane(nh), Ps = __builtin_amdgcn_readfirstlane(P);
        const int per = (nhs + Ps - 1) / Ps, full = per - 1;
        const Q4<R> *rec = (C::MOL ? V.m : V.p) + ph;
        // candidate c of the stage-A sequence -> tile index
        auto cand = [&](int c) -> int {
            if constexpr (heads) return (int)V.h[c];
            else return c;
        };
        const V3<R> mi = C::MOL ? Imol : Ipos;
        // NORM: -2 i and |i|^2 per lane (idle lanes: far away, never a hit)
        R mi2 = R(0), nx = R(0), ny = R(0), nz = R(0);
        if constexpr (NORM) {
            mi2 = mi.x * mi.x + mi.y * mi.y + mi.z * mi.z;
            nx = R(-2) * mi.x;
            ny = R(-2) * mi.y;
            nz = R(-2) * mi.z;
        }
        auto dist2 = [&](const Q4<R> &q) -> R {
            if constexpr (NORM) {
                return __builtin_fmaf(nx, q.x, __builtin_fmaf(ny, q.y, __builtin_fmaf(nz, q.z, mi2 + q.w)));
            } else {
                const R dx = mi.x - q.x, dy = mi.y - q.y, dz = mi.z - q.z;
                return dx * dx + dy * dy + dz * dz;
            }
        };
        auto test = [&](R r2) { return sizeof(R) == 4 ? r2 <= cut : r2 < cut; };
        int m0 = 0;
        for (; m0 + kUnroll <= full; m0 += kUnroll) {
            // drain when some lane's list could overflow in this group
            if (__any(cnt > C::LIST - kUnroll)) drain();
            R r2v[kUnroll];
            int sv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                Q4<R> q;
                if constexpr (heads) {
                    sv[u] = cand(ph + (m0 + u) * P);
                    q = V.m[sv[u]];
                } else {
                    sv[u] = ph + (m0 + u) * P;
                    q = rec[(m0 + u) * P];
                }
                r2v[u] = dist2(q);
            }
            if (COUNT && valid) ntest += kUnroll;
            // every candidate is written at the end of the list and kept only
            // on a hit (no branch; the drain test above keeps cnt + kUnroll
            // within the list)
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                T.lst[cnt][lane] = (unsigned short)sv[u];
                cnt += test(r2v[u]) ? 1 : 0;
            }
        }
        for (; m0 < per; ++m0) {  // the last rows, bounds-tested
            if (__any(cnt > C::LIST - 1)) drain();
            const int c = ph + m0 * P;
            const int sidx = cand(min(c, nh - 1));
            const Q4<R> q = C::MOL ? V.m[sidx] : V.p[sidx];
            const R r2 = dist2(q);
            const bool in = valid && c < nh;
            if (COUNT && in) ++ntest;
            T.lst[cnt][lane] = (unsigned short)sidx;
            cnt += (in && test(r2)) ? 1 : 0;
        }
        drain();
        __syncthreads();
        base += nst;
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) nhit += __shfl_xor(nhit, o, 64);
        if (lane == 0) atomicAdd(&a.counters[1 + Kind<TI, TJ>::k], nhit);
    }
}

__device__ __forceinline__ double shfl_xor_d(double v, int o) { return __shfl_xor(v, o, 64); }

// All i sites of one cell (global index) with their partners; adds the
// cell's per-lane macroscopic sums to wmac and writes a zero row for the cell
// (flush_macro later stores the sums of a group of cells in one row).
// span = 2: the cell and the next one in x together, over their common
// neighbourhood (the i sites of the second cell move into the first cell's
// FP32 frame). Returns whether the cells have i sites.
template <int TI, typename R, bool COUNT, bool MOLSITE, bool MULTI>
__device__ __forceinline__ bool pair_cell(const PairArgs &a, Tile<TI, R, MOLSITE, MULTI> &T,
                                          int cell, int span, int lane,
                                          unsigned long long &ntest, double (&wmac)[3]) {
    bool empty;
    {
        const int rep = cell_replica(a.R, cell);
        const Geom &g = a.R.geom[rep];
        const int cbase = a.R.cell_off[rep], local = cell - cbase;
        const SiteArrays &S = a.s[TI];
        const int i_begin = S.cstart[cell], i_end = S.cstart[cell + span];
        const int i_mid = span > 1 ? S.cstart[cell + 1] : i_end;  // the second cell's first site
        const int nxs = span + 2;
        empty = i_begin == i_end;
        const int cx = local % g.n[0], cy = (local / g.n[0]) % g.n[1],
                  cz = local / (g.n[0] * g.n[1]);
        R eps24 = R(0), sig2 = R(0), shift6 = R(0);
        if constexpr (TI == T_LJ) {
            if constexpr (sizeof(R) == 8) {
                eps24 = a.pp.eps24[0]; sig2 = a.pp.sig2[0]; shift6 = a.pp.shift6[0];
            } else {
                eps24 = a.pp.feps24[0]; sig2 = a.pp.fsig2[0]; shift6 = a.pp.fshift6[0];
            }
        }
        for (int chunk = i_begin; chunk < i_end; chunk += 64) {
            // lane = ph * nI + il: P = floor(64 / nI) phases of the chunk's nI
            // sites split the partner candidates; the 64 - P * nI lanes left
            // over idle (phase 0, never a hit). LJ kernels cap P at 8: their
            // sparse cells (vapour) have few candidates, and more phases only
            // lengthen the phase fold; electrostatic kernels, whose cells hold
            // a few sites of many candidates, gain from up to 16 (measured)
            const int nI = min(64, i_end - chunk);
            const int P = min(64 / nI, TI == T_LJ ? 8 : 16);
            // lane / nI through a float reciprocal (exact: lane < 64, nI <= 64,
            // and (lane + 0.5) / nI stays >= 1/128 away from an integer)
            const int ph0 = (int)(((float)lane + 0.5f) * __builtin_amdgcn_rcpf((float)nI));
            const int il = lane - ph0 * nI;
            const bool valid = ph0 < P;
            const int ph = valid ? ph0 : 0;
            const int i = chunk + (valid ? il : 0);
            ISite<R> I;
            I.pos = SiteView<R>::pos(S, i);
            I.mol = MOLSITE ? I.pos : SiteView<R>::mol(S, i);
            if constexpr (sizeof(R) == 4) {
                if (i >= i_mid) {  // second cell: one cell edge further in x
                    const float sx = (float)g.cl[0];
                    I.pos.x += sx;
                    I.mol.x += sx;
                }
            }
            I.idx = i;
            if constexpr (TI >= T_D) I.e = SiteView<R>::ori(S, i);
            if constexpr (TI == T_LJ) {
                I.ljid = S.ljid[i];
                I.ljrow = 4 * a.pp.nlj * I.ljid;
            } else
                I.par = SiteView<R>::par(S, i);
            Acc acc;
            for (int k = 0; k < 3; ++k) acc.f[k] = acc.v[k] = acc.m[k] = 0.0;
            acc.u = acc.vir = acc.rf = 0.0;
            if constexpr (TI == T_LJ) {
                partner_pass<T_LJ, T_LJ, R, MOLSITE, MULTI, COUNT>(a, g, cbase, I, valid, cx, cy, cz, nxs, ph, P,
                                                                   lane, T, acc, eps24, sig2,
                                                                   shift6, ntest);
            } else {
                if (a.present[T_C])
                    partner_pass<TI, T_C, R, MOLSITE, MULTI, COUNT>(a, g, cbase, I, valid, cx, cy, cz, nxs, ph, P,
                                                                    lane, T, acc, eps24, sig2,
                                                                    shift6, ntest);
                if (a.present[T_D])
                    partner_pass<TI, T_D, R, MOLSITE, MULTI, COUNT>(a, g, cbase, I, valid, cx, cy, cz, nxs, ph, P,
                                                                    lane, T, acc, eps24, sig2,
                                                                    shift6, ntest);
                if (a.present[T_Q])
                    partner_pass<TI, T_Q, R, MOLSITE, MULTI, COUNT>(a, g, cbase, I, valid, cx, cy, cz, nxs, ph, P,
                                                                    lane, T, acc, eps24, sig2,
                                                                    shift6, ntest);
            }
            // fold the phases of each site into its phase-0 lane, in phase
            // order (the virial sums below take the per-lane parts before
            // the fold)
            wmac[1] += acc.v[0] + acc.v[1] + acc.v[2];  // virial = trace of the site tensors
            if (P > 1) {
                // through LDS: the tile is free after the partner passes
                constexpr int NA = TI >= T_D ? 9 : 6;
                static_assert(sizeof(T) >= NA * 64 * sizeof(double), "fold scratch");
                double *fb = reinterpret_cast<double *>(&T);
                for (int k = 0; k < 3; ++k) {
                    fb[k * 64 + lane] = acc.f[k];
                    fb[(3 + k) * 64 + lane] = acc.v[k];
                    if constexpr (TI >= T_D) fb[(6 + k) * 64 + lane] = acc.m[k];
                }
                __syncthreads();
                if (valid && ph == 0)
                    for (int q = 1; q < P; ++q) {
                        const int src = lane + q * nI;
                        for (int k = 0; k < 3; ++k) {
                            acc.f[k] += fb[k * 64 + src];
                            acc.v[k] += fb[(3 + k) * 64 + src];
                            if constexpr (TI >= T_D) acc.m[k] += fb[(6 + k) * 64 + src];
                        }
                    }
                __syncthreads();
            }
            if (valid && ph == 0 && a.kvx) {
                // VCP1CLJRMM (VCP1CLJRMM.cpp:300-362): the site's force, times
                // dt/m, goes straight into its molecule's velocity
                a.kvx[i] = __dadd_rn(a.kvx[i], __dmul_rn(acc.f[0], a.kick));
                a.kvy[i] = __dadd_rn(a.kvy[i], __dmul_rn(acc.f[1], a.kick));
                a.kvz[i] = __dadd_rn(a.kvz[i], __dmul_rn(acc.f[2], a.kick));
            } else if (valid && ph == 0) {
                S.Fx[i] = acc.f[0]; S.Fy[i] = acc.f[1]; S.Fz[i] = acc.f[2];
                double vv[3] = {__dmul_rn(a.vscale, acc.v[0]), __dmul_rn(a.vscale, acc.v[1]),
                                __dmul_rn(a.vscale, acc.v[2])};
                if (a.thermo) {  // Homogeneous.cpp:139-145, then calcFM's own sum
                    const double w = a.thermo[kThermoW * cell_replica(a.R, cell) + kThVi];
                    for (int k = 0; k < 3; ++k) vv[k] = __dadd_rn(vv[k], w);
                }
                S.Vx[i] = vv[0];
                S.Vy[i] = vv[1];
                S.Vz[i] = vv[2];
                if constexpr (TI >= T_D) {
                    S.Mx[i] = acc.m[0]; S.My[i] = acc.m[1]; S.Mz[i] = acc.m[2];
                }
            }
            wmac[0] += acc.u;
            wmac[2] += acc.rf;
        }
    }
    // zero rows (the sums of the work item's cells go to one row at the end)
    if (lane == 0 && !a.accumulate)
        for (int k = 0; k < 4 * span; ++k) a.macro_part[4 * (size_t)cell + k] = 0.0;
    return !empty;
}

// The per-lane sums of a group of cells of one replica into the row of the
// group's last non-empty cell.
template <int TI>
__device__ __forceinline__ void flush_macro(const PairArgs &a, double (&wmac)[3], int cell,
                                            int lane) {
    // the reaction-field sum (wmac[2]) is non-zero only for dipole i sites
    for (int k = 0; k < (TI == T_D ? 3 : 2); ++k)
        for (int o = 32; o > 0; o >>= 1) wmac[k] += shfl_xor_d(wmac[k], o);
    if (lane == 0) {
        double *o = a.macro_part + 4 * (size_t)cell;
        const double v[4] = {(TI == T_LJ) ? wmac[0] : 0.0, (TI == T_LJ) ? 0.0 : wmac[0], wmac[1],
                             wmac[2]};
        for (int k = 0; k < 4; ++k) o[k] = a.accumulate ? o[k] + v[k] : v[k];
    }
    wmac[0] = wmac[1] = wmac[2] = 0.0;
}

// One 64-lane block (one wavefront) per work item of cpb consecutive cells.
// XCD-aware: hardware block b runs on XCD b % 8, and XCD x owns the
// contiguous item range [x, x + 1) * items_per_xcd, so neighbouring cells
// share its L2. A grid smaller than 8 * items_per_xcd makes the blocks
// persistent: each walks its XCD's items with stride gridDim.x / 8, and every
// block leaves after its last item.
template <int TI, typename R, bool COUNT, bool MOLSITE, bool MULTI>
__global__ __launch_bounds__(64) __attribute__((
    amdgpu_waves_per_eu(TileCfg<TI, R, MOLSITE, MULTI>::WAVES))) void k_pairs(PairArgs a) {
    __shared__ Tile<TI, R, MOLSITE, MULTI> T;
    const int lane = threadIdx.x;
    const int xcd = blockIdx.x & 7, stride = gridDim.x >> 3;
    unsigned long long ntest = 0;
    for (int item = blockIdx.x >> 3; item < a.items_per_xcd; item += stride) {
        const int c0 = (xcd * a.items_per_xcd + item) * a.cpb;
        const int c1 = min(c0 + a.cpb, a.ncell);
        double wmac[3] = {0.0, 0.0, 0.0};
        int last = -1, rep_prev = -1;
        for (int cell = c0; cell < c1;) {
            const int rep = cell_replica(a.R, cell);
            if (rep != rep_prev && last >= 0) {  // rows never mix replicas
                flush_macro<TI>(a, wmac, last, lane);
                last = -1;
            }
            rep_prev = rep;
            int span = 1;
            if constexpr (TileCfg<TI, R, MOLSITE, MULTI>::JOINT) {
                // pairs of cells in one x row of one replica, where the cell
                // edge exceeds the cutoff: the 9 cells of the common
                // neighbourhood that are not neighbours of one of the two are
                // then at least a cell edge, more than rc, from its sites, so
                // testing them adds no pair that ls1's traversal leaves out
                const Geom &g = a.R.geom[rep];
                const int local = cell - a.R.cell_off[rep];
                const double rc2 = TI == T_LJ ? a.rclj2 : a.rc2;
                if (cell + 1 < c1 && local % g.n[0] + 1 < g.n[0] &&
                    g.cl[0] * g.cl[0] > rc2 * (1.0 + 1e-9))
                    span = 2;
            }
            if (pair_cell<TI, R, COUNT, MOLSITE, MULTI>(a, T, cell, span, lane, ntest, wmac))
                last = cell + span - 1;
            cell += span;
        }
        if (last >= 0) flush_macro<TI>(a, wmac, last, lane);
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) ntest += __shfl_xor(ntest, o, 64);
        if (lane == 0) atomicAdd(&a.counters[0], ntest);
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
template <int T, typename R, bool C>
PairKernel pick(bool molsite, bool multi) {
    if (molsite) return multi ? k_pairs<T, R, C, true, true> : k_pairs<T, R, C, true, false>;
    return multi ? k_pairs<T, R, C, false, true> : k_pairs<T, R, C, false, false>;
}

template <int T> PairKernel pick_t(bool fp32, bool count, bool molsite, bool multi) {
    if (fp32) return count ? pick<T, float, true>(molsite, multi) : pick<T, float, false>(molsite, multi);
    return count ? pick<T, double, true>(molsite, multi) : pick<T, double, false>(molsite, multi);
}

PairKernel pair_kernel(int ti, bool fp32, bool count, bool molsite, bool multi) {
    switch (ti) {
    case T_LJ: return pick_t<T_LJ>(fp32, count, molsite, multi);
    case T_C: return pick_t<T_C>(fp32, count, false, false);
    case T_D: return pick_t<T_D>(fp32, count, false, false);
    case T_Q: return pick_t<T_Q>(fp32, count, false, false);
    }
    return nullptr;
}

}  // namespace mdx
