// mdx_adjoint.hip -- FP64 adjoint of the force path for LJ and charge sites.
//
// For the current cell lists (no stored pair tape: the pair geometry is
// recomputed exactly as the forward pair kernels enumerate it) this returns
//  * dU/dtheta and dW/dtheta for the pair parameters eps_ab, sigma_ab of LJ
//    lookup ids a, b (U = upot6/6 + upotXpoles, W = sum (m_i - m_j) . f),
//    with the truncation shift moving with its parameters where ls1's
//    Comp2Param rules tie it to them (Comp2Param.cpp:27-38,79-82,102-105);
//  * dU/dq and dW/dq for the charge of every (component, charge site);
//  * the force vector-Jacobian product vjp_i = sum_j dF_j/dr_i . c_j with
//    respect to molecule positions (orientations held fixed), for a
//    cotangent c on the molecule forces.
// Layout, as the forward pair kernels: one 64-lane block walks its share of
// an XCD's contiguous range of cells; the cell's i sites occupy nI lanes per
// phase and P = min(64 / nI, 8) phases split the partners; the 27 neighbour
// images are staged into LDS in chunks (site position, and molecule position
// where sites sit off their molecule's centre, with the image shift applied;
// the partner molecule's cotangent), a stage-A pass keeps per-lane hit lists
// of molecule-centre distances below the cutoff and stage B evaluates the
// derivative terms of the hits. Every site accumulates only its own terms
// (owner computes), so each unordered pair is visited from both sides: the
// parameter derivatives are halved per visit and summed over both index
// orders on the host. Parameter sums stay in registers (few LJ ids, charges)
// or per-lane LDS columns until a fixed-order per-block fold, then a
// fixed-order reduction over blocks, so results are deterministic.
#include <cstddef>

#include "mdx_adjoint.h"
#include "mdx_pair_common.h"

namespace mdx {

constexpr int kAdjList = 16;  // per-lane hit list entries
constexpr int kAdjUnroll = 4; // stage-A rows per drain check

// Partner sites per LDS chunk: 104 keeps a one-centre block within 10 KB of
// LDS (9.4 KB of tile and lists, 0.6 KB of tables for up to 2 LJ ids), so 16
// blocks (4 waves per SIMD) fit a CU's 160 KB and the registers are
// allocated for that occupancy, and most cells' 27 neighbour images fit 3
// chunks; the kernel with 4 LJ ids keeps 128, as its sixteen register sums
// exceed the 4-wave register budget anyway and its lane columns take the
// bytes of 128 site records.
template <int TI, int NLR> constexpr int adj_cap() { return TI == T_LJ && NLR == 4 ? 128 : 104; }

template <bool MS, int CAP> struct AdjTile {
    Q4<double> p[CAP];  // site position, i cell frame; w = LJ id or charge
    Q4<double> m[MS ? 1 : CAP];  // molecule position, i cell frame (MS: = p)
    Q4<double> c[CAP];  // cotangent on the partner molecule's force
    unsigned short lst[kAdjList][64];
    int j0[27], pre[28];
    double shift[27][3];
};

// NLR > 0: the LJ parameter sums of a lane (partner LJ ids < NLR) stay in
// registers until the chunk's fold instead of read-modify-writes of its LDS
// column per pair (charge sites always keep their two sums in registers)
// MS: every site sits at its molecule's centre (one-centre LJ), so the tile
// keeps no separate molecule positions
template <int TI, int NLR, bool MS>
__global__ __launch_bounds__(64) void k_adjoint(AdjArgs a) {
    constexpr int CAP = adj_cap<TI, NLR>();
    using Tile = AdjTile<MS, CAP>;
    __shared__ Tile T;
    extern __shared__ double lds[];
    // per-lane columns [W][64], the block table [K*W], lane keys [64], then
    // (LJ) the per-pair constants [nl * nl][kAdjPar]. Kernels that keep the
    // sums in registers fill the columns only for the fold after a chunk's
    // partner passes, when the tile is free: their columns use its bytes.
    constexpr bool ALIAS = TI != T_LJ || NLR > 0;
    static_assert(!ALIAS || offsetof(Tile, lst) >=
                                (TI == T_LJ ? 4 * NLR : 2) * 64 * sizeof(double),
                  "columns in the tile's site records");
    const int W = a.W, K = a.K, TW = K * W;
    double *col = ALIAS ? reinterpret_cast<double *>(&T) : lds;
    double *table = ALIAS ? lds : lds + (size_t)W * 64;
    int *lkey = (int *)(table + TW);
    double *ptab = (double *)(lkey + 64);
    const int lane = threadIdx.x;
    for (int e = lane; e < TW; e += 64) table[e] = 0.0;
    if constexpr (!ALIAS)
        for (int w = 0; w < W; ++w) col[w * 64 + lane] = 0.0;
    const SiteArrays &S = a.s;
    const double rc2 = TI == T_LJ ? a.g.rclj2 : a.g.rc2;
    const int nl = a.nlj;
    const bool vjp = a.cx != nullptr;
    if constexpr (TI == T_LJ) {
        // the parameter-only factors of the derivative terms, once per block:
        // 4 eps / sigma, 24 eps / sigma and the truncation-shift terms of
        // dU/deps and dU/dsigma (zero where no shift is tied to the pair)
        for (int kk = lane; kk < nl * nl; kk += 64) {
            const double e24 = a.pp.eps24[kk], s2 = a.pp.sig2[kk];
            const double eps = e24 / 24.0, sig = sqrt(s2);
            double cue = 0.0, cus = 0.0;
            const double rcs = a.shift_rc[kk];
            if (rcs > 0.0) {
                double s6 = s2 / (rcs * rcs);
                s6 = s6 * s6 * s6;
                cue = 4.0 * (s6 - s6 * s6);
                cus = 4.0 * eps * (6.0 * s6 - 12.0 * s6 * s6) / sig;
            }
            double *pt = ptab + (size_t)kk * kAdjPar;
            pt[0] = e24;
            pt[1] = s2;
            pt[2] = 4.0 * eps / sig;
            pt[3] = e24 / sig;
            pt[4] = cue;
            pt[5] = cus;
        }
        __syncthreads();
    }
    // cells: XCD-aware where the grid allows it (block b runs on XCD b % 8,
    // which walks a contiguous eighth of the cells, so neighbour sites are
    // shared through that XCD's L2), else a plain stride
    int cfirst = blockIdx.x, cend = a.ncell, cstride = gridDim.x;
    if (gridDim.x >= 8 && (gridDim.x & 7) == 0) {
        const int per = (a.ncell + 7) / 8, x = blockIdx.x & 7;
        cfirst = x * per + (blockIdx.x >> 3);
        cend = min(a.ncell, (x + 1) * per);
        cstride = gridDim.x >> 3;
    }
    for (int cell = cfirst; cell < cend; cell += cstride) {
        const int i_begin = S.cstart[cell], i_end = S.cstart[cell + 1];
        if (i_begin == i_end) continue;
        const int cx = cell % a.g.n[0], cy = (cell / a.g.n[0]) % a.g.n[1],
                  cz = cell / (a.g.n[0] * a.g.n[1]);
        // the 27 neighbour ranges and their prefix (same for every chunk)
        int nk = 0;
        if (lane < 27) {
            Nbr b = neighbour(a.g, cx, cy, cz, lane);
            int j0 = 0;
            if (b.ok) {
                j0 = S.cstart[b.cell];
                nk = S.cstart[b.cell + 1] - j0;
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
        for (int chunk = i_begin; chunk < i_end; chunk += 64) {
            const int nI = min(64, i_end - chunk);
            const int P = min(64 / nI, 8);
            const int ph0 = lane / nI;
            const int il = lane - ph0 * nI;
            const bool valid = ph0 < P;
            const int ph = valid ? ph0 : 0;
            const int i = chunk + (valid ? il : 0);
            // idle lanes test from far away and never hit; one-centre sites
            // (MS) sit at their molecule's position bit for bit (the site
            // offset is a rotated zero vector)
            const double mx = valid ? S.mx[i] : 1e300, my = S.my[i], mz = S.mz[i];
            const double px = MS ? mx : S.x[i], py = MS ? my : S.y[i], pz = MS ? mz : S.z[i];
            const int mi = a.site_is_mol ? i : S.mol[i];
            double ci[3] = {0, 0, 0};
            if (vjp) ci[0] = a.cx[mi], ci[1] = a.cy[mi], ci[2] = a.cz[mi];
            int key;
            double qi = 0.0;
            if constexpr (TI == T_LJ) {
                key = S.ljid[i];
            } else {
                key = a.mcid[mi] * kMaxSite + (i - S.mstart[mi]);
                qi = S.p[i];
            }
            double g[3] = {0, 0, 0};
            constexpr int NR = TI == T_LJ ? (NLR > 0 ? 4 * NLR : 1) : 2;
            double racc[NR];
            for (int w = 0; w < NR; ++w) racc[w] = 0.0;
            for (int base = 0; base < total; base += CAP) {
                const int nst = min(CAP, total - base);
                for (int t = lane; t < nst; t += 64) {
                    const int idx = base + t;
                    const int k = tile_k(T.pre, idx);
                    const int j = T.j0[k] + (idx - T.pre[k]);
                    const double sx = T.shift[k][0], sy = T.shift[k][1], sz = T.shift[k][2];
                    const double w = TI == T_LJ ? (double)S.ljid[j] : S.p[j];
                    T.p[t] = Q4<double>{S.x[j] + sx, S.y[j] + sy, S.z[j] + sz, w};
                    if constexpr (!MS)
                        T.m[t] = Q4<double>{S.mx[j] + sx, S.my[j] + sy, S.mz[j] + sz, 0.0};
                    if (vjp) {
                        const int mj = a.site_is_mol ? j : S.mol[j];
                        T.c[t] = Q4<double>{a.cx[mj], a.cy[mj], a.cz[mj], 0.0};
                    }
                }
                __syncthreads();
                int cnt = 0;
                // stage B: the derivative terms of the listed partners
                auto drain = [&]() {
                    int emax = cnt;  // the longest list: a scalar trip count
                    for (int o = 32; o > 0; o >>= 1) emax = max(emax, __shfl_xor(emax, o, 64));
                    emax = __builtin_amdgcn_readfirstlane(emax);
                    for (int e = 0; e < emax; ++e) {
                        if (e >= cnt) continue;
                        const int t = T.lst[e][lane];
                        const Q4<double> pj = T.p[t];
                        const double dx = px - pj.x, dy = py - pj.y, dz = pz - pj.z;
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        const double r2i = Math<double>::rcp(r2);
                        // (m_i - m_j) . d: r^2 itself where sites are their
                        // molecules' centres
                        double dmd = r2;
                        if constexpr (!MS) {
                            const Q4<double> mj4 = T.m[t];
                            dmd = (mx - mj4.x) * dx + (my - mj4.y) * dy + (mz - mj4.z) * dz;
                        }
                        double sc, sp;  // the force-derivative factors (VJP only)
                        if constexpr (TI == T_LJ) {
                            const int bj = (int)pj.w, kk = key * nl + bj;
                            const double *pt = ptab + (size_t)kk * kAdjPar;
                            const double e24 = pt[0], s2 = pt[1];
                            double l6 = s2 * r2i;
                            l6 = l6 * l6 * l6;
                            const double l12 = l6 * l6;
                            sc = e24 * r2i * (2.0 * l12 - l6);
                            sp = e24 * r2i * r2i * (4.0 * l6 - 14.0 * l12);
                            const double gUe = 4.0 * (l12 - l6) + pt[4];
                            const double gUs = pt[2] * (12.0 * l12 - 6.0 * l6) + pt[5];
                            const double gWe = dmd * 24.0 * r2i * (2.0 * l12 - l6);
                            const double gWs = dmd * pt[3] * r2i * (24.0 * l12 - 6.0 * l6);
                            if constexpr (NLR > 0) {
#pragma unroll
                                for (int b = 0; b < NLR; ++b) {
                                    // fma(h, x, s) with h = 0.5 or 0 rounds
                                    // once, as s + 0.5 x does, and leaves the
                                    // other ids' sums unchanged
                                    const double h = bj == b ? 0.5 : 0.0;
                                    racc[4 * b] = __builtin_fma(h, gUe, racc[4 * b]);
                                    racc[4 * b + 1] = __builtin_fma(h, gUs, racc[4 * b + 1]);
                                    racc[4 * b + 2] = __builtin_fma(h, gWe, racc[4 * b + 2]);
                                    racc[4 * b + 3] = __builtin_fma(h, gWs, racc[4 * b + 3]);
                                }
                            } else {
                                double *c0 = col + (size_t)(bj * 4) * 64 + lane;
                                c0[0] += 0.5 * gUe;
                                c0[64] += 0.5 * gUs;
                                c0[128] += 0.5 * gWe;
                                c0[192] += 0.5 * gWs;
                            }
                        } else {
                            const double qj = pj.w;
                            const double ri = Math<double>::rinv(r2, r2i), r3i = r2i * ri;
                            racc[0] += qj * ri;
                            racc[1] += qj * dmd * r3i;
                            sc = qi * qj * r3i;
                            sp = -1.5 * qi * qj * r3i * r2i;
                        }
                        if (vjp) {  // wave-uniform: the parameter-only call skips it
                            const Q4<double> cj = T.c[t];
                            const double dc[3] = {ci[0] - cj.x, ci[1] - cj.y, ci[2] - cj.z};
                            const double ddc = dx * dc[0] + dy * dc[1] + dz * dc[2];
                            g[0] += sc * dc[0] + 2.0 * sp * dx * ddc;
                            g[1] += sc * dc[1] + 2.0 * sp * dy * ddc;
                            g[2] += sc * dc[2] + 2.0 * sp * dz * ddc;
                        }
                    }
                    cnt = 0;
                };
                // stage A: molecule-centre cutoff, strictly less; the molecule
                // itself (distance 0) is skipped (SingleCellPolicy_). Row r
                // holds candidates ph + r P; rows below full lie inside
                // [0, nst) for every phase, so they run without bounds tests,
                // kAdjUnroll rows per drain check; every candidate is written
                // at the end of the list and kept only on a hit
                auto hit = [&](int m) -> bool {
                    const Q4<double> q = MS ? T.p[m] : T.m[m];
                    const double ex = mx - q.x, ey = my - q.y, ez = mz - q.z;
                    const double r2m = ex * ex + ey * ey + ez * ez;
                    return r2m < rc2 && r2m != 0.0;
                };
                const int nsts = __builtin_amdgcn_readfirstlane(nst);
                const int Ps = __builtin_amdgcn_readfirstlane(P);
                const int per = (nsts + Ps - 1) / Ps, full = per - 1;
                int row = 0;
                for (; row + kAdjUnroll <= full; row += kAdjUnroll) {
                    if (__any(cnt > kAdjList - kAdjUnroll)) drain();
                    bool h[kAdjUnroll];
#pragma unroll
                    for (int u = 0; u < kAdjUnroll; ++u) h[u] = hit(ph + (row + u) * P);
#pragma unroll
                    for (int u = 0; u < kAdjUnroll; ++u) {
                        T.lst[cnt][lane] = (unsigned short)(ph + (row + u) * P);
                        cnt += h[u] ? 1 : 0;
                    }
                }
                for (; row < per; ++row) {  // the last rows, bounds-tested
                    if (__any(cnt > kAdjList - 1)) drain();
                    const int m = ph + row * P;
                    const int mc = min(m, nst - 1);
                    const bool h = hit(mc) && m < nst;
                    T.lst[cnt][lane] = (unsigned short)mc;
                    cnt += h ? 1 : 0;
                }
                drain();
                __syncthreads();
            }
            // the phases of each site into its phase-0 lane, in phase order
            // (every shuffle reads the lanes' own partial sums)
            if (vjp) {
                const double own[3] = {g[0], g[1], g[2]};
                const int Ps = __builtin_amdgcn_readfirstlane(P);
                for (int q = 1; q < Ps; ++q) {  // src < P nI <= 64
                    const int src = il + q * nI;
                    g[0] += __shfl(own[0], src, 64);
                    g[1] += __shfl(own[1], src, 64);
                    g[2] += __shfl(own[2], src, 64);
                }
            }
            if (valid && ph == 0 && a.gx) {
                a.gx[i] = g[0];
                a.gy[i] = g[1];
                a.gz[i] = g[2];
            }
            // the register sums into the lane's columns (W <= NR of them)
            if constexpr (TI != T_LJ || NLR > 0) {
#pragma unroll
                for (int w = 0; w < NR; ++w)
                    if (w < W) col[w * 64 + lane] = racc[w];
            }
            lkey[lane] = valid ? key : -1;
            __syncthreads();
            // fixed-order fold of the lane columns into the block table. With
            // TW <= 32 outputs, output e = (key kq, column w) is summed by G
            // lanes, each over 64 / G source lanes (G the largest power of
            // two with G TW <= 64), and the G partial sums are added in lane
            // order; otherwise one lane per output runs over all 64 lanes
            if (TW > 0 && TW <= 32) {
                int G = 2;
                while (2 * G * TW <= 64) G *= 2;
                const int e = lane / G, part = lane - e * G, span = 64 / G;
                double acc = 0.0;
                if (e < TW) {
                    const int kq = e / W, w = e - kq * W;
                    for (int l = part * span; l < (part + 1) * span; ++l)
                        if (lkey[l] == kq) acc += col[w * 64 + l];
                }
                double tot = acc;  // used by the first lane of each group
                for (int q = 1; q < G; ++q) tot += __shfl(acc, min(lane + q, 63), 64);
                if (e < TW && part == 0) table[e] += tot;
            } else {
                for (int e = lane; e < TW; e += 64) {
                    const int kq = e / W, w = e - kq * W;
                    double acc = 0.0;
                    for (int l = 0; l < 64; ++l)
                        if (lkey[l] == kq) acc += col[w * 64 + l];
                    table[e] += acc;
                }
            }
            __syncthreads();
            if constexpr (!ALIAS)
                for (int w = 0; w < W; ++w) col[w * 64 + lane] = 0.0;
        }
    }
    for (int e = lane; e < TW; e += 64) a.part[(size_t)blockIdx.x * TW + e] = table[e];
}

// vjp of molecule i = sum of its LJ and charge site terms (current order).
__global__ void k_adjoint_mol(int n, const int *__restrict__ ms0, const int *__restrict__ ms1,
                              const double *g0x, const double *g0y, const double *g0z,
                              const double *g1x, const double *g1y, const double *g1z,
                              double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3] = {0, 0, 0};
    if (g0x)
        for (int s = ms0[i]; s < ms0[i + 1]; ++s) v[0] += g0x[s], v[1] += g0y[s], v[2] += g0z[s];
    if (g1x)
        for (int s = ms1[i]; s < ms1[i + 1]; ++s) v[0] += g1x[s], v[1] += g1y[s], v[2] += g1z[s];
    out[i] = v[0];
    out[n + i] = v[1];
    out[2 * n + i] = v[2];
}

// out[d * n + i] = in[orig[i] * 3 + d]: upload-order AoS to current-order SoA.
__global__ void k_gather_orig3(int n, const int *__restrict__ orig, const double *__restrict__ in,
                               double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = orig[i];
    out[i] = in[3 * o];
    out[n + i] = in[3 * o + 1];
    out[2 * n + i] = in[3 * o + 2];
}

template <bool MS> AdjKernel adjoint_lj(int nlj) {
    if (nlj <= 2) return k_adjoint<T_LJ, 2, MS>;
    if (nlj <= 4) return k_adjoint<T_LJ, 4, MS>;
    return k_adjoint<T_LJ, 0, MS>;
}

// Whether the kernel adjoint_kernel picks keeps its lane columns in its tile
// (then the launch's dynamic LDS holds no columns).
bool adjoint_columns_in_tile(int ti, int nlj) { return ti != T_LJ || nlj <= 4; }

AdjKernel adjoint_kernel(int ti, int nlj, bool sites_at_centre) {
    if (ti != T_LJ) return k_adjoint<T_C, 0, false>;
    return sites_at_centre ? adjoint_lj<true>(nlj) : adjoint_lj<false>(nlj);
}
AdjMolKernel adjoint_mol_kernel() { return k_adjoint_mol; }
GatherKernel gather_orig3_kernel() { return k_gather_orig3; }

}  // namespace mdx
