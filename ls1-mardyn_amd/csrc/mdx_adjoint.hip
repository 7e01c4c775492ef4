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
// Layout: one 64-lane block walks a strided set of cells; lane = (i-site,
// j-phase) as in the forward kernels; j sites are read straight from the
// coalesced SoA site arrays. Every site accumulates only its own terms
// (owner computes), so each unordered pair is visited from both sides: the
// parameter derivatives are halved per visit and summed over both index
// orders on the host. Parameter sums go through per-lane LDS columns and a
// fixed-order per-block fold, then a fixed-order reduction over blocks, so
// results are deterministic.
#include "mdx_adjoint.h"
#include "mdx_pair_common.h"

namespace mdx {

template <int TI>
__global__ __launch_bounds__(64) void k_adjoint(AdjArgs a) {
    extern __shared__ double lds[];
    // per-lane columns [W][64], the block table [K*W], lane keys [64]
    const int W = a.W, K = a.K, TW = K * W;
    double *col = lds;
    double *table = lds + (size_t)W * 64;
    int *lkey = (int *)(table + TW);
    const int lane = threadIdx.x;
    for (int e = lane; e < TW; e += 64) table[e] = 0.0;
    for (int w = 0; w < W; ++w) col[w * 64 + lane] = 0.0;
    const SiteArrays &S = a.s;
    const double rc2 = TI == T_LJ ? a.g.rclj2 : a.g.rc2;
    const int nl = a.nlj;
    for (int cell = blockIdx.x; cell < a.ncell; cell += gridDim.x) {
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
            const double px = S.x[i], py = S.y[i], pz = S.z[i];
            const double mx = S.mx[i], my = S.my[i], mz = S.mz[i];
            const int mi = S.mol[i];
            double ci[3] = {0, 0, 0};
            if (a.cx) ci[0] = a.cx[mi], ci[1] = a.cy[mi], ci[2] = a.cz[mi];
            int key;
            double qi = 0.0;
            if constexpr (TI == T_LJ) {
                key = S.ljid[i];
            } else {
                key = a.mcid[mi] * kMaxSite + (i - S.mstart[mi]);
                qi = S.p[i];
            }
            double g[3] = {0, 0, 0};
            for (int k = 0; k < 27; ++k) {
                Nbr b = neighbour(a.g, cx, cy, cz, k);
                if (!b.ok) continue;
                const int j0 = S.cstart[b.cell], j1 = S.cstart[b.cell + 1];
                for (int j = j0 + ph; j < j1; j += P) {
                    if (!valid) break;
                    const double dmx = mx - (S.mx[j] + b.s[0]), dmy = my - (S.my[j] + b.s[1]),
                                 dmz = mz - (S.mz[j] + b.s[2]);
                    const double r2m = dmx * dmx + dmy * dmy + dmz * dmz;
                    if (!(r2m < rc2)) continue;
                    if (k == 13 && r2m == 0.0) continue;
                    const double dx = px - (S.x[j] + b.s[0]), dy = py - (S.y[j] + b.s[1]),
                                 dz = pz - (S.z[j] + b.s[2]);
                    const double r2 = dx * dx + dy * dy + dz * dz, r2i = 1.0 / r2;
                    const double dmd = dmx * dx + dmy * dy + dmz * dz;
                    const int mj = S.mol[j];
                    double dc[3] = {0, 0, 0};
                    if (a.cx) dc[0] = ci[0] - a.cx[mj], dc[1] = ci[1] - a.cy[mj], dc[2] = ci[2] - a.cz[mj];
                    const double ddc = dx * dc[0] + dy * dc[1] + dz * dc[2];
                    double sc, sp;
                    if constexpr (TI == T_LJ) {
                        const int bj = S.ljid[j], kk = key * nl + bj;
                        const double e24 = a.pp.eps24[kk], s2 = a.pp.sig2[kk];
                        const double eps = e24 / 24.0, sig = sqrt(s2);
                        double l6 = s2 * r2i;
                        l6 = l6 * l6 * l6;
                        const double l12 = l6 * l6;
                        sc = e24 * r2i * (2.0 * l12 - l6);
                        sp = e24 * r2i * r2i * (4.0 * l6 - 14.0 * l12);
                        double gUe = 4.0 * (l12 - l6), gUs = 4.0 * eps * (12.0 * l12 - 6.0 * l6) / sig;
                        const double rcs = a.shift_rc[kk];
                        if (rcs > 0.0) {
                            double s6 = s2 / (rcs * rcs);
                            s6 = s6 * s6 * s6;
                            gUe += 4.0 * (s6 - s6 * s6);
                            gUs += 4.0 * eps * (6.0 * s6 - 12.0 * s6 * s6) / sig;
                        }
                        const double gWe = dmd * 24.0 * r2i * (2.0 * l12 - l6);
                        const double gWs = dmd * e24 * r2i * (24.0 * l12 - 6.0 * l6) / sig;
                        double *c0 = col + (size_t)(bj * 4) * 64 + lane;
                        c0[0] += 0.5 * gUe;
                        c0[64] += 0.5 * gUs;
                        c0[128] += 0.5 * gWe;
                        c0[192] += 0.5 * gWs;
                    } else {
                        const double qj = S.p[j];
                        const double ri = sqrt(r2i), r3i = r2i * ri;
                        col[lane] += qj * ri;
                        col[64 + lane] += qj * dmd * r3i;
                        sc = qi * qj * r3i;
                        sp = -1.5 * qi * qj * r3i * r2i;
                    }
                    g[0] += sc * dc[0] + 2.0 * sp * dx * ddc;
                    g[1] += sc * dc[1] + 2.0 * sp * dy * ddc;
                    g[2] += sc * dc[2] + 2.0 * sp * dz * ddc;
                }
            }
            for (int o = Ipad; o < 64; o <<= 1)
                for (int d = 0; d < 3; ++d) g[d] += __shfl_xor(g[d], o, 64);
            if (valid && ph == 0 && a.gx) {
                a.gx[i] = g[0];
                a.gy[i] = g[1];
                a.gz[i] = g[2];
            }
            lkey[lane] = valid ? key : -1;
            __syncthreads();
            // fixed-order fold of the lane columns into the block table
            for (int e = lane; e < TW; e += 64) {
                const int kq = e / W, w = e - kq * W;
                double acc = 0.0;
                for (int l = 0; l < 64; ++l)
                    if (lkey[l] == kq) acc += col[w * 64 + l];
                table[e] += acc;
            }
            __syncthreads();
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

AdjKernel adjoint_kernel(int ti) { return ti == T_LJ ? k_adjoint<T_LJ> : k_adjoint<T_C>; }
AdjMolKernel adjoint_mol_kernel() { return k_adjoint_mol; }
GatherKernel gather_orig3_kernel() { return k_gather_orig3; }

}  // namespace mdx
