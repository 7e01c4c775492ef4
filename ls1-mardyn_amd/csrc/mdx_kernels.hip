// mdx_kernels.hip -- gfx950 kernels of the short-range force path.
//
// Step = leapfrog_pre (+ periodic wrap) -> bin_count -> scan -> bin_scatter
//        -> bin_gather (stable counting sort) -> site scans -> build_sites
//        -> pair kernels (one per i-site type) -> calc_fm -> leapfrog_post.
// Systems of one-centre LJ molecules only: bin_gather_point (gather + site
// cache) -> LJ pair kernel (writes molecule F, Vi) -> leapfrog_post.
//
// Pair kernels: one 64-lane wavefront per cell. Lane = (i-site, j-phase):
// the cell's n sites of the i type occupy n lanes per phase and the
// min(floor(64/n), 8) phases stride over the j sites of the 27 neighbour cells
// (periodic images through shift vectors, no halo copies). Stage A tests the
// molecule-centre cutoff of every candidate and appends hits to a per-lane
// list in LDS; stage B evaluates the site-pair bodies for the dense list, so
// lanes do not idle on the 85% of candidates outside the cutoff. Every site
// accumulates only its own force, virial and torque (owner computes), so no
// atomics and no colouring are needed and results are deterministic; each
// unordered pair is evaluated from both sides and the macroscopic sums are
// halved. Phases are folded through LDS at the end of the cell.
#include "mdx_internal.h"

namespace mdx {

// ------------------------------------------------------------------ helpers
// The quaternion helpers and the Leapfrog updates round every product and sum
// on its own (fp contract off, as the oracle is built), so a fused pass over
// two half steps and the separate passes give the same bits.

// Quaternion::rotate (src/molecules/Quaternion.cpp:43-64)
__device__ __forceinline__ void qrot(const double q[4], const double d[3], double o[3]) {
#pragma clang fp contract(off)
    o[0] = o[1] = o[2] = 0.0;
    if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) return;
    double ww = q[0] * q[0], xx = q[1] * q[1], yy = q[2] * q[2], zz = q[3] * q[3];
    double wx = q[0] * q[1], wy = q[0] * q[2], wz = q[0] * q[3];
    double xy = q[1] * q[2], xz = q[1] * q[3], yz = q[2] * q[3];
    o[0] = (ww + xx - yy - zz) * d[0] + 2. * (xy - wz) * d[1] + 2. * (wy + xz) * d[2];
    o[1] = 2. * (wz + xy) * d[0] + (ww - xx + yy - zz) * d[1] + 2. * (yz - wx) * d[2];
    o[2] = 2. * (xz - wy) * d[0] + 2. * (wx + yz) * d[1] + (ww - xx - yy + zz) * d[2];
}

// Quaternion::rotateinv (Quaternion.cpp:66-86)
__device__ __forceinline__ void qrotinv(const double q[4], const double d[3], double o[3]) {
#pragma clang fp contract(off)
    o[0] = o[1] = o[2] = 0.0;
    if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) return;
    double ww = q[0] * q[0], xx = q[1] * q[1], yy = q[2] * q[2], zz = q[3] * q[3];
    double wx = q[0] * q[1], wy = q[0] * q[2], wz = q[0] * q[3];
    double xy = q[1] * q[2], xz = q[1] * q[3], yz = q[2] * q[3];
    o[0] = (ww + xx - yy - zz) * d[0] + 2. * (xy + wz) * d[1] + 2. * (xz - wy) * d[2];
    o[1] = 2. * (xy - wz) * d[0] + (ww - xx + yy - zz) * d[1] + 2. * (yz + wx) * d[2];
    o[2] = 2. * (xz + wy) * d[0] + 2. * (yz - wx) * d[1] + (ww - xx - yy + zz) * d[2];
}

// Quaternion::differentiate (Quaternion.cpp:98-103)
__device__ __forceinline__ void qdiff(const double q[4], const double w[3], double dq[4]) {
#pragma clang fp contract(off)
    dq[0] = .5 * (-q[1] * w[0] - q[2] * w[1] - q[3] * w[2]);
    dq[1] = .5 * (q[0] * w[0] - q[3] * w[1] + q[2] * w[2]);
    dq[2] = .5 * (q[3] * w[0] + q[0] * w[1] - q[1] * w[2]);
    dq[3] = .5 * (-q[2] * w[0] + q[1] * w[1] + q[0] * w[2]);
}

// CellBorderAndFlagManager::getBorder (CellBorderAndFlagManager.h:134-152).
// Contraction off: the cell id must match the reference bit for bit.
__device__ __forceinline__ double cell_border(const Geom &g, int idx, int d) {
#pragma clang fp contract(off)
    if (idx == 0) return g.halo_min[d];
    if (idx == 1) return g.bbmin[d];
    if (idx == g.cpd[d] - 1) return g.bbmax[d];
    if (idx == g.cpd[d]) return g.halo_max[d];
    return idx * g.cl[d] + g.halo_min[d];
}

// LinkedCells::getCellIndexOfPoint (LinkedCells.cpp:870-923): 3-D ls1 index
// (halo layer included).
__device__ __forceinline__ void ls1_cell3(const Geom &g, const double p[3], int ci[3]) {
#pragma clang fp contract(off)
    double lp[3] = {p[0], p[1], p[2]};
    for (int d = 0; d < 3; ++d) {
        if (lp[d] <= g.halo_min[d])
            lp[d] += g.cl[d] * 0.5;
        else if (lp[d] >= g.halo_max[d])
            lp[d] -= g.cl[d] * 0.5;
        int v = (int)floor((lp[d] - g.bbmin[d]) * g.rcl[d]) + 1;
        v = v < 0 ? 0 : v;
        v = v > g.cpd[d] - 1 ? g.cpd[d] - 1 : v;
        ci[d] = v;
    }
    bool inside = true;
    for (int d = 0; d < 3; ++d)
        inside = inside && cell_border(g, ci[d], d) <= lp[d] && lp[d] < cell_border(g, ci[d] + 1, d);
    if (!inside) {
        for (int d = 0; d < 3; ++d) {
            if (lp[d] < cell_border(g, ci[d], d))
                ci[d]--;
            else if (lp[d] >= cell_border(g, ci[d] + 1, d))
                ci[d]++;
        }
    }
}

// Inner cell index from an ls1 3-D index (x fastest, like
// LinkedCells::cellIndexOf3DIndex, LinkedCells.cpp:925-927).
__device__ __forceinline__ int inner_index(const Geom &g, const int ci[3]) {
    int x = min(max(ci[0] - 1, 0), g.n[0] - 1);
    int y = min(max(ci[1] - 1, 0), g.n[1] - 1);
    int z = min(max(ci[2] - 1, 0), g.n[2] - 1);
    return (z * g.n[1] + y) * g.n[0] + x;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// --------------------------------------------------------------- Leapfrog

// Rotational half of FullMolecule::upd_preF (FullMolecule.cpp:340-360):
// L += dt/2 M, q advanced with the half-step angular velocity, normalised.
// Point molecules skip it: their invI and torque are 0, so L is unchanged and
// q only re-normalised (which k_build_sites does every step, as ls1's
// setupSoACache does).
__device__ __forceinline__ void rotate_pre(MolArrays &s, const double *__restrict__ M,
                                           const Component &c, double dt, int n, int i) {
#pragma clang fp contract(off)
    const double dt_halve = .5 * dt;
    double q[4] = {s.qw[i], s.qx[i], s.qy[i], s.qz[i]};
    double L[3] = {s.Lx[i], s.Ly[i], s.Lz[i]};
    double w[3], qh[4], qi[4];
    qrotinv(q, L, w);
    for (int d = 0; d < 3; ++d) w[d] *= c.invI[d];
    qdiff(q, w, qh);
    for (int k = 0; k < 4; ++k) qh[k] = qh[k] * dt_halve + q[k];
    double qc = 1. / sqrt(qh[0] * qh[0] + qh[1] * qh[1] + qh[2] * qh[2] + qh[3] * qh[3]);
    for (int k = 0; k < 4; ++k) qh[k] *= qc;
    for (int d = 0; d < 3; ++d) L[d] += dt_halve * M[d * n + i];
    qrotinv(qh, L, w);
    for (int d = 0; d < 3; ++d) w[d] *= c.invI[d];
    qdiff(qh, w, qi);
    for (int k = 0; k < 4; ++k) q[k] += qi[k] * dt;
    qc = 1. / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] *= qc;
    s.qw[i] = q[0]; s.qx[i] = q[1]; s.qy[i] = q[2]; s.qz[i] = q[3];
    s.Lx[i] = L[0]; s.Ly[i] = L[1]; s.Lz[i] = L[2];
}

// FullMolecule::upd_preF (FullMolecule.cpp:331-361), then the periodic wrap of
// DomainDecompBase::handleDomainLeavingParticles (DomainDecompBase.cpp:199-249),
// for molecule i with its velocity v and force f in registers (f null:
// PositionCellProcessorRMM, r += dt v only). Writes r and v.
__device__ __forceinline__ void lf_pre_mol(MolArrays &s, const double *__restrict__ M,
                                           const Component &c, const Geom &g, double dt, int n,
                                           int i, double v[3], const double *f) {
#pragma clang fp contract(off)
    const double dt_halve = .5 * dt;
    const double dtInv2m = dt_halve / c.m;
    double r[3] = {s.rx[i], s.ry[i], s.rz[i]};
    if (r[0] != r[0]) return;  // deleted at an outflow wall: no longer moved
    if (f && (g.reflect[0] | g.reflect[1] | g.reflect[2] | g.outflow[0] | g.outflow[1] |
              g.outflow[2])) {
        // BoundaryHandler::processGlobalWallLeavingParticles (BoundaryHandler.cpp:
        // 82-134; called before upd_preF, Simulation.cpp:1114), walls in the
        // order +x, -x, +y, -y, +z, -z: a molecule in the wall's slab of width
        // rc (RegionUtils.cpp:11-45, inBox l <= r < u) whose predicted position
        // r + dt (v + dt/2m F) reaches the wall (RegionUtils.cpp:47-56) gets
        // v = -v - 2 dt/2m F at a reflecting wall, so the coming half kick
        // leaves it moving away at -(v + dt/2m F), and is deleted at an
        // outflow wall (later walls no longer see it)
        bool inside = true;
        for (int d = 0; d < 3; ++d) inside = inside && r[d] >= g.bbmin[d] && r[d] < g.bbmax[d];
        bool dead = false;
        for (int d = 0; d < 3 && inside && !dead; ++d) {
            if (!g.reflect[d] && !g.outflow[d]) continue;
            const double adj = dtInv2m * f[d];
            // +wall: slab [max(max - rc, min), max)
            if (r[d] >= fmax(g.bbmax[d] - g.wall_rc, g.bbmin[d]) &&
                r[d] + dt * (v[d] + adj) >= g.bbmax[d]) {
                if (g.reflect[d])
                    v[d] = -v[d] - adj - adj;
                else
                    dead = true;
            }
            // -wall: slab [min, min(min + rc, max))
            if (!dead && r[d] < fmin(g.bbmin[d] + g.wall_rc, g.bbmax[d]) &&
                r[d] + dt * (v[d] + adj) <= g.bbmin[d]) {
                if (g.reflect[d])
                    v[d] = -v[d] - adj - adj;
                else
                    dead = true;
            }
        }
        if (dead) {
            s.rx[i] = s.ry[i] = s.rz[i] = __builtin_nan("");
            return;
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (f) v[d] += dtInv2m * f[d];
        r[d] += dt * v[d];
    }
    if (!c.point) rotate_pre(s, M, c, dt, n, i);
    for (int d = 0; d < 3; ++d) {
        if (!g.pdim[d]) continue;
        double x = r[d];
        if (x < g.bbmin[d]) {
            x = x + g.L[d];
            if (x >= g.bbmax[d]) x = nextafter(g.bbmax[d], -INFINITY);
        } else if (x >= g.bbmax[d]) {
            x = x - g.L[d];
            if (x <= g.bbmin[d]) x = g.bbmin[d];
        }
        r[d] = x;
    }
    s.rx[i] = r[0]; s.ry[i] = r[1]; s.rz[i] = r[2];
    s.vx[i] = v[0]; s.vy[i] = v[1]; s.vz[i] = v[2];
}

__global__ __launch_bounds__(256) void k_leapfrog_pre(MolArrays s, const double *__restrict__ F,
                                                      const double *__restrict__ M,
                                                      const Component *__restrict__ comps, Reps R,
                                                      double dt, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3] = {s.vx[i], s.vy[i], s.vz[i]};
    double f[3] = {0., 0., 0.};
    if (F) f[0] = F[i], f[1] = F[n + i], f[2] = F[2 * n + i];
    lf_pre_mol(s, M, comps[s.cid[i]], R.geom[mol_replica(R, i)], dt, n, i, v, F ? f : nullptr);
}

// FullMolecule::upd_postF (FullMolecule.cpp:363-387) for molecule i: v (in
// registers) and L advanced, {m v^2, I w^2} added to the sums.
__device__ __forceinline__ void lf_post_mol(MolArrays &s, const double *__restrict__ M,
                                            const Component &c, double dt, int n, int i,
                                            double v[3], const double f[3], double &mv2,
                                            double &iw2) {
#pragma clang fp contract(off)
    const double dt_halve = 0.5 * dt;
    const double dtInv2m = dt_halve / c.m;
    double v2 = 0.;
    for (int d = 0; d < 3; ++d) {
        v[d] += dtInv2m * f[d];
        v2 += v[d] * v[d];
    }
    mv2 = c.m * v2;
    if (!c.point) {  // point molecules: L unchanged, I w^2 = 0
        double L[3] = {s.Lx[i], s.Ly[i], s.Lz[i]};
        for (int d = 0; d < 3; ++d) L[d] += dt_halve * M[d * n + i];
        double q[4] = {s.qw[i], s.qx[i], s.qy[i], s.qz[i]};
        double w[3];
        qrotinv(q, L, w);
        for (int d = 0; d < 3; ++d) {
            w[d] *= c.invI[d];
            iw2 += c.I[d] * w[d] * w[d];
        }
        s.Lx[i] = L[0]; s.Ly[i] = L[1]; s.Lz[i] = L[2];
    }
}

// Block partial sums {sum m v^2, sum I w^2, N, rotDOF} (Leapfrog.cpp:63-147).
__device__ __forceinline__ void lf_block_sums(double mv2, double iw2, double cnt, double rot,
                                              double *partial) {
    __shared__ double red[4][4];
    mv2 = wave_sum(mv2);
    iw2 = wave_sum(iw2);
    cnt = wave_sum(cnt);
    rot = wave_sum(rot);
    int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) {
        red[wv][0] = mv2; red[wv][1] = iw2; red[wv][2] = cnt; red[wv][3] = rot;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) a += red[k][threadIdx.x];
        partial[blockIdx.x * 4 + threadIdx.x] = a;
    }
}

// The second half kick of every molecule with the step's sums. Block b
// covers molecules [chunk[b], chunk[b] + 256) clipped to chunk_end[b]: chunks
// never straddle two replicas.
__global__ __launch_bounds__(256) void k_leapfrog_post(MolArrays s, const double *__restrict__ F,
                                                       const double *__restrict__ M,
                                                       const Component *__restrict__ comps,
                                                       double dt, int n,
                                                       const int *__restrict__ chunk,
                                                       const int *__restrict__ chunk_end,
                                                       double *partial, int rmm, int open) {
    int i = chunk[blockIdx.x] + threadIdx.x;
    double mv2 = 0, iw2 = 0, cnt = 0, rot = 0;
    const bool deleted = open && i < chunk_end[blockIdx.x] && s.rx[i] != s.rx[i];
    if (deleted) {
        // deleted at an outflow wall: not moved, not counted
    } else if (rmm && i < chunk_end[blockIdx.x]) {
        // VelocityCellProcessorRMM (VelocityCellProcessorRMM.h:60-110): the
        // kick happened in the force kernel; sum v^2 (no mass factor, as the
        // reference's reduced-memory mode sums it) and N, no rotation
        const double v[3] = {s.vx[i], s.vy[i], s.vz[i]};
        mv2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        cnt = 1.0;
    } else if (i < chunk_end[blockIdx.x]) {
        const Component &c = comps[s.cid[i]];
        double v[3] = {s.vx[i], s.vy[i], s.vz[i]};
        const double f[3] = {F[i], F[n + i], F[2 * n + i]};
        lf_post_mol(s, M, c, dt, n, i, v, f, mv2, iw2);
        cnt = 1.0;
        rot = c.rotdof;
        s.vx[i] = v[0]; s.vy[i] = v[1]; s.vz[i] = v[2];
    }
    lf_block_sums(mv2, iw2, cnt, rot, partial);
}

// Step n's second half kick (with its sums) fused with step n + 1's first
// half kick and drift, for consecutive steps without a thermostat between
// them: the same operations in the same order as k_leapfrog_post followed by
// k_leapfrog_pre, with v and F read once.
__global__ __launch_bounds__(256) void k_leapfrog_post_pre(MolArrays s,
                                                           const double *__restrict__ F,
                                                           const double *__restrict__ M,
                                                           const Component *__restrict__ comps,
                                                           Reps R, double dt, int n,
                                                           const int *__restrict__ chunk,
                                                           const int *__restrict__ chunk_end,
                                                           double *partial) {
    int i = chunk[blockIdx.x] + threadIdx.x;
    double mv2 = 0, iw2 = 0, cnt = 0, rot = 0;
    if (i < chunk_end[blockIdx.x]) {
        const Component &c = comps[s.cid[i]];
        double v[3] = {s.vx[i], s.vy[i], s.vz[i]};
        const double f[3] = {F[i], F[n + i], F[2 * n + i]};
        lf_post_mol(s, M, c, dt, n, i, v, f, mv2, iw2);
        cnt = 1.0;
        rot = c.rotdof;
        lf_pre_mol(s, M, c, R.geom[mol_replica(R, i)], dt, n, i, v, f);
    }
    lf_block_sums(mv2, iw2, cnt, rot, partial);
}

// Fixed-order reduction of per-block partials (width values per block):
// block b of the launch sums the contiguous range [b*per, (b+1)*per) of the
// input blocks into out[b*width + w]. One block with per = nblocks gives the
// final sum; a many-block pass followed by a one-block pass is the two-level
// form. Deterministic.
__global__ __launch_bounds__(256) void k_reduce_partials(const double *__restrict__ partial,
                                                         int nblocks, int width, int per,
                                                         double *out) {
    __shared__ double red[256];
    const int b0 = blockIdx.x * per, b1 = min(nblocks, b0 + per);
    for (int w = 0; w < width; ++w) {
        double a = 0;
        for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) a += partial[b * width + w];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[blockIdx.x * width + w] = red[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------- binning

__global__ void k_fill_int(int *a, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_double(double *a, int n, double v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Cell of every molecule (ls1 index -> inner index) and its arrival rank in
// the cell. Molecules are stored in cell order, so a wave's lanes form a few
// runs of equal cells: the first lane of each run takes one atomic for the
// whole run, and all runs' atomics go out in the same instruction (one
// round trip per wave; a cell that shows up in two runs gets two disjoint
// rank ranges).
__global__ __launch_bounds__(256) void k_bin_count(MolArrays s, Reps R, int n, int *cell,
                                                   int *rank, int *count, int *ls1id) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool in = i < n;
    int c = -1;
    if (in) {
        double p[3] = {s.rx[i], s.ry[i], s.rz[i]};
        if (p[0] != p[0] && R.nrep == 1) {
            // deleted at an outflow wall: the bin behind the last cell, which
            // no traversal visits
            c = R.cell_off[1];
            ls1id[i] = -1;
        } else {
            const int rep = mol_replica(R, i);
            const Geom &g = R.geom[rep];
            int ci[3];
            ls1_cell3(g, p, ci);
            c = R.cell_off[rep] + inner_index(g, ci);
            ls1id[i] = (ci[2] * g.cpd[1] + ci[1]) * g.cpd[0] + ci[0];
        }
        cell[i] = c;
    }
    const int cprev = __shfl_up(c, 1, 64);
    const bool head = in && (lane == 0 || cprev != c);
    const unsigned long long heads = __ballot(head);
    const unsigned long long valid = __ballot(in);
    // my run's head: the highest head at or below my lane
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all for 63)
    const unsigned long long below = heads & upto;
    const int h = in ? 63 - __clzll((long long)below) : 0;
    int base = 0;
    if (head) {
        // the run ends at the next head or at the last valid lane
        const unsigned long long after = heads & ~upto;
        const int end = after ? __ffsll((long long)after) - 1 : 64 - __clzll((long long)valid);
        base = atomicAdd(&count[c], end - lane);
    }
    base = __shfl(base, h, 64);
    if (in) rank[i] = base + (lane - h);
}

__global__ __launch_bounds__(256) void k_bin_scatter(int n, const int *__restrict__ cell,
                                                     const int *__restrict__ rank,
                                                     const int *__restrict__ start, int *perm) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[cell[i]] + rank[i]] = i;
}

// Destination of entry p of perm (cells in order, each cell's members in
// arrival order): its cell's start plus its stable rank in the cell, the
// number of members with a smaller previous index (which makes the sort
// deterministic and keeps molecules that stay in their cell in order). The
// block's cells occupy one contiguous range of perm, staged in LDS when it
// fits (the rank loops then read LDS); every thread of the block must call
// this. Returns -1 for p >= n.
constexpr int kGatherWin = 1024;
__device__ __forceinline__ int stable_dest(int p, int n, const int *__restrict__ cell,
                                           const int *__restrict__ perm,
                                           const int *__restrict__ start, int *win, int *src_i,
                                           int *cell_of) {
    const int p0 = blockIdx.x * blockDim.x, p1 = min(p0 + (int)blockDim.x, n);
    const int w0 = start[cell[perm[p0]]], w1 = start[cell[perm[p1 - 1]] + 1];
    const bool staged = w1 - w0 <= kGatherWin;
    if (staged)
        for (int k = w0 + (int)threadIdx.x; k < w1; k += blockDim.x) win[k - w0] = perm[k];
    __syncthreads();
    if (p >= n) return -1;
    const int i = perm[p];
    const int c = cell[i];
    const int b = start[c], e = start[c + 1];
    int r = 0;
    if (staged)
        for (int k = b; k < e; ++k) r += win[k - w0] < i;
    else
        for (int k = b; k < e; ++k) r += perm[k] < i;
    *src_i = i;
    *cell_of = c;
    return b + r;
}

// The permutation of the state into cell order (i -> d); q and L are left to
// the caller.
__device__ __forceinline__ void gather_common(int i, int d, int c, const MolArrays &src,
                                              const MolArrays &dst, int *newcell,
                                              const int *__restrict__ ls1id, int *newls1id) {
    dst.rx[d] = src.rx[i]; dst.ry[d] = src.ry[i]; dst.rz[d] = src.rz[i];
    dst.vx[d] = src.vx[i]; dst.vy[d] = src.vy[i]; dst.vz[d] = src.vz[i];
    dst.cid[d] = src.cid[i];
    dst.orig[d] = src.orig[i];
    newcell[d] = c;
    newls1id[d] = ls1id[i];
}

__global__ __launch_bounds__(256) void k_bin_gather(int n, MolArrays src, MolArrays dst,
                                                    const int *__restrict__ cell,
                                                    const int *__restrict__ perm,
                                                    const int *__restrict__ start, int *newcell,
                                                    const int *__restrict__ ls1id, int *newls1id,
                                                    const Component *__restrict__ comps,
                                                    int *nsite0, int *nsite1, int *nsite2,
                                                    int *nsite3) {
    __shared__ int win[kGatherWin];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int i, c;
    const int d = stable_dest(p, n, cell, perm, start, win, &i, &c);
    if (d < 0) return;
    gather_common(i, d, c, src, dst, newcell, ls1id, newls1id);
    dst.qw[d] = src.qw[i]; dst.qx[d] = src.qx[i]; dst.qy[d] = src.qy[i]; dst.qz[d] = src.qz[i];
    dst.Lx[d] = src.Lx[i]; dst.Ly[d] = src.Ly[i]; dst.Lz[d] = src.Lz[i];
    const Component &cp = comps[src.cid[i]];
    nsite0[d] = cp.n[0];
    nsite1[d] = cp.n[1];
    nsite2[d] = cp.n[2];
    nsite3[d] = cp.n[3];
}

// The same for systems whose every component is one LJ centre at the
// molecule's centre and nothing else: site j is molecule j, so the gather
// also does the SoA cache build of k_build_sites for them: the FP32 position
// relative to the cell corner and the LJ lookup id. Site FP64 positions,
// forces and virials are the molecule arrays themselves. q and L of such
// molecules never change and stay in upload order (q's normalisation of
// setupSoACache, FullMolecule.cpp:719, is applied by k_normalize_q).
__global__ __launch_bounds__(256) void k_bin_gather_point(
    int n, MolArrays src, MolArrays dst, const int *__restrict__ cell,
    const int *__restrict__ perm, const int *__restrict__ start, int *newcell,
    const int *__restrict__ ls1id, int *newls1id, const Component *__restrict__ comps, Reps R,
    float *fx, float *fy, float *fz, int *ljid) {
    __shared__ int win[kGatherWin];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int i, c;
    const int d = stable_dest(p, n, cell, perm, start, win, &i, &c);
    if (d < 0) return;
    gather_common(i, d, c, src, dst, newcell, ls1id, newls1id);
    const int rep = c < R.cell_off[R.nrep] ? cell_replica(R, c) : 0;  // deleted: the last bin
    const Geom &g = R.geom[rep];
    const int local = c - R.cell_off[rep];
    const int cx = local % g.n[0], cy = (local / g.n[0]) % g.n[1], cz = local / (g.n[0] * g.n[1]);
    fx[d] = (float)(src.rx[i] - (g.bbmin[0] + cx * g.cl[0]));
    fy[d] = (float)(src.ry[i] - (g.bbmin[1] + cy * g.cl[1]));
    fz[d] = (float)(src.rz[i] - (g.bbmin[2] + cz * g.cl[2]));
    ljid[d] = comps[src.cid[i]].lj_base;
}

// q <- N^k(q), N = Quaternion::normalize (q scaled by 1/sqrt(|q|^2)), for
// molecules whose q is touched by nothing else (one-centre systems, where ls1
// normalises every molecule at each of the k container updates). N reaches a
// fixed point or a two-cycle after a few applications (|q|^2 is then 1 to
// within an ulp or two); the rest of the k are skipped, or reduced to their
// parity for a two-cycle. The loop is capped at kMaxNormalize applications,
// far beyond where the sequence settles.
constexpr long long kMaxNormalize = 64;
__global__ __launch_bounds__(256) void k_normalize_q(int n, MolArrays s, long long k) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double q[4] = {s.qw[i], s.qx[i], s.qy[i], s.qz[i]}, qp[4] = {0, 0, 0, 0};
    for (long long t = 0; t < k && t < kMaxNormalize; ++t) {
        const double qs = 1. / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        double qn[4];
        for (int c = 0; c < 4; ++c) qn[c] = q[c] * qs;
        bool same = true, cyc = t > 0;
        for (int c = 0; c < 4; ++c) {
            same = same && qn[c] == q[c];
            cyc = cyc && qn[c] == qp[c];
        }
        if (same) break;
        if (cyc) {  // q <-> qn: the remaining k - t applications alternate
            if ((k - t) & 1)
                for (int c = 0; c < 4; ++c) q[c] = qn[c];
            break;
        }
        for (int c = 0; c < 4; ++c) {
            qp[c] = q[c];
            q[c] = qn[c];
        }
    }
    s.qw[i] = q[0]; s.qx[i] = q[1]; s.qy[i] = q[2]; s.qz[i] = q[3];
}

// ------------------------------------------------------------- scan (int)

// Exclusive scan of blocks of 1024 ints (256 threads x 4); block totals out.
__global__ __launch_bounds__(256) void k_scan_block(ScanArgs a) {
    __shared__ int ws[4];
    const int *__restrict__ in = a.in[blockIdx.y];
    int *out = a.out[blockIdx.y];
    const int n = a.n;
    int base = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        t += v[k];
    }
    // inclusive scan of t across the block
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = t;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[wv] = x;
    __syncthreads();
    int woff = 0;
    for (int k = 0; k < wv; ++k) woff += ws[k];
    int excl = woff + x - t;
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = excl;
        excl += v[k];
    }
    if (threadIdx.x == 255) a.aux[blockIdx.y][blockIdx.x] = woff + x;
}

__global__ __launch_bounds__(256) void k_scan_add(ScanArgs a) {
    int *out = a.out[blockIdx.y];
    const int n = a.n;
    int i = blockIdx.x * 1024 + threadIdx.x;
    int off = a.aux[blockIdx.y][blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        int j = i + k * 256;
        if (j < n) out[j] += off;
    }
}

// out[n] = the total of array blockIdx.y (out holds n + 1 ints).
__global__ void k_scan_total(ScanArgs a) {
    const int *in = a.in[blockIdx.y];
    int *out = a.out[blockIdx.y];
    const int n = a.n;
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = (n > 0) ? out[n - 1] + in[n - 1] : 0;
}

// For array t = blockIdx.y: cstart_t[c] = mstart_t[start[c]] for c in
// [0, ncell] (start[ncell] = n), with mstart_t = a.in[t], cstart_t = a.out[t].
__global__ void k_cell_site_start(int ncell, const int *__restrict__ start, ScanArgs a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    a.out[blockIdx.y][c] = a.in[blockIdx.y][start[c]];
}

// --------------------------------------------------- SoA site cache build

// The block's copy of the component table (ncomp records) in LDS: the site
// loops read offsets, axes and parameters per site, and a component's record
// is then one LDS read away instead of a global load after the cid load.
// Every thread of the block must call it (it synchronises the block).
__device__ __forceinline__ const Component *stage_comps(const Component *__restrict__ comps,
                                                        int ncomp) {
    extern __shared__ double comp_lds[];
    static_assert(sizeof(Component) % sizeof(double) == 0, "component record size");
    const double *src = reinterpret_cast<const double *>(comps);
    const int nw = ncomp * (int)(sizeof(Component) / sizeof(double));
    for (int w = threadIdx.x; w < nw; w += blockDim.x) comp_lds[w] = src[w];
    __syncthreads();
    return reinterpret_cast<const Component *>(comp_lds);
}

// FullMolecule::setupSoACache (FullMolecule.cpp:711-767): normalise q (the
// reference mutates the state here, Quaternion.h:40-42), rotate the body-frame
// sites, record molecule position and parameters; positions also relative to
// the molecule's cell corner in FP32.
// forms: bit 0 the FP64 site records (positions, molecule positions,
// orientations, parameters) the FP64 pair kernels and the adjoint read, bit 1
// the FP32 ones of the FP32-pair form. first: the re-bin's build, which also
// normalises q and writes what both forms read (molecule index of each site,
// LJ ids, charge head counts); a later call with first = 0 adds a form the
// re-bin did not build, from the same (normalised) q.
__global__ __launch_bounds__(256) void k_build_sites(int n, MolArrays s, Reps R,
                                                     const int *__restrict__ mcell,
                                                     const Component *__restrict__ gcomps,
                                                     int ncomp, int forms, int first,
                                                     SiteArrays s0, SiteArrays s1,
                                                     SiteArrays s2, SiteArrays s3) {
    const Component *comps = stage_comps(gcomps, ncomp);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double q[4] = {s.qw[i], s.qx[i], s.qy[i], s.qz[i]};
    if (first) {
        double qs = 1. / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int k = 0; k < 4; ++k) q[k] *= qs;
        s.qw[i] = q[0]; s.qx[i] = q[1]; s.qy[i] = q[2]; s.qz[i] = q[3];
    }
    const bool f64 = forms & 1, f32 = forms & 2;
    const Component &c = comps[s.cid[i]];
    double r[3] = {s.rx[i], s.ry[i], s.rz[i]};
    const int rep = mol_replica(R, i);
    const Geom &g = R.geom[rep];
    int cell = mcell[i] - R.cell_off[rep];
    int cx = cell % g.n[0], cy = (cell / g.n[0]) % g.n[1], cz = cell / (g.n[0] * g.n[1]);
    double corner[3] = {g.bbmin[0] + cx * g.cl[0], g.bbmin[1] + cy * g.cl[1],
                        g.bbmin[2] + cz * g.cl[2]};
    float mrel[3] = {(float)(r[0] - corner[0]), (float)(r[1] - corner[1]),
                     (float)(r[2] - corner[2])};
    SiteArrays *S[4] = {&s0, &s1, &s2, &s3};
#pragma unroll
    for (int t = 0; t < kTypes; ++t) {
        SiteArrays &A = *S[t];
        int base = A.mstart[i];
        for (int k = 0; k < c.n[t]; ++k) {
            int j = base + k;
            double d[3], pos[3];
            qrot(q, c.body[t][k], d);
            for (int a = 0; a < 3; ++a) pos[a] = d[a] + r[a];
            if (first) {
                A.mol[j] = i;
                if (t == T_LJ) A.ljid[j] = c.lj_base + k;
                if (t == T_C) A.nhead[j] = k == 0 ? c.n[T_C] : 0;
            }
            if (f64) {
                A.x[j] = pos[0]; A.y[j] = pos[1]; A.z[j] = pos[2];
                A.mx[j] = r[0]; A.my[j] = r[1]; A.mz[j] = r[2];
                if (t != T_LJ) A.p[j] = c.param[t][k];
            }
            if (f32) {
                A.fx[j] = (float)(pos[0] - corner[0]);
                A.fy[j] = (float)(pos[1] - corner[1]);
                A.fz[j] = (float)(pos[2] - corner[2]);
                A.fmx[j] = mrel[0]; A.fmy[j] = mrel[1]; A.fmz[j] = mrel[2];
                if (t != T_LJ) A.fp[j] = (float)c.param[t][k];
            }
            if (t >= T_D) {
                double e[3];
                qrot(q, c.e[t][k], e);
                if (f64) { A.ex[j] = e[0]; A.ey[j] = e[1]; A.ez[j] = e[2]; }
                if (f32) { A.fex[j] = (float)e[0]; A.fey[j] = (float)e[1]; A.fez[j] = (float)e[2]; }
            }
        }
    }
}

// FullMolecule::calcFM (FullMolecule.cpp:548-631): F = sum F_site,
// M = sum d_site x F_site + sum M_site(D, Q), Vi = 0.5 sum V_site.
__global__ __launch_bounds__(256) void k_calc_fm(int n, MolArrays s,
                                                 const Component *__restrict__ gcomps,
                                                 int ncomp, SiteArrays s0, SiteArrays s1,
                                                 SiteArrays s2, SiteArrays s3, double *F,
                                                 double *M, double *Vi, Reps R,
                                                 const double *__restrict__ thermo) {
    const Component *comps = stage_comps(gcomps, ncomp);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Component &c = comps[s.cid[i]];
    // point molecules: every site offset is 0, so no d x F term needs q
    double q[4] = {1, 0, 0, 0};
    if (!c.point) q[0] = s.qw[i], q[1] = s.qx[i], q[2] = s.qy[i], q[3] = s.qz[i];
    double f[3] = {0, 0, 0}, m[3] = {0, 0, 0}, tm[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
    SiteArrays *S[4] = {&s0, &s1, &s2, &s3};
#pragma unroll
    for (int t = 0; t < kTypes; ++t) {
        SiteArrays &A = *S[t];
        int base = A.mstart[i];
        for (int k = 0; k < c.n[t]; ++k) {
            int j = base + k;
            double d[3];
            qrot(q, c.body[t][k], d);
            double fs[3] = {A.Fx[j], A.Fy[j], A.Fz[j]};
            f[0] += fs[0]; f[1] += fs[1]; f[2] += fs[2];
            m[0] += d[1] * fs[2] - d[2] * fs[1];
            m[1] += d[2] * fs[0] - d[0] * fs[2];
            m[2] += d[0] * fs[1] - d[1] * fs[0];
            tv[0] += A.Vx[j]; tv[1] += A.Vy[j]; tv[2] += A.Vz[j];
            if (t >= T_D) {
                tm[0] += A.Mx[j]; tm[1] += A.My[j]; tm[2] += A.Mz[j];
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        F[d * n + i] = f[d];
        M[d * n + i] = m[d] + tm[d];
        // Homogeneous::calculateLongRange adds VirialCorr / (3N) to every
        // molecule before calcFM adds its sites (Homogeneous.cpp:139-145)
        Vi[d * n + i] =
            thermo ? __dadd_rn(__dmul_rn(0.5, tv[d]), thermo[kThermoW * mol_replica(R, i) + kThVi])
                   : 0.5 * tv[d];
    }
}

// Macroscopic sums of one traversal (VectorizedCellProcessor::endTraversal,
// VectorizedCellProcessor.cpp:122-155). raw = pair-kernel sums over both
// sides of every pair {u6, uXpoles, virial, rf}; halved here. myRF carries
// the reference's negative sign (VectorizedCellProcessor.cpp:2727).
// Step record: Upot, virial, summv2, sumIw2, Upot_LJ, Upot_rest.
// One thread per replica; raw, macro are [nrep][4], the step record is
// [step][nrep][6].
__global__ void k_finalize_macro(const double *raw, double *macro, double *step_out,
                                 const int *step_idx, int nrep,
                                 const double *__restrict__ thermo) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    raw += 4 * r;
    macro += 4 * r;
    double u6 = 0.5 * raw[0], ux = 0.5 * raw[1], vir = 0.5 * raw[2], rf = -0.5 * raw[3];
    macro[0] = u6;
    macro[1] = ux;
    macro[2] = vir;
    macro[3] = rf;
    double *rec = step_out + 6 * ((size_t)step_idx[0] * nrep + r);
    rec[0] = u6 / 6.0 + ux + rf;
    rec[1] = vir + 3.0 * rf;
    if (thermo) {  // Domain::calculateGlobalValues (Domain.cpp:187-188)
        rec[0] = rec[0] + thermo[kThermoW * r + kThUcorr];
        rec[1] = rec[1] + thermo[kThermoW * r + kThWcorr];
    }
    rec[4] = u6 / 6.0;
    rec[5] = ux + rf;
}

__global__ void k_finalize_lf(const double *sums, double *step_out, const int *step_idx,
                              int nrep, double *thermo, int heuristics) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    double *rec = step_out + 6 * ((size_t)step_idx[0] * nrep + r);
    rec[2] = sums[4 * r];
    rec[3] = sums[4 * r + 1];
    if (!thermo) return;
    // Domain::calculateGlobalValues, global thermostat, Tfactor 1
    // (Domain.cpp:219-313): temperature and scaling factors; the explosion
    // heuristic replaces them when one is below MIN_BETA = 0.9 (its error
    // counter never rises above 0 in the reference, so it stays armed)
    double *th = thermo + kThermoW * r;
    const double Ti = th[kThTarget];
    const double summv2 = sums[4 * r], N = sums[4 * r + 2], rot = sums[4 * r + 3];
    const double sumIw2 = rot > 0 ? sums[4 * r + 1] : 0.0;
    th[kThT] = N > 0 ? (summv2 + sumIw2) / (3 * N + rot) : Ti;
    double bt = 1.0, br = 1.0;
    if (Ti > 0.0 && N > 0) {
        bt = pow(3.0 * N * Ti / summv2, 0.4);
        br = sumIw2 == 0.0 ? 1.0 : pow(rot * Ti / sumIw2, 0.4);
    }
    double explode = 0.0;
    if ((bt < 0.9 || br < 0.9) && heuristics) {
        explode = 1.0;
        bt = 1.0;
        br = pow(br, 0.0091);
    }
    th[kThBetaT] = bt;
    th[kThBetaR] = br;
    th[kThExplode] = explode;
}

// The explosion capping (Domain.cpp:263-297: a molecule's translational
// kinetic energy above 10 Ti, or rotational above 3 rotDOF Ti, is scaled down,
// v with F and D with M) when the step's heuristic fired, then
// VelocityScalingThermostat::apply for the global thermostat without a
// directed velocity (VelocityScalingThermostat.cpp:70-90): v *= beta_trans,
// D *= beta_rot, for every molecule of a replica with a target temperature.
__global__ __launch_bounds__(256) void k_velocity_scale(MolArrays s, double *F, double *M,
                                                        const Component *__restrict__ comps,
                                                        Reps R, const double *__restrict__ thermo,
                                                        int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *th = thermo + kThermoW * mol_replica(R, i);
    const double Ti = th[kThTarget];
    if (!(Ti > 0.0)) return;
    const Component &c = comps[s.cid[i]];
    const double bt = th[kThBetaT], br = th[kThBetaR];
    const bool explode = th[kThExplode] != 0.0;
    double v[3] = {s.vx[i], s.vy[i], s.vz[i]};
    // L is rewritten only where it changes: beta_rot is exactly 1 in systems
    // without rotational degrees of freedom, whose L stays in upload order
    bool rot = br != 1.0 || (explode && c.rotdof > 0);
    double L[3] = {0., 0., 0.};
    if (rot) L[0] = s.Lx[i], L[1] = s.Ly[i], L[2] = s.Lz[i];
    if (explode) {
        const double limit = 10.0 * Ti;
        const double ut = 0.5 * c.m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (ut > limit) {
            const double vc = sqrt(limit / ut);
            for (int d = 0; d < 3; ++d) {
                v[d] *= vc;
                F[d * n + i] *= vc;
            }
        }
        if (c.rotdof > 0) {
            const double lr = 3.0 * c.rotdof * Ti;
            const double q[4] = {s.qw[i], s.qx[i], s.qy[i], s.qz[i]};
            double w[3], Iw2 = 0.;
            qrotinv(q, L, w);
            for (int d = 0; d < 3; ++d) {
                w[d] *= c.invI[d];
                Iw2 += c.I[d] * w[d] * w[d];
            }
            const double ur = 0.5 * Iw2;
            if (ur > lr) {
                const double dc = sqrt(lr / ur);
                for (int d = 0; d < 3; ++d) {
                    L[d] *= dc;
                    M[d * n + i] *= dc;
                }
            }
        }
    }
    s.vx[i] = v[0] * bt; s.vy[i] = v[1] * bt; s.vz[i] = v[2] * bt;
    if (rot) {
        s.Lx[i] = L[0] * br; s.Ly[i] = L[1] * br; s.Lz[i] = L[2] * br;
    }
}

// Segmented fixed-order sum: block r adds rows [off[r], off[r+1]) of a
// [rows][width] array into out[r][width] (one block per replica).
__global__ __launch_bounds__(256) void k_seg_reduce(const double *__restrict__ in, int width,
                                                    const int *__restrict__ off, double *out) {
    __shared__ double red[256];
    const int b0 = off[blockIdx.x], b1 = off[blockIdx.x + 1];
    for (int w = 0; w < width; ++w) {
        double a = 0;
        for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) a += in[(size_t)b * width + w];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * width + w] = red[0];
        __syncthreads();
    }
}

__global__ void k_step_advance(int *step_idx) {
    if (threadIdx.x == 0 && blockIdx.x == 0) step_idx[0] += 1;
}

// Download helper: scatter a current-order array back to upload order.
__global__ void k_scatter_orig(int n, const int *__restrict__ orig, const double *__restrict__ a,
                               double *out, int stride_in, int width) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = orig[i];
    for (int w = 0; w < width; ++w) out[o * width + w] = a[w * stride_in + i];
}

}  // namespace mdx
