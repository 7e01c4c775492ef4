// mdx_capi.hip -- the C ABI of include/ls1amd.h: context, device buffers,
// model set-up (Component / Comp2Param restated) and the launch sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ls1amd.h"
#include "mdx_adjoint.h"
#include "mdx_internal.h"

namespace mdx {
__global__ void k_leapfrog_pre(MolArrays, const double *, const double *, const Component *, Reps,
                               double, int);
__global__ void k_leapfrog_post_pre(MolArrays, const double *, const double *, const Component *,
                                    Reps, double, int, const int *, const int *, double *);
__global__ void k_leapfrog_post(MolArrays, const double *, const double *, const Component *,
                                double, int, const int *, const int *, double *, int, int);
__global__ void k_seg_reduce(const double *, int, const int *, double *);
__global__ void k_reduce_partials(const double *, int, int, int, double *);
__global__ void k_fill_int(int *, int, int);
__global__ void k_fill_double(double *, int, double);
__global__ void k_bin_count(MolArrays, Reps, int, int *, int *, int *, int *);
__global__ void k_bin_scatter(int, const int *, const int *, const int *, int *);
__global__ void k_bin_gather(int, MolArrays, MolArrays, const int *, const int *, const int *,
                             int *, const int *, int *, const Component *, int *, int *, int *,
                             int *);
__global__ void k_bin_gather_point(int, MolArrays, MolArrays, const int *, const int *,
                                   const int *, int *, const int *, int *, const Component *, Reps,
                                   float *, float *, float *, int *);
__global__ void k_scan_block(ScanArgs);
__global__ void k_scan_add(ScanArgs);
__global__ void k_scan_total(ScanArgs);
__global__ void k_cell_site_start(int, const int *, ScanArgs);
__global__ void k_build_sites(int, MolArrays, Reps, const int *, const Component *, int, int,
                              int, SiteArrays, SiteArrays, SiteArrays, SiteArrays);
__global__ void k_calc_fm(int, MolArrays, const Component *, int, SiteArrays, SiteArrays,
                          SiteArrays, SiteArrays, double *, double *, double *, Reps,
                          const double *);
__global__ void k_velocity_scale(MolArrays, double *, double *, const Component *, Reps,
                                 const double *, int);
__global__ void k_scatter_orig(int, const int *, const double *, double *, int, int);
__global__ void k_normalize_q(int, MolArrays, long long);
__global__ void k_finalize_macro(const double *, double *, double *, const int *, int,
                                 const double *);
__global__ void k_finalize_lf(const double *, double *, const int *, int, double *, int);
__global__ void k_step_advance(int *);
__global__ void k_enum_pairs(Geom, int, int *);
}  // namespace mdx

using namespace mdx;

namespace {
thread_local std::string g_create_error;
constexpr int kBlock = 256;
inline int nblk(long n, int b = kBlock) { return (int)std::max<long>(1, (n + b - 1) / b); }

// Scratch of the entry points that allocate nothing (mdx_adjoint_noalloc):
// 64 MiB per device, held from the library's first use on a device until the
// process ends. Host code never uses the array's name as an address: each
// device's copy is resolved once, by mdx_create on that device, with
// hipGetSymbolAddress, and the resolved pointer is what kernels receive.
constexpr size_t kScratchDoubles = (size_t)8 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

struct mdx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool have_model = false, have_domain = false, have_state = false;
    int precision = MDX_PREC_FP64;
    bool counting = false, timing = false;
    // model
    Model model{};
    Component *d_comps = nullptr;
    std::vector<double> site_eps, site_sig, site_shift6, site_rcs, mix;
    bool has_mix = false;
    bool lj_at_origin = false;  // every LJ site sits at its molecule's centre
    // every component is one LJ centre at its centre and nothing else: site i
    // is molecule i, the gather builds the site cache, the pair kernel writes
    // the molecule force and virial (no calcFM pass)
    bool point_lj = false;
    // point_lj: q and L never change (no torque, invI = 0, beta_rot = 1) and
    // stay in upload order, shared by both state buffers; the per-update
    // normalisation of q (setupSoACache) is applied when q is read
    // (q_renorm = container updates since the last application)
    int64_t q_renorm = 0;
    // the cell lists (site starts, site cache) describe the current state:
    // false after an upload until the first container update
    bool binned = false;
    // NVT tail of the step (Domain::calculateGlobalValues, Homogeneous LRC,
    // velocity-scaling thermostat), off by default (the NVE path)
    bool lrc = false, thermostat = false;
    int heuristics = 1;
    std::vector<double> T_target;    // [nrep]; <= 0: no thermostat for that replica
    std::vector<double> comp_count;  // [nrep][ncomp] molecules per component
    double *d_thermo = nullptr;      // [nrep][kThermoW]
    // reduced-memory one-centre mode (VCP1CLJRMM + LeapfrogRMM): velocities at
    // half steps, the kick inside the pair kernel; 0 = off, else its dt
    double rmm_dt = 0.0;
    std::vector<double> eps24, sig2, shift6;
    std::vector<double> shift_rc;  // radius of the shift tied to (eps, sigma); 0 = none
    double *d_eps24 = nullptr, *d_sig2 = nullptr, *d_shift6 = nullptr, *d_shift_rc = nullptr;
    float *d_feps24 = nullptr, *d_fsig2 = nullptr, *d_fshift6 = nullptr;
    double *d_dpar = nullptr;  // interleaved {eps24, sig2, shift6, 0} per id pair
    float *d_fpar = nullptr;
    double eps_rf = 1.0;
    // domain: nrep replicas (1 for a single system; g = replica 0)
    Geom g{};
    double rc = 0, rc_lj = 0;
    int ncell = 0;      // global cells over all replicas
    int nrep = 1;
    std::vector<Geom> geoms;
    std::vector<int> cell_off;   // [nrep + 1]
    std::vector<int64_t> rep_n;  // molecules per replica (nrep > 1)
    Geom *d_geoms = nullptr;
    int *d_cell_off = nullptr, *d_cell_rep = nullptr;  // cell_rep: null for one replica
    int *d_mol_rep = nullptr;                          // state; null for one replica
    int *d_lf_chunk = nullptr, *d_lf_chunk_end = nullptr, *d_lf_off = nullptr;
    int lf_nchunk = 0;
    // pair-kernel launch shape: cells per work item, blocks per XCD
    // (0: one block per work item)
    int cpb = 1, blocks_per_xcd = 0;
    // state
    int n = 0;
    int nsites[kTypes] = {0, 0, 0, 0};
    MolArrays mol[2]{};
    int cur = 0;
    double *F = nullptr, *M = nullptr, *Vi = nullptr;
    int *cell = nullptr, *rank = nullptr, *count = nullptr, *start = nullptr, *perm = nullptr,
        *mcell = nullptr, *ls1id = nullptr, *mls1id = nullptr;
    double *red_tmp = nullptr;  // first level of the two-level reductions
    double *dl_tmp = nullptr;   // download staging (4 n)
    double *adj_site = nullptr, *adj_cot = nullptr, *adj_part = nullptr, *adj_out = nullptr;
    double *scratch = nullptr;  // this device's copy of the library scratch array
    int *nsite_mol[kTypes] = {nullptr, nullptr, nullptr, nullptr};
    int *ident = nullptr;  // 0 .. n (point_lj: molecule index of each site, site starts)
    SiteArrays sites[kTypes]{};
    // per scan level: kTypes segments of seg ints (block sums, then their scan)
    struct ScanLevel {
        int *p;
        size_t seg;
    };
    std::vector<ScanLevel> scan_tmp;
    int site_forms = 0;  // site record forms built for the current re-bin (1 FP64, 2 FP32)
    double *macro_part = nullptr, *lf_part = nullptr, *red = nullptr, *step_out = nullptr;
    int *step_idx = nullptr;
    int cap_steps = 0;
    unsigned long long *d_counters = nullptr;
    std::vector<uint64_t> ids;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double pair_ms = 0, total_ms = 0;
    // per i-site type pair kernel: an event before each launch and one after
    // the last, and the last timed traversal's times
    hipEvent_t kev[kTypes + 1] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int kev_order[kTypes] = {-1, -1, -1, -1};
    double kernel_ms[kTypes] = {0, 0, 0, 0};
    // captured steps per kind (full step, step fused with the next one's
    // first half, last step of a fused run) and state-buffer parity
    hipGraphExec_t gexec[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    double graph_dt = 0;
    std::vector<void *> allocs;        // lifetime of the context
    std::vector<void *> state_allocs;  // re-made by each upload
};

namespace {

int fail(mdx_context *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(ctx, MDX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != MDX_OK) return rc_; \
    } while (0)

template <typename T> int dalloc(mdx_context *c, T **p, size_t count, bool state = false) {
    void *q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, MDX_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    (state ? c->state_allocs : c->allocs).push_back(q);
    *p = (T *)q;
    return MDX_OK;
}

void free_list(std::vector<void *> &v) {
    for (void *p : v) (void)hipFree(p);
    v.clear();
}

void drop_graphs(mdx_context *c) {
    for (auto &gk : c->gexec)
        for (auto &g : gk)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

// Component::calculateLJshift (Component.cpp:118-122)
double lj_shift6(double eps, double sigma, double rc) {
    double s2 = sigma * sigma / (rc * rc);
    double s6 = s2 * s2 * s2;
    return 24.0 * eps * (s6 - s6 * s6);
}

// Parses the ls1 ASCII-header component records. Mass and moments follow
// Component::updateMassInertia (Component.cpp:151-179: only LJ centres and
// charges carry mass) and FullMolecule::updateMassInertia (FullMolecule.h:
// 88-99: invI = 0 for a zero moment); positive header moments override
// (ASCIIReader.cpp:197-202).
int parse_model(mdx_context *c, int ncomp, const double *desc, int len) {
    if (ncomp < 1 || ncomp > kMaxComp) return fail(c, MDX_ERR_ARG, "ncomp out of range [1,16]");
    Model &m = c->model;
    memset(&m, 0, sizeof(m));
    m.ncomp = ncomp;
    c->site_eps.clear();
    c->site_sig.clear();
    c->site_shift6.clear();
    c->site_rcs.clear();
    int p = 0, base = 0;
    for (int ci = 0; ci < ncomp; ++ci) {
        Component &cp = m.comp[ci];
        if (p + 4 > len) return fail(c, MDX_ERR_ARG, "component records truncated");
        for (int t = 0; t < kTypes; ++t) {
            cp.n[t] = (int)desc[p + t];
            if (cp.n[t] < 0 || cp.n[t] > kMaxSite)
                return fail(c, MDX_ERR_ARG, "site count out of range [0,8]");
        }
        p += 4;
        if (p + 8 * cp.n[0] + 5 * cp.n[1] + 7 * cp.n[2] + 7 * cp.n[3] + 3 > len)
            return fail(c, MDX_ERR_ARG, "component records truncated");
        double I6[3] = {0, 0, 0};
        int rotdof = 0;
        auto add_mass = [&](const double *r, double sm) {
            cp.m += sm;
            I6[0] += sm * (r[1] * r[1] + r[2] * r[2]);
            I6[1] += sm * (r[0] * r[0] + r[2] * r[2]);
            I6[2] += sm * (r[0] * r[0] + r[1] * r[1]);
            rotdof = 3;
            for (int d = 0; d < 3; ++d)
                if (I6[d] == 0.) --rotdof;
        };
        for (int s = 0; s < cp.n[0]; ++s, p += 8) {
            for (int d = 0; d < 3; ++d) cp.body[0][s][d] = desc[p + d];
            c->site_eps.push_back(desc[p + 4]);
            c->site_sig.push_back(desc[p + 5]);
            c->site_shift6.push_back(desc[p + 7] != 0. ? lj_shift6(desc[p + 4], desc[p + 5],
                                                                   desc[p + 6])
                                                       : 0.);
            c->site_rcs.push_back(desc[p + 7] != 0. ? desc[p + 6] : 0.);
            add_mass(desc + p, desc[p + 3]);
        }
        for (int s = 0; s < cp.n[1]; ++s, p += 5) {
            for (int d = 0; d < 3; ++d) cp.body[1][s][d] = desc[p + d];
            cp.param[1][s] = desc[p + 4];
            add_mass(desc + p, desc[p + 3]);
        }
        for (int t = 2; t < 4; ++t)
            for (int s = 0; s < cp.n[t]; ++s, p += 7) {
                for (int d = 0; d < 3; ++d) {
                    cp.body[t][s][d] = desc[p + d];
                    cp.e[t][s][d] = desc[p + 3 + d];
                }
                cp.param[t][s] = desc[p + 6];
            }
        for (int d = 0; d < 3; ++d) {
            cp.I[d] = desc[p + d] > 0. ? desc[p + d] : I6[d];
            cp.invI[d] = cp.I[d] != 0. ? 1. / cp.I[d] : 0.;
        }
        p += 3;
        cp.rotdof = rotdof;
        cp.point = cp.n[2] == 0 && cp.n[3] == 0 && cp.invI[0] == 0. && cp.invI[1] == 0. &&
                   cp.invI[2] == 0.;
        for (int t = 0; t < 2; ++t)
            for (int k = 0; k < cp.n[t]; ++k)
                for (int d = 0; d < 3; ++d)
                    if (cp.body[t][k][d] != 0.) cp.point = 0;
        cp.lj_base = base;
        base += cp.n[0];
        // massless components (only dipoles / quadrupoles) are valid for the
        // force traversal; FullMolecule::upd_preF asserts m > 0 only when
        // such a molecule is integrated (FullMolecule.cpp:332)
    }
    m.nlj = base;
    c->lj_at_origin = true;
    for (int ci = 0; ci < ncomp; ++ci)
        for (int s = 0; s < m.comp[ci].n[0]; ++s)
            for (int d = 0; d < 3; ++d)
                if (m.comp[ci].body[0][s][d] != 0.0) c->lj_at_origin = false;
    c->point_lj = c->lj_at_origin;
    for (int ci = 0; ci < ncomp; ++ci) {
        const Component &cp = m.comp[ci];
        if (cp.n[0] != 1 || cp.n[1] != 0 || cp.n[2] != 0 || cp.n[3] != 0) c->point_lj = false;
    }
    return MDX_OK;
}

// Comp2Param::initialize, LJ part (Comp2Param.cpp:11-110), in the layout of
// VectorizedCellProcessor's _eps_sig / _shift6 (VectorizedCellProcessor.cpp:
// 40-80): same component -> site i's own shift6; different components ->
// Lorentz-Berthelot with (xi, eta), always shifted at rcLJ.
void build_tables(mdx_context *c) {
    const Model &m = c->model;
    const int n = m.nlj;
    const std::vector<double> &eps = c->site_eps, &sig = c->site_sig;
    c->eps24.assign((size_t)n * n, 0.);
    c->sig2.assign((size_t)n * n, 0.);
    c->shift6.assign((size_t)n * n, 0.);
    c->shift_rc.assign((size_t)n * n, 0.);
    int k = 0;
    std::vector<double> xi(kMaxComp * kMaxComp, 1.0), eta(kMaxComp * kMaxComp, 1.0);
    for (int i = 0; i < m.ncomp; ++i)
        for (int j = i + 1; j < m.ncomp; ++j, ++k)
            if (c->has_mix) {
                xi[i * kMaxComp + j] = c->mix[2 * k];
                eta[i * kMaxComp + j] = c->mix[2 * k + 1];
            }
    for (int ci = 0; ci < m.ncomp; ++ci) {
        const Component &a = m.comp[ci];
        for (int si = 0; si < a.n[0]; ++si)
            for (int sj = 0; sj < a.n[0]; ++sj) {
                int ii = a.lj_base + si, jj = a.lj_base + sj;
                double s2 = .5 * (sig[ii] + sig[jj]);
                s2 *= s2;
                c->eps24[(size_t)ii * n + jj] = 24. * sqrt(eps[ii] * eps[jj]);
                c->sig2[(size_t)ii * n + jj] = s2;
                c->shift6[(size_t)ii * n + jj] = c->site_shift6[ii];
                // only a site's own entry has its shift built from the pair's
                // (eps, sigma); other same-component entries borrow site ii's
                if (si == sj) c->shift_rc[(size_t)ii * n + jj] = c->site_rcs[ii];
            }
        for (int cj = ci + 1; cj < m.ncomp; ++cj) {
            const Component &b = m.comp[cj];
            double x = xi[ci * kMaxComp + cj], h = eta[ci * kMaxComp + cj];
            for (int pass = 0; pass < 2; ++pass) {
                const Component &p1 = pass ? b : a, &p2 = pass ? a : b;
                for (int s1 = 0; s1 < p1.n[0]; ++s1)
                    for (int s2i = 0; s2i < p2.n[0]; ++s2i) {
                        int ii = p1.lj_base + s1, jj = p2.lj_base + s2i;
                        double e24 = 24. * (x * sqrt(eps[ii] * eps[jj]));
                        double s2 = 0.5 * (h * (sig[ii] + sig[jj]));
                        s2 *= s2;
                        double sr2 = s2 / (c->rc_lj * c->rc_lj);
                        double sr6 = sr2 * sr2 * sr2;
                        c->eps24[(size_t)ii * n + jj] = e24;
                        c->sig2[(size_t)ii * n + jj] = s2;
                        c->shift6[(size_t)ii * n + jj] = e24 * (sr6 - sr6 * sr6);
                        c->shift_rc[(size_t)ii * n + jj] = c->rc_lj;
                    }
            }
        }
    }
}

int upload_tables(mdx_context *c) {
    size_t nn = c->eps24.size();
    if (nn == 0) return MDX_OK;
    std::vector<float> fe(nn), fs(nn), fh(nn), fpar(4 * nn, 0.f);
    std::vector<double> dpar(4 * nn, 0.0);
    for (size_t i = 0; i < nn; ++i) {
        fe[i] = (float)c->eps24[i];
        fs[i] = (float)c->sig2[i];
        fh[i] = (float)c->shift6[i];
        dpar[4 * i] = c->eps24[i];
        dpar[4 * i + 1] = c->sig2[i];
        dpar[4 * i + 2] = c->shift6[i];
        fpar[4 * i] = fe[i];
        fpar[4 * i + 1] = fs[i];
        fpar[4 * i + 2] = fh[i];
    }
    HIPCHK(c, hipMemcpy(c->d_dpar, dpar.data(), 4 * nn * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_fpar, fpar.data(), 4 * nn * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_eps24, c->eps24.data(), nn * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_sig2, c->sig2.data(), nn * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_shift6, c->shift6.data(), nn * 8, hipMemcpyHostToDevice));
    if (c->shift_rc.size() == nn)
        HIPCHK(c, hipMemcpy(c->d_shift_rc, c->shift_rc.data(), nn * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_feps24, fe.data(), nn * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_fsig2, fs.data(), nn * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_fshift6, fh.data(), nn * 4, hipMemcpyHostToDevice));
    return MDX_OK;
}

// LinkedCells ctor (cell_mode 0, LinkedCells.cpp:38-106) or rebuild
// (cell_mode 1, float cell edge, LinkedCells.cpp:132-201).
int make_geom(mdx_context *c, const double *box, double rc, int periodic, int cell_mode, Geom &g,
              long &ncell_out) {
    memset(&g, 0, sizeof(g));
    long ncell = 1;
    for (int d = 0; d < 3; ++d) {
        if (!(box[d] > 0)) return fail(c, MDX_ERR_ARG, "box lengths must be positive");
        g.bbmin[d] = 0.0;
        g.bbmax[d] = box[d];
        g.L[d] = box[d];
        double rcc = cell_mode ? (double)(float)rc : rc;
        int nb = (int)floor(box[d] / rcc);
        if (cell_mode) {
            // rebuild: one cell per dimension is allowed (the halo images are
            // then copies of that cell), none is not (LinkedCells.cpp:153-164)
            if (nb == 0)
                return fail(c, MDX_ERR_ARG,
                            "region too small: no cell of edge rc fits (LinkedCells.cpp:160-164)");
        } else {
            if (nb == 0) nb = 1;
            if (nb < 2)
                return fail(c, MDX_ERR_ARG,
                            "bounding box too small for the cell length (LinkedCells needs at "
                            "least 2 cells per dimension, LinkedCells.cpp:79-96)");
        }
        g.n[d] = nb;
        g.cpd[d] = nb + 2;
        g.cl[d] = box[d] / nb;
        g.rcl[d] = nb / box[d];
        g.halo_min[d] = g.bbmin[d] - g.cl[d];
        g.halo_max[d] = g.bbmax[d] + g.cl[d];
        ncell *= nb;
    }
    if (ncell > (1L << 28)) return fail(c, MDX_ERR_CAPACITY, "too many cells");
    for (int d = 0; d < 3; ++d) {
        g.pdim[d] = periodic ? 1 : 0;
        g.reflect[d] = 0;
    }
    g.wall_rc = rc;
    ncell_out = ncell;
    return MDX_OK;
}

void set_cutoff_terms(mdx_context *c) {
    for (Geom *g : {&c->g}) {
        g->rc2 = c->rc * c->rc;
        g->rclj2 = c->rc_lj * c->rc_lj;
        // VectorizedCellProcessor ctor, VectorizedCellProcessor.cpp:21
        g->epsRFInvrc3 =
            2. * (c->eps_rf - 1.) / ((c->rc * c->rc * c->rc) * (2. * c->eps_rf + 1.));
    }
    for (Geom &g : c->geoms) {
        g.rc2 = c->g.rc2;
        g.rclj2 = c->g.rclj2;
        g.epsRFInvrc3 = c->g.epsRFInvrc3;
    }
}

// Homogeneous::init + calculateLongRange (src/longRange/Homogeneous.cpp:21-146)
// for count[i] molecules of component i in a box of volume `volume`:
// out = {UpotCorr, VirialCorr, VirialCorr / (3N)}. Only LJ table entries with
// shift6 == 0 contribute, as the reference reads uLJshift6 from Comp2Param;
// MySelbstTerm is the reaction-field self term of each molecule's dipole
// (charges' first moment plus point dipoles). The radial integrals are the
// reference's _TICC/_TICS/_TISS (Homogeneous.cpp:158-196).
namespace lrc {
double ticc_u(int n, double rc, double s2) { return -pow(rc, 2 * n + 3) / (pow(s2, n) * (2 * n + 3)); }
double tics_u(int n, double rc, double s2, double t) {
    return -(pow(rc + t, 2 * n + 3) - pow(rc - t, 2 * n + 3)) * rc /
               (4 * pow(s2, n) * t * (n + 1) * (2 * n + 3)) +
           (pow(rc + t, 2 * n + 4) - pow(rc - t, 2 * n + 4)) /
               (4 * pow(s2, n) * t * (n + 1) * (2 * n + 3) * (2 * n + 4));
}
double tiss_u(int n, double rc, double s2, double t1, double t2) {
    const double tp = t1 + t2, tm = t1 - t2;
    return -(pow(rc + tp, 2 * n + 4) - pow(rc + tm, 2 * n + 4) - pow(rc - tm, 2 * n + 4) +
             pow(rc - tp, 2 * n + 4)) *
               rc / (8 * pow(s2, n) * t1 * t2 * (n + 1) * (2 * n + 3) * (2 * n + 4)) +
           (pow(rc + tp, 2 * n + 5) - pow(rc + tm, 2 * n + 5) - pow(rc - tm, 2 * n + 5) +
            pow(rc - tp, 2 * n + 5)) /
               (8 * pow(s2, n) * t1 * t2 * (n + 1) * (2 * n + 3) * (2 * n + 4) * (2 * n + 5));
}
double ticc_v(int n, double rc, double s2) { return 2 * n * ticc_u(n, rc, s2); }
double tics_v(int n, double rc, double s2, double t) {
    return -(pow(rc + t, 2 * n + 2) - pow(rc - t, 2 * n + 2)) * rc * rc /
               (4 * pow(s2, n) * t * (n + 1)) -
           3 * tics_u(n, rc, s2, t);
}
double tiss_v(int n, double rc, double s2, double t1, double t2) {
    const double tp = t1 + t2, tm = t1 - t2;
    return -(pow(rc + tp, 2 * n + 3) - pow(rc + tm, 2 * n + 3) - pow(rc - tm, 2 * n + 3) +
             pow(rc - tp, 2 * n + 3)) *
               rc * rc / (8 * pow(s2, n) * t1 * t2 * (n + 1) * (2 * n + 3)) -
           3 * tiss_u(n, rc, s2, t1, t2);
}
}  // namespace lrc

int lrc_terms(mdx_context *c, const double *count, double volume, double *out) {
    const Model &m = c->model;
    const int nl = m.nlj;
    double upot = 0., vir = 0., self = 0., N = 0.;
    for (int i = 0; i < m.ncomp; ++i) {
        const Component &ci = m.comp[i];
        N += count[i];
        double cb[3] = {0, 0, 0};
        for (int s = 0; s < ci.n[T_C]; ++s)
            for (int d = 0; d < 3; ++d) cb[d] += ci.param[T_C][s] * ci.body[T_C][s][d];
        for (int s = 0; s < ci.n[T_D]; ++s) {
            double ev = 0.;
            for (int d = 0; d < 3; ++d) ev += ci.e[T_D][s][d] * ci.e[T_D][s][d];
            const double norm = 1.0 / sqrt(ev);
            for (int d = 0; d < 3; ++d) cb[d] += ci.param[T_D][s] * ci.e[T_D][s][d] * norm;
        }
        double my2 = 0.;
        for (int d = 0; d < 3; ++d) my2 += cb[d] * cb[d];
        self += my2 * count[i];
        for (int j = 0; j < m.ncomp; ++j) {
            const Component &cj = m.comp[j];
            for (int si = 0; si < ci.n[T_LJ]; ++si) {
                const double *a = ci.body[T_LJ][si];
                const double t1 = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
                for (int sj = 0; sj < cj.n[T_LJ]; ++sj) {
                    const double *b = cj.body[T_LJ][sj];
                    const double t2 = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
                    if (t1 + t2 >= c->rc_lj)
                        return fail(c, MDX_ERR_ARG, "long-range correction: rcLJ too small");
                    const size_t k = (size_t)(ci.lj_base + si) * nl + cj.lj_base + sj;
                    if (c->shift6[k] != 0.0) continue;
                    const double fac = count[i] * count[j] * c->eps24[k], s2 = c->sig2[k];
                    const double rc = c->rc_lj;
                    if (t1 == 0. && t2 == 0.) {
                        upot += fac * (lrc::ticc_u(-6, rc, s2) - lrc::ticc_u(-3, rc, s2));
                        vir += fac * (lrc::ticc_v(-6, rc, s2) - lrc::ticc_v(-3, rc, s2));
                    } else if (t1 != 0. && t2 != 0.) {
                        upot += fac * (lrc::tiss_u(-6, rc, s2, t1, t2) - lrc::tiss_u(-3, rc, s2, t1, t2));
                        vir += fac * (lrc::tiss_v(-6, rc, s2, t1, t2) - lrc::tiss_v(-3, rc, s2, t1, t2));
                    } else {
                        const double t = t2 == 0. ? t1 : t2;
                        upot += fac * (lrc::tics_u(-6, rc, s2, t) - lrc::tics_u(-3, rc, s2, t));
                        vir += fac * (lrc::tics_v(-6, rc, s2, t) - lrc::tics_v(-3, rc, s2, t));
                    }
                }
            }
        }
    }
    if (N <= 0) {
        out[0] = out[1] = out[2] = 0.;
        return MDX_OK;
    }
    const double rho = N / volume;
    const double fac = M_PI * rho / (3. * N);
    const double epsRFInvrc3 =
        2. * (c->eps_rf - 1.) / ((c->rc * c->rc * c->rc) * (2. * c->eps_rf + 1.));
    const double myself = -0.5 * epsRFInvrc3 * self;
    out[0] = fac * upot + myself;
    out[1] = -fac * vir + 3. * myself;
    out[2] = out[1] / (3. * N);
    return MDX_OK;
}

// Per-replica NVT record on the device: target T and the long-range terms;
// the step outputs start at T 0, betas 1, no explosion.
int upload_thermo(mdx_context *c) {
    const int R = c->nrep, nc = c->model.ncomp;
    std::vector<double> th((size_t)kThermoW * R, 0.0);
    for (int r = 0; r < R; ++r) {
        double *t = th.data() + (size_t)kThermoW * r;
        t[kThTarget] = r < (int)c->T_target.size() ? c->T_target[r] : 0.0;
        t[kThBetaT] = t[kThBetaR] = 1.0;
        if (c->lrc && (int)c->comp_count.size() == R * nc) {  // counts of the uploaded state
            const Geom &g = c->geoms[r];
            TRY(lrc_terms(c, c->comp_count.data() + (size_t)nc * r, g.L[0] * g.L[1] * g.L[2],
                          t + kThUcorr));
        }
    }
    HIPCHK(c, hipMemcpy(c->d_thermo, th.data(), 8 * th.size(), hipMemcpyHostToDevice));
    return MDX_OK;
}

// Replica geometries on the device; reduction buffers sized for nrep.
int upload_reps(mdx_context *c) {
    const int R = c->nrep;
    TRY(dalloc(c, &c->d_geoms, R));
    TRY(dalloc(c, &c->d_cell_off, R + 1));
    TRY(dalloc(c, &c->red, 12 * (size_t)R));
    TRY(dalloc(c, &c->d_thermo, (size_t)kThermoW * R));
    HIPCHK(c, hipMemcpy(c->d_geoms, c->geoms.data(), sizeof(Geom) * R, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_cell_off, c->cell_off.data(), 4 * (size_t)(R + 1),
                        hipMemcpyHostToDevice));
    c->d_cell_rep = nullptr;
    if (R > 1) {
        std::vector<int> cr((size_t)c->ncell);
        for (int r = 0; r < R; ++r)
            for (int k = c->cell_off[r]; k < c->cell_off[r + 1]; ++k) cr[k] = r;
        TRY(dalloc(c, &c->d_cell_rep, c->ncell));
        HIPCHK(c, hipMemcpy(c->d_cell_rep, cr.data(), 4 * cr.size(), hipMemcpyHostToDevice));
    }
    return MDX_OK;
}

Reps reps(const mdx_context *c) {
    Reps r;
    r.geom = c->d_geoms;
    r.cell_off = c->d_cell_off;
    r.mol_rep = c->d_mol_rep;
    r.cell_rep = c->d_cell_rep;
    r.nrep = c->nrep;
    return r;
}

// share: take q and L from this buffer instead of allocating them
int alloc_mol(mdx_context *c, MolArrays &m, int n, const MolArrays *share = nullptr) {
    double **dp[] = {&m.rx, &m.ry, &m.rz, &m.vx, &m.vy, &m.vz, &m.qw,
                     &m.qx, &m.qy, &m.qz, &m.Lx, &m.Ly, &m.Lz};
    for (int k = 0; k < 13; ++k) {
        if (share && k >= 6) {
            double *const src[] = {share->qw, share->qx, share->qy, share->qz,
                                   share->Lx, share->Ly, share->Lz};
            *dp[k] = src[k - 6];
        } else {
            TRY(dalloc(c, dp[k], n, true));
        }
    }
    TRY(dalloc(c, &m.cid, n, true));
    TRY(dalloc(c, &m.orig, n, true));
    return MDX_OK;
}

int alloc_sites(mdx_context *c, SiteArrays &s, int cap, int ncell, int n, int type) {
    memset(&s, 0, sizeof(s));
    if (c->point_lj) {  // the rest are views of molecule arrays (site_view)
        if (type != T_LJ || cap == 0) return MDX_OK;
        float **fp[] = {&s.fx, &s.fy, &s.fz};
        for (float **p : fp) TRY(dalloc(c, p, cap, true));
        return dalloc(c, &s.ljid, cap, true);
    }
    TRY(dalloc(c, &s.cstart, ncell + 1, true));
    TRY(dalloc(c, &s.mstart, n + 1, true));
    if (cap == 0) return MDX_OK;
    TRY(dalloc(c, &s.mol, cap, true));
    double **dp[] = {&s.x, &s.y, &s.z, &s.mx, &s.my, &s.mz,
                     &s.Fx, &s.Fy, &s.Fz, &s.Vx, &s.Vy, &s.Vz};
    for (double **p : dp) TRY(dalloc(c, p, cap, true));
    float **fp[] = {&s.fx, &s.fy, &s.fz, &s.fmx, &s.fmy, &s.fmz};
    for (float **p : fp) TRY(dalloc(c, p, cap, true));
    if (type == T_LJ) {
        TRY(dalloc(c, &s.ljid, cap, true));
    } else {
        TRY(dalloc(c, &s.p, cap, true));
        TRY(dalloc(c, &s.fp, cap, true));
    }
    if (type == T_C) TRY(dalloc(c, &s.nhead, cap, true));
    if (type >= T_D) {
        double **ep[] = {&s.ex, &s.ey, &s.ez, &s.Mx, &s.My, &s.Mz};
        for (double **p : ep) TRY(dalloc(c, p, cap, true));
        float **fe[] = {&s.fex, &s.fey, &s.fez};
        for (float **p : fe) TRY(dalloc(c, p, cap, true));
    }
    return MDX_OK;
}

int alloc_scan(mdx_context *c, int n) {
    c->scan_tmp.clear();
    int len = std::max(n, 1);
    while (true) {
        int nb = (len + 1023) / 1024;
        const size_t seg = (size_t)((nb + 1023) / 1024) * 1024 * 2 + 2048;
        int *p;
        TRY(dalloc(c, &p, seg * kTypes, true));
        c->scan_tmp.push_back({p, seg});
        if (nb <= 1) break;
        len = nb;
    }
    return MDX_OK;
}

// Exclusive scans of na arrays of a.n ints each in one set of launches
// (array t on blockIdx.y), out_t[n] = total of in_t, on the context stream.
int scan_exclusive(mdx_context *c, const ScanArgs &a, int na, int level = 0) {
    const int n = a.n;
    const int nb = std::max(1, (n + 1023) / 1024);
    if (level >= (int)c->scan_tmp.size()) return fail(c, MDX_ERR_CAPACITY, "scan depth");
    const mdx_context::ScanLevel &L = c->scan_tmp[level];
    ScanArgs blk = a, sub = a, add = a;
    sub.n = nb;
    for (int t = 0; t < na; ++t) {
        int *bsum = L.p + (size_t)t * L.seg;
        int *bscan = bsum + ((nb + 1023) / 1024) * 1024;
        blk.aux[t] = bsum;
        sub.in[t] = bsum;
        sub.out[t] = bscan;
        add.aux[t] = bscan;
    }
    hipLaunchKernelGGL(k_scan_block, dim3(nb, na), dim3(256), 0, c->stream, blk);
    if (nb > 1) {
        TRY(scan_exclusive(c, sub, na, level + 1));
        hipLaunchKernelGGL(k_scan_add, dim3(nb, na), dim3(256), 0, c->stream, add);
    }
    if (level == 0) hipLaunchKernelGGL(k_scan_total, dim3(1, na), dim3(1), 0, c->stream, a);
    return MDX_OK;
}

// Work items of cpb consecutive cells in 8 contiguous ranges (one per XCD);
// one block per item, or blocks_per_xcd persistent blocks per XCD.
// The quadrupole kernel takes cells two at a time, so its items are twice as
// long (same number of cell pairs per item as single cells per item).
int cells_per_item(const mdx_context *c, int t) { return t == T_Q ? 2 * c->cpb : c->cpb; }

int pair_grid(const mdx_context *c, int *items_per_xcd, int t = T_LJ) {
    const int cpb = cells_per_item(c, t);
    const int items = (c->ncell + cpb - 1) / cpb;
    *items_per_xcd = (items + 7) / 8;
    int per = *items_per_xcd;
    if (c->blocks_per_xcd > 0) per = std::min(per, c->blocks_per_xcd);
    return 8 * std::max(per, 1);
}

int n_present(const mdx_context *c) {
    int k = 0;
    for (int t = 0; t < kTypes; ++t) k += c->nsites[t] > 0;
    return k;
}

// Sum of nb blocks x width partials into out[width], fixed order.
void launch_reduce(mdx_context *c, const double *in, int nb, int width, double *out) {
    const int g1 = 256;
    if (nb <= 4 * g1) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->stream, in, nb, width, nb,
                           out);
        return;
    }
    const int per = (nb + g1 - 1) / g1;
    hipLaunchKernelGGL(k_reduce_partials, dim3(g1), dim3(256), 0, c->stream, in, nb, width, per,
                       c->red_tmp);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->stream, c->red_tmp, g1, width,
                       g1, out);
}

// The site arrays of type t as the kernels see them. For point_lj systems
// the FP64 site and molecule positions are the current molecule positions,
// the site outputs are the molecule force and virial rows, site i belongs to
// molecule i and the per-cell site starts are the molecule starts.
SiteArrays site_view(const mdx_context *c, int t) {
    SiteArrays s = c->sites[t];
    if (!c->point_lj || t != T_LJ) return s;
    const MolArrays &m = c->mol[c->cur];
    const size_t n = c->n;
    s.x = s.mx = m.rx;
    s.y = s.my = m.ry;
    s.z = s.mz = m.rz;
    s.fmx = s.fx;
    s.fmy = s.fy;
    s.fmz = s.fz;
    s.mol = s.mstart = c->ident;
    s.cstart = c->start;
    s.Fx = c->F;
    s.Fy = c->F + n;
    s.Fz = c->F + 2 * n;
    s.Vx = c->Vi;
    s.Vy = c->Vi + n;
    s.Vz = c->Vi + 2 * n;
    return s;
}

// The site record forms (1 FP64, 2 FP32) a re-bin builds: its precision's.
int rebin_forms(const mdx_context *c) {
    return c->precision == MDX_PREC_FP32_FP64ACC ? 2 : 1;
}

// Re-bin (LinkedCells::update + exchangeMolecules' periodic wrap, which
// leapfrog_pre already applied) and SoA cache build (setupSoACache).
int launch_update(mdx_context *c) {
    const int n = c->n;
    if (c->point_lj) ++c->q_renorm;
    MolArrays &src = c->mol[c->cur], &dst = c->mol[1 - c->cur];
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(c->ncell + 2)), dim3(kBlock), 0, c->stream, c->count,
                       c->ncell + 2, 0);
    hipLaunchKernelGGL(k_bin_count, dim3(nblk(n)), dim3(kBlock), 0, c->stream, src, reps(c), n,
                       c->cell, c->rank, c->count, c->ls1id);
    // the cells and the bin of deleted molecules behind them: start[ncell] is
    // the live count, start[ncell + 1] = n
    ScanArgs cells{};
    cells.in[0] = c->count;
    cells.out[0] = c->start;
    cells.n = c->ncell + 1;
    TRY(scan_exclusive(c, cells, 1));
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblk(n)), dim3(kBlock), 0, c->stream, n, c->cell,
                       c->rank, c->start, c->perm);
    if (c->point_lj) {
        const SiteArrays &s = c->sites[T_LJ];
        hipLaunchKernelGGL(k_bin_gather_point, dim3(nblk(n)), dim3(kBlock), 0, c->stream, n, src,
                           dst, c->cell, c->perm, c->start, c->mcell, c->ls1id, c->mls1id,
                           c->d_comps, reps(c), s.fx, s.fy, s.fz, s.ljid);
        c->cur = 1 - c->cur;
        c->binned = true;
        HIPCHK(c, hipGetLastError());
        return MDX_OK;
    }
    hipLaunchKernelGGL(k_bin_gather, dim3(nblk(n)), dim3(kBlock), 0, c->stream, n, src, dst,
                       c->cell, c->perm, c->start, c->mcell, c->ls1id, c->mls1id, c->d_comps,
                       c->nsite_mol[0],
                       c->nsite_mol[1], c->nsite_mol[2], c->nsite_mol[3]);
    c->cur = 1 - c->cur;
    // per site type: molecule site starts (scan of the per-molecule site
    // counts), then cell site starts, all present types in one launch each
    ScanArgs ms{}, cs{};
    int na = 0;
    for (int t = 0; t < kTypes; ++t) {
        if (c->nsites[t] == 0) continue;
        ms.in[na] = c->nsite_mol[t];
        ms.out[na] = c->sites[t].mstart;
        cs.in[na] = c->sites[t].mstart;
        cs.out[na] = c->sites[t].cstart;
        ++na;
    }
    ms.n = n;
    if (na > 0) {
        TRY(scan_exclusive(c, ms, na));
        hipLaunchKernelGGL(k_cell_site_start, dim3(nblk(c->ncell + 1), na), dim3(kBlock), 0,
                           c->stream, c->ncell, c->start, cs);
    }
    // the site records of the current precision form only; another form is
    // added on demand (ensure_sites)
    const int nc = c->model.ncomp;
    const int forms = rebin_forms(c);
    hipLaunchKernelGGL(k_build_sites, dim3(nblk(n)), dim3(kBlock), nc * sizeof(Component),
                       c->stream, n, c->mol[c->cur], reps(c), c->mcell, c->d_comps, nc, forms,
                       1, c->sites[0], c->sites[1], c->sites[2], c->sites[3]);
    c->site_forms = forms;
    c->binned = true;
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

// Traversals read the cell lists: a state not binned since its upload is
// binned first (LinkedCells' update before the first traversal).
int ensure_binned(mdx_context *c) {
    if (!c->binned) TRY(launch_update(c));
    return MDX_OK;
}

// The site records of form bit(s) need (1 FP64, 2 FP32) for the current
// re-bin: the re-bin builds its precision's form; a traversal in the other
// precision or the (FP64) adjoint adds the missing one from the same q.
// One-centre systems read molecule arrays and the gather's FP32 cache.
int ensure_sites(mdx_context *c, int need) {
    if (c->point_lj || (c->site_forms & need) == need) return MDX_OK;
    const int nc = c->model.ncomp, n = c->n;
    hipLaunchKernelGGL(k_build_sites, dim3(nblk(n)), dim3(kBlock), nc * sizeof(Component),
                       c->stream, n, c->mol[c->cur], reps(c), c->mcell, c->d_comps, nc,
                       need & ~c->site_forms, 0, c->sites[0], c->sites[1], c->sites[2],
                       c->sites[3]);
    c->site_forms |= need;
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

// kick != 0 (reduced-memory mode): the pair kernel adds kick * F to the
// velocities instead of storing F and Vi.
int launch_traverse(mdx_context *c, double kick = 0.0) {
    TRY(ensure_sites(c, rebin_forms(c)));
    PairArgs a;
    memset(&a, 0, sizeof(a));
    a.R = reps(c);
    a.rc2 = c->g.rc2;
    a.rclj2 = c->g.rclj2;
    a.epsRFInvrc3 = c->g.epsRFInvrc3;
    for (int t = 0; t < kTypes; ++t) {
        a.s[t] = site_view(c, t);
        a.present[t] = c->nsites[t] > 0;
    }
    a.vscale = c->point_lj ? 0.5 : 1.0;
    a.thermo = (c->lrc && c->point_lj) ? c->d_thermo : nullptr;
    a.mrx = c->mol[c->cur].rx;
    a.mry = c->mol[c->cur].ry;
    a.mrz = c->mol[c->cur].rz;
    if (kick != 0.0) {
        const MolArrays &m = c->mol[c->cur];
        a.kvx = m.vx;
        a.kvy = m.vy;
        a.kvz = m.vz;
        a.kick = kick;
    }
    a.pp.eps24 = c->d_eps24;
    a.pp.sig2 = c->d_sig2;
    a.pp.shift6 = c->d_shift6;
    a.pp.feps24 = c->d_feps24;
    a.pp.fsig2 = c->d_fsig2;
    a.pp.fshift6 = c->d_fshift6;
    a.pp.dpar = c->d_dpar;
    a.pp.fpar = c->d_fpar;
    a.pp.nlj = c->model.nlj;
    a.ncell = c->ncell;
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    // one [ncell][4] partial array: the first i-site type writes each cell's
    // row, later types add to it (the launches are ordered on the stream)
    int slot = 0;
    for (int t = 0; t < kTypes; ++t) {
        if (c->nsites[t] == 0) continue;
        a.macro_part = c->macro_part;
        a.accumulate = slot > 0;
        a.counters = c->d_counters + 8 * t;  // each kernel its own row
        a.cpb = cells_per_item(c, t);
        const int grid = pair_grid(c, &a.items_per_xcd, t);
        PairKernel k = pair_kernel(t, c->precision == MDX_PREC_FP32_FP64ACC, c->counting,
                                   t == T_LJ && c->lj_at_origin, c->model.nlj > 1);
        if (c->timing) {
            HIPCHK(c, hipEventRecord(c->kev[slot], c->stream));
            c->kev_order[slot] = t;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, c->stream, a);
        ++slot;
    }
    if (c->timing) {
        HIPCHK(c, hipEventRecord(c->kev[slot], c->stream));
        for (int k = slot; k < kTypes; ++k) c->kev_order[k] = -1;
    }
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    const int R = c->nrep;
    if (slot == 0)
        hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, c->stream, c->red, 4 * R, 0.0);
    else if (R == 1)
        launch_reduce(c, c->macro_part, c->ncell, 4, c->red);
    else
        hipLaunchKernelGGL(k_seg_reduce, dim3(R), dim3(256), 0, c->stream, c->macro_part, 4,
                           c->d_cell_off, c->red);
    // point_lj: the pair kernel wrote F and Vi; M stays 0 (no site offsets)
    if (!c->point_lj)
        hipLaunchKernelGGL(k_calc_fm, dim3(nblk(c->n)), dim3(kBlock),
                           c->model.ncomp * sizeof(Component), c->stream, c->n,
                           c->mol[c->cur], c->d_comps, c->model.ncomp, c->sites[0],
                           c->sites[1], c->sites[2], c->sites[3], c->F, c->M, c->Vi, reps(c),
                           c->lrc ? c->d_thermo : nullptr);
    hipLaunchKernelGGL(k_finalize_macro, dim3((R + 63) / 64), dim3(64), 0, c->stream, c->red,
                       c->red + 4 * R, c->step_out, c->step_idx, R,
                       c->lrc ? c->d_thermo : nullptr);
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

// The domain has an outflow wall: molecules can be deleted, the live count
// can drop below the uploaded one.
bool open_domain(const mdx_context *c) {
    return c->g.outflow[0] | c->g.outflow[1] | c->g.outflow[2];
}

int launch_lf_pre(mdx_context *c, double dt) {
    // reduced-memory mode: PositionCellProcessorRMM, r += dt v only
    hipLaunchKernelGGL(k_leapfrog_pre, dim3(nblk(c->n)), dim3(kBlock), 0, c->stream,
                       c->mol[c->cur], c->rmm_dt != 0.0 ? nullptr : c->F, c->M, c->d_comps,
                       reps(c), dt, c->n);
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

int launch_lf_post(mdx_context *c, double dt) {
    const int R = c->nrep, nb = c->lf_nchunk;
    hipLaunchKernelGGL(k_leapfrog_post, dim3(nb), dim3(kBlock), 0, c->stream, c->mol[c->cur], c->F,
                       c->M, c->d_comps, dt, c->n, c->d_lf_chunk, c->d_lf_chunk_end, c->lf_part,
                       c->rmm_dt != 0.0 ? 1 : 0, open_domain(c) ? 1 : 0);
    if (R == 1)
        launch_reduce(c, c->lf_part, nb, 4, c->red + 8);
    else
        hipLaunchKernelGGL(k_seg_reduce, dim3(R), dim3(256), 0, c->stream, c->lf_part, 4,
                           c->d_lf_off, c->red + 8 * R);
    hipLaunchKernelGGL(k_finalize_lf, dim3((R + 63) / 64), dim3(64), 0, c->stream, c->red + 8 * R,
                       c->step_out, c->step_idx, R, c->thermostat ? c->d_thermo : nullptr,
                       c->heuristics);
    if (c->thermostat)
        hipLaunchKernelGGL(k_velocity_scale, dim3(nblk(c->n)), dim3(kBlock), 0, c->stream,
                           c->mol[c->cur], c->F, c->M, c->d_comps, reps(c), c->d_thermo, c->n);
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

// Step n's second half kick and step n + 1's first half kick and drift in
// one pass (no thermostat acts between them).
int launch_lf_post_pre(mdx_context *c, double dt) {
    const int R = c->nrep, nb = c->lf_nchunk;
    hipLaunchKernelGGL(k_leapfrog_post_pre, dim3(nb), dim3(kBlock), 0, c->stream, c->mol[c->cur],
                       c->F, c->M, c->d_comps, reps(c), dt, c->n, c->d_lf_chunk,
                       c->d_lf_chunk_end, c->lf_part);
    if (R == 1)
        launch_reduce(c, c->lf_part, nb, 4, c->red + 8);
    else
        hipLaunchKernelGGL(k_seg_reduce, dim3(R), dim3(256), 0, c->stream, c->lf_part, 4,
                           c->d_lf_off, c->red + 8 * R);
    hipLaunchKernelGGL(k_finalize_lf, dim3((R + 63) / 64), dim3(64), 0, c->stream, c->red + 8 * R,
                       c->step_out, c->step_idx, R, nullptr, c->heuristics);
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

// Kinds of step launch: a whole step; a step whose first half kick was done
// by the previous one and which does the next one's (fused); the last step
// of such a run.
enum { kStepFull = 0, kStepFused = 1, kStepLast = 2 };

int launch_step(mdx_context *c, double dt, int kind = kStepFull) {
    if (kind == kStepFull) TRY(launch_lf_pre(c, dt));
    TRY(launch_update(c));
    if (c->rmm_dt != 0.0)  // LeapfrogRMM: one full kick dt/m inside the force kernel
        TRY(launch_traverse(c, dt / c->model.comp[0].m));
    else
        TRY(launch_traverse(c));
    if (kind == kStepFused)
        TRY(launch_lf_post_pre(c, dt));
    else
        TRY(launch_lf_post(c, dt));
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, c->stream, c->step_idx);
    HIPCHK(c, hipGetLastError());
    return MDX_OK;
}

int need(mdx_context *c, bool model, bool domain, bool state) {
    if (model && !c->have_model) return fail(c, MDX_ERR_STATE, "no components set");
    if (domain && !c->have_domain) return fail(c, MDX_ERR_STATE, "no domain set");
    if (state && !c->have_state) return fail(c, MDX_ERR_STATE, "no molecules uploaded");
    return MDX_OK;
}

int ensure_steps(mdx_context *c, int nsteps) {
    if (nsteps + 1 <= c->cap_steps) return MDX_OK;
    int cap = std::max(nsteps + 1, 2 * c->cap_steps);
    double *p;
    TRY(dalloc(c, &p, (size_t)6 * cap * c->nrep));
    c->step_out = p;  // the old buffer stays owned by the context until destroy
    c->cap_steps = cap;
    drop_graphs(c);
    return MDX_OK;
}

int state_setup(mdx_context *c, int64_t n64, const int32_t *cid_host) {
    if (n64 <= 0 || n64 > (1 << 27)) return fail(c, MDX_ERR_ARG, "molecule count out of range");
    if (c->nrep > 1) {
        int64_t tot = 0;
        for (int64_t k : c->rep_n) tot += k;
        if (tot != n64) return fail(c, MDX_ERR_ARG, "n differs from the sum of replica counts");
    }
    const int n = (int)n64;
    drop_graphs(c);
    c->rmm_dt = 0.0;  // a new state starts with full-step velocities
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_list(c->state_allocs);
    c->n = n;
    for (int t = 0; t < kTypes; ++t) c->nsites[t] = 0;
    for (int i = 0; i < n; ++i) {
        int k = cid_host[i];
        if (k < 0 || k >= c->model.ncomp) return fail(c, MDX_ERR_ARG, "component id out of range");
        for (int t = 0; t < kTypes; ++t) c->nsites[t] += c->model.comp[k].n[t];
    }
    for (int t = 0; t < kTypes; ++t)
        if (c->nsites[t] >= (1 << 27)) return fail(c, MDX_ERR_CAPACITY, "too many sites");
    c->binned = false;
    TRY(alloc_mol(c, c->mol[0], n));
    TRY(alloc_mol(c, c->mol[1], n, c->point_lj ? &c->mol[0] : nullptr));
    c->q_renorm = 0;
    TRY(dalloc(c, &c->F, 3 * (size_t)n, true));
    TRY(dalloc(c, &c->M, 3 * (size_t)n, true));
    TRY(dalloc(c, &c->Vi, 3 * (size_t)n, true));
    int *ip[] = {nullptr};
    (void)ip;
    TRY(dalloc(c, &c->cell, n, true));
    TRY(dalloc(c, &c->rank, n, true));
    TRY(dalloc(c, &c->perm, n, true));
    TRY(dalloc(c, &c->mcell, n, true));
    TRY(dalloc(c, &c->ls1id, n, true));
    TRY(dalloc(c, &c->mls1id, n, true));
    // one bin per cell, one behind the last cell for molecules deleted at an
    // outflow wall, and the scan total
    TRY(dalloc(c, &c->count, c->ncell + 2, true));
    TRY(dalloc(c, &c->start, c->ncell + 2, true));
    for (int t = 0; t < kTypes; ++t) {
        c->nsite_mol[t] = nullptr;
        if (!c->point_lj) TRY(dalloc(c, &c->nsite_mol[t], n + 1, true));
        TRY(alloc_sites(c, c->sites[t], c->nsites[t], c->ncell, n, t));
    }
    if (c->point_lj) {
        std::vector<int> id((size_t)n + 1);
        for (int i = 0; i <= n; ++i) id[i] = i;
        TRY(dalloc(c, &c->ident, (size_t)n + 1, true));
        HIPCHK(c, hipMemcpy(c->ident, id.data(), 4 * id.size(), hipMemcpyHostToDevice));
    }
    TRY(alloc_scan(c, std::max(n, c->ncell) + 1));
    TRY(dalloc(c, &c->macro_part, (size_t)c->ncell * 4, true));
    // molecule chunks of <= kBlock that never straddle two replicas
    {
        std::vector<int> ch, che, off(1, 0);
        std::vector<int64_t> rn = c->nrep > 1 ? c->rep_n : std::vector<int64_t>(1, n);
        int64_t m0 = 0;
        for (int64_t cnt : rn) {
            for (int64_t k = 0; k < cnt; k += kBlock) {
                ch.push_back((int)(m0 + k));
                che.push_back((int)(m0 + cnt));
            }
            m0 += cnt;
            off.push_back((int)ch.size());
        }
        if (ch.empty()) {
            ch.push_back(0);
            che.push_back(0);
        }
        c->lf_nchunk = (int)ch.size();
        TRY(dalloc(c, &c->d_lf_chunk, ch.size(), true));
        TRY(dalloc(c, &c->d_lf_chunk_end, ch.size(), true));
        TRY(dalloc(c, &c->d_lf_off, off.size(), true));
        HIPCHK(c, hipMemcpy(c->d_lf_chunk, ch.data(), 4 * ch.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_lf_chunk_end, che.data(), 4 * che.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_lf_off, off.data(), 4 * off.size(), hipMemcpyHostToDevice));
        TRY(dalloc(c, &c->lf_part, (size_t)c->lf_nchunk * 4, true));
    }
    c->d_mol_rep = nullptr;
    if (c->nrep > 1) {
        std::vector<int> mr((size_t)n);
        int64_t m0 = 0;
        for (int r = 0; r < c->nrep; ++r) {
            for (int64_t k = 0; k < c->rep_n[r]; ++k) mr[m0 + k] = r;
            m0 += c->rep_n[r];
        }
        TRY(dalloc(c, &c->d_mol_rep, n, true));
        HIPCHK(c, hipMemcpy(c->d_mol_rep, mr.data(), 4 * mr.size(), hipMemcpyHostToDevice));
    }
    TRY(dalloc(c, &c->dl_tmp, (size_t)n * 4, true));
    c->adj_site = c->adj_cot = c->adj_part = c->adj_out = nullptr;  // made on first use
    HIPCHK(c, hipMemsetAsync(c->F, 0, 3 * (size_t)n * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->M, 0, 3 * (size_t)n * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->Vi, 0, 3 * (size_t)n * 8, c->stream));
    c->cur = 0;
    // molecules per component and replica (the long-range correction's N_i)
    {
        const int nc = c->model.ncomp;
        c->comp_count.assign((size_t)c->nrep * nc, 0.0);
        int64_t m0 = 0;
        for (int r = 0; r < c->nrep; ++r) {
            const int64_t cnt = c->nrep > 1 ? c->rep_n[r] : n64;
            for (int64_t k = 0; k < cnt; ++k) c->comp_count[(size_t)r * nc + cid_host[m0 + k]] += 1.0;
            m0 += cnt;
        }
    }
    return upload_thermo(c);
}

}  // namespace

extern "C" {

const char *mdx_last_error(const mdx_context *ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int mdx_create(int device, mdx_context **out) {
    if (!out) return MDX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "no HIP device";
        return MDX_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
        g_create_error = "device index out of range";
        return MDX_ERR_ARG;
    }
    mdx_context *c = new mdx_context();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        g_create_error = "hipSetDevice / hipStreamCreate failed";
        delete c;
        return MDX_ERR_HIP;
    }
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    for (auto &e : c->kev) (void)hipEventCreate(&e);
    if (dalloc(c, &c->red_tmp, 256 * 8) || dalloc(c, &c->d_counters, 8 * kTypes) ||
        dalloc(c, &c->step_idx, 1) ||
        dalloc(c, &c->d_comps, kMaxComp) || ensure_steps(c, 1)) {
        g_create_error = c->err;
        free_list(c->allocs);
        delete c;
        return MDX_ERR_HIP;
    }
    (void)hipMemset(c->step_idx, 0, sizeof(int));
    (void)hipMemset(c->d_counters, 0, 8 * kTypes * sizeof(unsigned long long));
    if (device < kMaxDevices) {
        // once per device, with that device current
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[device] && g_scratch_err[device] == hipSuccess) {
            void *p = nullptr;
            g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (g_scratch_err[device] == hipSuccess)
                g_scratch_ptr[device] = (double *)p;
            else
                (void)hipGetLastError();  // keep the failure out of later launch checks
        }
        c->scratch = g_scratch_ptr[device];
    }
    *out = c;
    return MDX_OK;
}

int mdx_destroy(mdx_context *c) {
    if (!c) return MDX_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    free_list(c->state_allocs);
    free_list(c->allocs);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->kev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MDX_OK;
}

int mdx_set_components(mdx_context *c, int ncomp, const double *desc, int desc_len,
                       const double *mix, double eps_rf) {
    if (!c || !desc) return MDX_ERR_ARG;
    if (c->have_state) return fail(c, MDX_ERR_STATE, "components must be set before upload");
    TRY(parse_model(c, ncomp, desc, desc_len));
    if (c->model.nlj > 4096) return fail(c, MDX_ERR_CAPACITY, "too many LJ centres");
    c->eps_rf = eps_rf;
    c->has_mix = mix != nullptr;
    c->mix.assign(mix ? mix : desc, mix ? mix + ncomp * (ncomp - 1) : desc);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(c->d_comps, c->model.comp, sizeof(Component) * ncomp,
                        hipMemcpyHostToDevice));
    int nn = std::max(1, c->model.nlj * c->model.nlj);
    TRY(dalloc(c, &c->d_eps24, nn));
    TRY(dalloc(c, &c->d_sig2, nn));
    TRY(dalloc(c, &c->d_shift6, nn));
    TRY(dalloc(c, &c->d_shift_rc, nn));
    TRY(dalloc(c, &c->d_feps24, nn));
    TRY(dalloc(c, &c->d_fsig2, nn));
    TRY(dalloc(c, &c->d_fshift6, nn));
    TRY(dalloc(c, &c->d_dpar, 4 * nn));
    TRY(dalloc(c, &c->d_fpar, 4 * nn));
    c->have_model = true;
    c->have_domain = false;
    return MDX_OK;
}

int mdx_set_domain(mdx_context *c, const double *box, double rc, double rc_lj, int periodic,
                   int cell_mode) {
    if (!c || !box) return MDX_ERR_ARG;
    if (!c->have_model) return fail(c, MDX_ERR_STATE, "set components before the domain");
    if (c->have_state) return fail(c, MDX_ERR_STATE, "domain must be set before upload");
    if (!(rc > 0) || !(rc_lj > 0)) return fail(c, MDX_ERR_ARG, "cutoffs must be positive");
    if (rc_lj > rc)
        return fail(c, MDX_ERR_ARG, "rcLJ larger than the cell cutoff rc is not supported");
    c->rc = rc;
    c->rc_lj = rc_lj;
    long nc = 0;
    TRY(make_geom(c, box, rc, periodic, cell_mode, c->g, nc));
    c->ncell = (int)nc;
    c->nrep = 1;
    c->geoms.assign(1, c->g);
    c->cell_off = {0, c->ncell};
    c->rep_n.clear();
    drop_graphs(c);
    c->cap_steps = 0;
    TRY(ensure_steps(c, 1));
    set_cutoff_terms(c);
    c->T_target.clear();  // thermostat settings belong to a domain / replica set
    c->comp_count.clear();
    c->lrc = c->thermostat = false;
    TRY(upload_reps(c));
    TRY(upload_thermo(c));
    build_tables(c);
    TRY(upload_tables(c));
    c->have_domain = true;
    return MDX_OK;
}

int mdx_set_replicas(mdx_context *c, int nrep, const double *boxes, const int64_t *counts,
                     double rc, double rc_lj, int periodic, int cell_mode) {
    if (!c || !boxes || !counts || nrep < 1) return MDX_ERR_ARG;
    if (!c->have_model) return fail(c, MDX_ERR_STATE, "set components before the replicas");
    if (c->have_state) return fail(c, MDX_ERR_STATE, "replicas must be set before upload");
    if (!(rc > 0) || !(rc_lj > 0)) return fail(c, MDX_ERR_ARG, "cutoffs must be positive");
    if (rc_lj > rc)
        return fail(c, MDX_ERR_ARG, "rcLJ larger than the cell cutoff rc is not supported");
    c->rc = rc;
    c->rc_lj = rc_lj;
    std::vector<Geom> gs((size_t)nrep);
    std::vector<int> off(1, 0);
    long tot = 0;
    for (int r = 0; r < nrep; ++r) {
        if (counts[r] < 0) return fail(c, MDX_ERR_ARG, "negative replica count");
        long nc = 0;
        TRY(make_geom(c, boxes + 3 * r, rc, periodic, cell_mode, gs[r], nc));
        tot += nc;
        if (tot > (1L << 28)) return fail(c, MDX_ERR_CAPACITY, "too many cells");
        off.push_back((int)tot);
    }
    c->geoms = gs;
    c->g = gs[0];
    c->cell_off = off;
    c->ncell = (int)tot;
    c->nrep = nrep;
    c->rep_n.assign(counts, counts + nrep);
    drop_graphs(c);
    c->cap_steps = 0;  // step records are [step][nrep][6]: re-size for nrep
    TRY(ensure_steps(c, 1));
    set_cutoff_terms(c);
    build_tables(c);
    TRY(upload_tables(c));
    c->T_target.clear();  // thermostat settings belong to a domain / replica set
    c->comp_count.clear();
    c->lrc = c->thermostat = false;
    TRY(upload_reps(c));
    TRY(upload_thermo(c));
    c->have_domain = true;
    return MDX_OK;
}

int mdx_set_pair_config(mdx_context *c, int cells_per_block, int blocks_per_xcd) {
    if (!c || cells_per_block < 1 || cells_per_block > 64 || blocks_per_xcd < 0)
        return fail(c, MDX_ERR_ARG, "cells_per_block in [1, 64], blocks_per_xcd >= 0");
    c->cpb = cells_per_block;
    c->blocks_per_xcd = blocks_per_xcd;
    drop_graphs(c);
    return MDX_OK;
}

int mdx_set_thermostat(mdx_context *c, int nrep, const double *T_target, int lrc,
                       int explosion_heuristics) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    if (nrep != c->nrep) return fail(c, MDX_ERR_ARG, "nrep differs from the replica count");
    const bool any_t = T_target && std::any_of(T_target, T_target + nrep, [](double t) { return t > 0; });
    if (open_domain(c) && (any_t || lrc))
        return fail(c, MDX_ERR_ARG,
                    "the thermostat and the long-range correction keep the molecule count fixed; "
                    "the domain has outflow walls");
    HIPCHK(c, hipSetDevice(c->device));
    c->T_target.assign((size_t)nrep, 0.0);
    c->thermostat = false;
    if (T_target)
        for (int r = 0; r < nrep; ++r) {
            c->T_target[r] = T_target[r];
            if (T_target[r] > 0.0) c->thermostat = true;
        }
    c->lrc = lrc != 0;
    c->heuristics = explosion_heuristics != 0;
    drop_graphs(c);
    return upload_thermo(c);
}

int mdx_set_boundaries(mdx_context *c, const int32_t *types) {
    if (!c || !types) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    bool outflow = false;
    for (int d = 0; d < 3; ++d) {
        if (types[d] != MDX_BOUNDARY_PERIODIC && types[d] != MDX_BOUNDARY_REFLECTING &&
            types[d] != MDX_BOUNDARY_OUTFLOW)
            return fail(c, MDX_ERR_ARG, "unknown boundary type");
        outflow = outflow || types[d] == MDX_BOUNDARY_OUTFLOW;
    }
    if (c->rmm_dt != 0.0)
        return fail(c, MDX_ERR_STATE, "walls need the stored forces (not in reduced-memory mode)");
    if (outflow && c->nrep > 1)
        return fail(c, MDX_ERR_ARG, "outflow walls serve single systems (not batched replicas)");
    if (outflow && (c->thermostat || c->lrc))
        return fail(c, MDX_ERR_ARG,
                    "outflow walls change the molecule count; the thermostat and the long-range "
                    "correction keep it fixed");
    HIPCHK(c, hipSetDevice(c->device));
    for (Geom *g : {&c->g}) {
        for (int d = 0; d < 3; ++d) {
            g->pdim[d] = types[d] == MDX_BOUNDARY_PERIODIC;
            g->reflect[d] = types[d] == MDX_BOUNDARY_REFLECTING;
            g->outflow[d] = types[d] == MDX_BOUNDARY_OUTFLOW;
        }
    }
    for (Geom &g : c->geoms)
        for (int d = 0; d < 3; ++d) {
            g.pdim[d] = c->g.pdim[d];
            g.reflect[d] = c->g.reflect[d];
            g.outflow[d] = c->g.outflow[d];
        }
    HIPCHK(c, hipMemcpy(c->d_geoms, c->geoms.data(), sizeof(Geom) * c->nrep, hipMemcpyHostToDevice));
    drop_graphs(c);
    return MDX_OK;
}

int mdx_set_rmm(mdx_context *c, double dt) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    if (dt == 0.0) {
        c->rmm_dt = 0.0;
        drop_graphs(c);
        return MDX_OK;
    }
    if (!(dt > 0.0)) return fail(c, MDX_ERR_ARG, "dt must be positive");
    if (c->model.ncomp != 1 || !c->point_lj)
        return fail(c, MDX_ERR_ARG, "reduced-memory mode: one component of one LJ centre only");
    for (int d = 0; d < 3; ++d)
        if (c->g.reflect[d] || c->g.outflow[d])
            return fail(c, MDX_ERR_STATE, "reduced-memory mode stores no forces for the wall rule");
    if (c->rmm_dt != 0.0) return fail(c, MDX_ERR_STATE, "reduced-memory mode is already on");
    HIPCHK(c, hipSetDevice(c->device));
    drop_graphs(c);
    // Simulation::prepare_start (Simulation.cpp:902-937): the initial force
    // traversal with dt/m halved moves the velocities to the half step
    TRY(launch_update(c));
    TRY(launch_traverse(c, 0.5 * (dt / c->model.comp[0].m)));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->rmm_dt = dt;
    return MDX_OK;
}

int mdx_thermostat_state(mdx_context *c, double *out) {
    if (!c || !out) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_thermo, sizeof(double) * kThermoW * c->nrep,
                        hipMemcpyDeviceToHost));
    return MDX_OK;
}

int mdx_autotune(mdx_context *c, int reps, int32_t *chosen, double *best_ms) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    HIPCHK(c, hipSetDevice(c->device));
    TRY(ensure_binned(c));
    if (reps < 1) reps = 3;
    const bool was_timing = c->timing, was_counting = c->counting;
    c->timing = true;
    c->counting = false;
    // persistent grids below one full round of blocks (16 one-wave blocks
    // per CU, 512 per XCD) leave CUs idle: 64-256 blocks per XCD measured
    // 1.6-6x slower than one block per item on the mixture and the replicas
    const int cpbs[] = {1, 2, 4, 8};
    const int bpx[] = {0, 512, 1024};
    std::vector<std::pair<int, int>> cand;
    for (int cp : cpbs)
        for (int b : bpx) {
            c->cpb = cp;
            c->blocks_per_xcd = b;
            int ipx = 0;
            (void)pair_grid(c, &ipx);
            if (b > 0 && b >= ipx) continue;  // same as one block per item
            cand.push_back({cp, b});
        }
    // rounds over all candidates (the first warms up); a shape's time is its
    // fastest run, the least noisy estimate, and interleaving the candidates
    // keeps slow drifts of the clock from favouring the ones timed first
    std::vector<double> t(cand.size(), 1e300);
    for (int k = 0; k <= reps; ++k)
        for (size_t i = 0; i < cand.size(); ++i) {
            c->cpb = cand[i].first;
            c->blocks_per_xcd = cand[i].second;
            HIPCHK(c, hipMemsetAsync(c->step_idx, 0, sizeof(int), c->stream));
            TRY(launch_traverse(c));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            if (k > 0) t[i] = std::min(t[i], (double)ms);
        }
    double best = 1e300;
    int bc = c->cpb, bb = c->blocks_per_xcd;
    for (size_t i = 0; i < cand.size(); ++i)
        if (t[i] < best) {
            best = t[i];
            bc = cand[i].first;
            bb = cand[i].second;
        }
    c->cpb = bc;
    c->blocks_per_xcd = bb;
    c->timing = was_timing;
    c->counting = was_counting;
    drop_graphs(c);
    if (chosen) {
        chosen[0] = bc;
        chosen[1] = bb;
    }
    if (best_ms) *best_ms = best;
    return MDX_OK;
}

int mdx_set_precision(mdx_context *c, int precision) {
    if (!c) return MDX_ERR_ARG;
    if (precision != MDX_PREC_FP64 && precision != MDX_PREC_FP32_FP64ACC)
        return fail(c, MDX_ERR_ARG, "unknown precision");
    if (precision != c->precision) drop_graphs(c);
    c->precision = precision;
    return MDX_OK;
}

int mdx_set_lj_tables(mdx_context *c, int n_ljc, const double *eps24, const double *sig2,
                      const double *shift6) {
    if (!c || !eps24 || !sig2 || !shift6) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    if (n_ljc != c->model.nlj) return fail(c, MDX_ERR_ARG, "n_ljc differs from the model");
    size_t nn = (size_t)n_ljc * n_ljc;
    c->eps24.assign(eps24, eps24 + nn);
    c->sig2.assign(sig2, sig2 + nn);
    c->shift6.assign(shift6, shift6 + nn);
    return upload_tables(c);
}

int mdx_upload(mdx_context *c, int64_t n, const double *r, const double *v, const double *q,
               const double *L, const int32_t *cid, const uint64_t *ids) {
    if (!c || !r || !cid) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    HIPCHK(c, hipSetDevice(c->device));
    TRY(state_setup(c, n, cid));
    const int nn = (int)n;
    std::vector<double> buf((size_t)nn);
    MolArrays &m = c->mol[0];
    auto put = [&](const double *src, int stride, int k, double dflt, double *dst) -> int {
        for (int i = 0; i < nn; ++i) buf[i] = src ? src[(size_t)i * stride + k] : dflt;
        HIPCHK(c, hipMemcpy(dst, buf.data(), (size_t)nn * 8, hipMemcpyHostToDevice));
        return MDX_OK;
    };
    TRY(put(r, 3, 0, 0., m.rx));
    TRY(put(r, 3, 1, 0., m.ry));
    TRY(put(r, 3, 2, 0., m.rz));
    TRY(put(v, 3, 0, 0., m.vx));
    TRY(put(v, 3, 1, 0., m.vy));
    TRY(put(v, 3, 2, 0., m.vz));
    TRY(put(q, 4, 0, 1., m.qw));
    TRY(put(q, 4, 1, 0., m.qx));
    TRY(put(q, 4, 2, 0., m.qy));
    TRY(put(q, 4, 3, 0., m.qz));
    TRY(put(L, 3, 0, 0., m.Lx));
    TRY(put(L, 3, 1, 0., m.Ly));
    TRY(put(L, 3, 2, 0., m.Lz));
    HIPCHK(c, hipMemcpy(m.cid, cid, (size_t)nn * 4, hipMemcpyHostToDevice));
    std::vector<int> orig((size_t)nn);
    for (int i = 0; i < nn; ++i) orig[i] = i;
    HIPCHK(c, hipMemcpy(m.orig, orig.data(), (size_t)nn * 4, hipMemcpyHostToDevice));
    c->ids.assign((size_t)nn, 0);
    for (int i = 0; i < nn; ++i) c->ids[i] = ids ? ids[i] : (uint64_t)(i + 1);
    c->have_state = true;
    return MDX_OK;
}

int mdx_upload_device(mdx_context *c, int64_t n, const double *r, const double *v,
                      const double *q, const double *L, const int32_t *cid) {
    if (!c || !r || !v || !q || !L || !cid) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int32_t> hc((size_t)n);
    HIPCHK(c, hipMemcpy(hc.data(), cid, (size_t)n * 4, hipMemcpyDeviceToHost));
    TRY(state_setup(c, n, hc.data()));
    const int nn = (int)n;
    MolArrays &m = c->mol[0];
    double *dst[] = {m.rx, m.ry, m.rz, m.vx, m.vy, m.vz, m.qw, m.qx, m.qy, m.qz, m.Lx, m.Ly, m.Lz};
    const double *srcs[] = {r, r + nn, r + 2 * (size_t)nn, v, v + nn, v + 2 * (size_t)nn,
                            q, q + nn, q + 2 * (size_t)nn, q + 3 * (size_t)nn,
                            L, L + nn, L + 2 * (size_t)nn};
    // device inputs are structure-of-arrays: r = [x(n) | y(n) | z(n)] etc.
    for (int k = 0; k < 13; ++k)
        HIPCHK(c, hipMemcpyAsync(dst[k], srcs[k], (size_t)nn * 8, hipMemcpyDeviceToDevice,
                                 c->stream));
    HIPCHK(c, hipMemcpyAsync(m.cid, cid, (size_t)nn * 4, hipMemcpyDeviceToDevice, c->stream));
    std::vector<int> orig((size_t)nn);
    for (int i = 0; i < nn; ++i) orig[i] = i;
    HIPCHK(c, hipMemcpy(m.orig, orig.data(), (size_t)nn * 4, hipMemcpyHostToDevice));
    c->ids.assign((size_t)nn, 0);
    for (int i = 0; i < nn; ++i) c->ids[i] = (uint64_t)(i + 1);
    c->have_state = true;
    return MDX_OK;
}

}  // extern "C"

namespace {

// Every uploaded molecule (live or deleted) in upload order.
int download_all(mdx_context *c, double *r, double *v, double *q, double *L, double *F,
                 double *M, double *Vi, uint64_t *ids) {
    HIPCHK(c, hipSetDevice(c->device));
    const int nn = c->n;
    MolArrays &m = c->mol[c->cur];
    double *tmp = c->dl_tmp;
    if (c->point_lj && c->q_renorm > 0) {  // the pending normalisations of q
        hipLaunchKernelGGL(k_normalize_q, dim3(nblk(nn)), dim3(kBlock), 0, c->stream, nn, m,
                           (long long)c->q_renorm);
        HIPCHK(c, hipGetLastError());
        c->q_renorm = 0;
    }
    // upload-order arrays (point_lj: q, L) are copied as they are
    auto get_plain = [&](const double *a, double *out) -> int {
        HIPCHK(c, hipMemcpyAsync(out, a, (size_t)nn * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MDX_OK;
    };
    auto get = [&](const double *a, int width, int stride_in, double *out) -> int {
        if (!out) return MDX_OK;
        hipLaunchKernelGGL(k_scatter_orig, dim3(nblk(nn)), dim3(kBlock), 0, c->stream, nn, m.orig,
                           a, tmp, stride_in, width);
        HIPCHK(c, hipMemcpyAsync(out, tmp, (size_t)nn * width * 8, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MDX_OK;
    };
    // the molecule arrays are separate allocations; gather them pairwise
    auto get3 = [&](double *x, double *y, double *z, int width, double *out) -> int {
        if (!out) return MDX_OK;
        double *parts[4] = {x, y, z, nullptr};
        std::vector<double> col((size_t)nn);
        for (int w = 0; w < width; ++w) {
            TRY(get(parts[w], 1, nn, col.data()));
            for (int i = 0; i < nn; ++i) out[(size_t)i * width + w] = col[i];
        }
        return MDX_OK;
    };
    TRY(get3(m.rx, m.ry, m.rz, 3, r));
    TRY(get3(m.vx, m.vy, m.vz, 3, v));
    auto get_rot = [&](double *const *parts, int width, double *out) -> int {
        if (!out) return MDX_OK;
        std::vector<double> col((size_t)nn);
        for (int w = 0; w < width; ++w) {
            if (c->point_lj)
                TRY(get_plain(parts[w], col.data()));
            else
                TRY(get(parts[w], 1, nn, col.data()));
            for (int i = 0; i < nn; ++i) out[(size_t)i * width + w] = col[i];
        }
        return MDX_OK;
    };
    double *const lparts[3] = {m.Lx, m.Ly, m.Lz}, *const qparts[4] = {m.qw, m.qx, m.qy, m.qz};
    TRY(get_rot(lparts, 3, L));
    TRY(get_rot(qparts, 4, q));
    TRY(get(c->F, 3, nn, F));
    TRY(get(c->M, 3, nn, M));
    TRY(get(c->Vi, 3, nn, Vi));
    if (ids)
        for (int i = 0; i < nn; ++i) ids[i] = c->ids[i];
    return MDX_OK;
}

// Upload-order flags of the molecules not deleted at an outflow wall (a
// deleted molecule's position is NaN); returns their count.
int64_t live_mask(mdx_context *c, std::vector<char> &alive, int *rc) {
    std::vector<double> r(3 * (size_t)c->n);
    *rc = download_all(c, r.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    alive.assign((size_t)c->n, 1);
    int64_t k = 0;
    for (int i = 0; i < c->n; ++i) {
        alive[i] = r[3 * (size_t)i] == r[3 * (size_t)i];
        k += alive[i];
    }
    return k;
}

}  // namespace

extern "C" {

int mdx_download(mdx_context *c, int64_t n, double *r, double *v, double *q, double *L, double *F,
                 double *M, double *Vi, uint64_t *ids) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    if (!open_domain(c)) {
        if (n != c->n) return fail(c, MDX_ERR_ARG, "n differs from the uploaded count");
        return download_all(c, r, v, q, L, F, M, Vi, ids);
    }
    // outflow walls: the survivors, in upload order
    const size_t N = c->n;
    std::vector<double> rr(3 * N), vv(v ? 3 * N : 0), qq(q ? 4 * N : 0), LL(L ? 3 * N : 0),
        FF(F ? 3 * N : 0), MM(M ? 3 * N : 0), VV(Vi ? 3 * N : 0);
    std::vector<uint64_t> ii(ids ? N : 0);
    auto dp = [](std::vector<double> &a) { return a.empty() ? nullptr : a.data(); };
    TRY(download_all(c, rr.data(), dp(vv), dp(qq), dp(LL), dp(FF), dp(MM), dp(VV),
                     ids ? ii.data() : nullptr));
    int64_t live = 0;
    for (size_t i = 0; i < N; ++i) live += rr[3 * i] == rr[3 * i];
    if (n != live)
        return fail(c, MDX_ERR_ARG,
                    "n differs from the live molecule count " + std::to_string(live) +
                        " (mdx_molecule_count)");
    size_t k = 0;
    for (size_t i = 0; i < N; ++i) {
        if (rr[3 * i] != rr[3 * i]) continue;
        auto put = [&](double *out, const std::vector<double> &a, int w) {
            if (out)
                for (int d = 0; d < w; ++d) out[k * w + d] = a[i * w + d];
        };
        put(r, rr, 3);
        put(v, vv, 3);
        put(q, qq, 4);
        put(L, LL, 3);
        put(F, FF, 3);
        put(M, MM, 3);
        put(Vi, VV, 3);
        if (ids) ids[k] = ii[i];
        ++k;
    }
    return MDX_OK;
}

int mdx_molecule_count(mdx_context *c, int64_t *n) {
    if (!c || !n) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    if (!open_domain(c)) {
        *n = c->n;
        return MDX_OK;
    }
    std::vector<char> alive;
    int rc = MDX_OK;
    const int64_t k = live_mask(c, alive, &rc);
    TRY(rc);
    *n = k;
    return MDX_OK;
}

int mdx_update_container(mdx_context *c) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    HIPCHK(c, hipSetDevice(c->device));
    // molecules must be uploaded inside the box; leapfrog_pre wraps them
    // during stepping
    return launch_update(c);
}

int mdx_traverse(mdx_context *c, double *macro) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    HIPCHK(c, hipSetDevice(c->device));
    TRY(ensure_binned(c));
    if (c->counting)
        HIPCHK(c, hipMemsetAsync(c->d_counters, 0, 8 * kTypes * sizeof(unsigned long long),
                                 c->stream));
    HIPCHK(c, hipMemsetAsync(c->step_idx, 0, sizeof(int), c->stream));
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    // in the reduced-memory mode the traversal is VCP1CLJRMM's: force + kick
    TRY(launch_traverse(c, c->rmm_dt != 0.0 ? c->rmm_dt / c->model.comp[0].m : 0.0));
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    if (macro) {
        TRY(0);
        HIPCHK(c, hipMemcpyAsync(macro, c->red + 4 * c->nrep, 4 * sizeof(double) * c->nrep,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->timing) {
        float a = 0, b = 0;
        HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
        c->pair_ms = a;
        c->total_ms = b;
        for (int t = 0; t < kTypes; ++t) c->kernel_ms[t] = 0.0;
        for (int k = 0; k < kTypes && c->kev_order[k] >= 0; ++k) {
            float e = 0;
            HIPCHK(c, hipEventElapsedTime(&e, c->kev[k], c->kev[k + 1]));
            c->kernel_ms[c->kev_order[k]] = e;
        }
    }
    return MDX_OK;
}

int mdx_leapfrog_pre(mdx_context *c, double dt) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    HIPCHK(c, hipSetDevice(c->device));
    return launch_lf_pre(c, dt);
}

int mdx_leapfrog_post(mdx_context *c, double dt, double *sums) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->step_idx, 0, sizeof(int), c->stream));
    TRY(launch_lf_post(c, dt));
    if (sums) {
        HIPCHK(c, hipMemcpyAsync(sums, c->red + 8 * c->nrep, 4 * sizeof(double) * c->nrep,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDX_OK;
}

int mdx_step(mdx_context *c, int nsteps, double dt, double *per_step, int use_graph) {
    if (!c || nsteps < 0) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    if (c->rmm_dt != 0.0 && dt != c->rmm_dt)
        return fail(c, MDX_ERR_ARG, "reduced-memory mode: dt differs from mdx_set_rmm's");
    HIPCHK(c, hipSetDevice(c->device));
    TRY(ensure_steps(c, nsteps));
    const int64_t q_renorm0 = c->q_renorm;
    HIPCHK(c, hipMemsetAsync(c->step_idx, 0, sizeof(int), c->stream));
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    // Consecutive steps without a thermostat between them share one pass
    // over the molecules for step n's second and step n + 1's first half
    // kick (NVE; the reduced-memory mode kicks inside the force kernel).
    const bool fuse = nsteps > 1 && !c->thermostat && c->rmm_dt == 0.0 && !open_domain(c);
    auto kind_of = [&](int s) {
        if (!fuse) return (int)kStepFull;
        return s < nsteps - 1 ? (int)kStepFused : (int)kStepLast;
    };
    if (fuse) TRY(launch_lf_pre(c, dt));
    if (use_graph && !c->counting && !c->timing) {
        if (c->graph_dt != dt) drop_graphs(c);
        for (int s = 0; s < nsteps; ++s) {
            const int parity = c->cur, kind = kind_of(s);
            hipGraphExec_t &gx = c->gexec[kind][parity];
            if (!gx) {
                hipGraph_t graph;
                HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int rc = launch_step(c, dt, kind);
                hipError_t e = hipStreamEndCapture(c->stream, &graph);
                if (rc != MDX_OK) return rc;
                if (e != hipSuccess)
                    return fail(c, MDX_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
                HIPCHK(c, hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0));
                (void)hipGraphDestroy(graph);
                c->graph_dt = dt;
                c->cur = parity;  // capture recorded but did not run the step
            }
            HIPCHK(c, hipGraphLaunch(gx, c->stream));
            c->cur = 1 - parity;
            // the replayed re-bin rebuilt only its precision's site records
            if (!c->point_lj) c->site_forms = rebin_forms(c);
        }
    } else {
        for (int s = 0; s < nsteps; ++s) TRY(launch_step(c, dt, kind_of(s)));
    }
    // one container update per step (a captured step counted none)
    if (c->point_lj) c->q_renorm = q_renorm0 + nsteps;
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    if (per_step && nsteps > 0)
        HIPCHK(c, hipMemcpyAsync(per_step, c->step_out,
                                 (size_t)nsteps * 6 * c->nrep * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->timing) {
        float b = 0;
        HIPCHK(c, hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
        c->total_ms = b;
    }
    return MDX_OK;
}

int mdx_synchronize(mdx_context *c) {
    if (!c) return MDX_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDX_OK;
}

int mdx_cell_ids(mdx_context *c, int64_t n, int64_t *out) {
    if (!c || !out) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    if (c->nrep > 1) return fail(c, MDX_ERR_ARG, "not available for batched replicas");
    HIPCHK(c, hipSetDevice(c->device));
    TRY(ensure_binned(c));
    const size_t N = c->n;
    std::vector<int> tmp(N), orig(N), ls1(N);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->mls1id, N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(orig.data(), c->mol[c->cur].orig, N * 4, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < N; ++i) ls1[orig[i]] = tmp[i];
    // deleted molecules (outflow walls) carry -1 and are left out
    int64_t live = 0;
    for (size_t i = 0; i < N; ++i) live += ls1[i] >= 0;
    if (n != live) return fail(c, MDX_ERR_ARG, "n differs from the live molecule count");
    int64_t k = 0;
    for (size_t i = 0; i < N; ++i)
        if (ls1[i] >= 0) out[k++] = ls1[i];
    return MDX_OK;
}

int mdx_cell_geometry(mdx_context *c, int32_t *cells_per_dim, double *cell_length) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    if (c->nrep > 1) return fail(c, MDX_ERR_ARG, "not available for batched replicas");
    for (int d = 0; d < 3; ++d) {
        if (cells_per_dim) cells_per_dim[d] = c->g.cpd[d];
        if (cell_length) cell_length[d] = c->g.cl[d];
    }
    return MDX_OK;
}

int mdx_cell_pairs(mdx_context *c, int32_t *pairs, int64_t max_pairs, int64_t *count) {
    if (!c || !count) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    if (c->nrep > 1) return fail(c, MDX_ERR_ARG, "not available for batched replicas");
    HIPCHK(c, hipSetDevice(c->device));
    long tot = (long)c->ncell * 27;
    int *d;
    TRY(dalloc(c, &d, (size_t)tot * 2));
    hipLaunchKernelGGL(k_enum_pairs, dim3(nblk(tot)), dim3(kBlock), 0, c->stream, c->g, c->ncell, d);
    std::vector<int> h((size_t)tot * 2);
    HIPCHK(c, hipMemcpyAsync(h.data(), d, h.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(d);
    c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void *)d), c->allocs.end());
    // each unordered pair is visited from both cells: keep it once, non-halo
    // cell first as C08BasedTraversals::processBaseCell hands it over
    std::vector<std::pair<int, int>> v;
    v.reserve((size_t)tot);
    auto halo = [&](int e) {
        int x = e % c->g.cpd[0], y = (e / c->g.cpd[0]) % c->g.cpd[1],
            z = e / (c->g.cpd[0] * c->g.cpd[1]);
        return x == 0 || y == 0 || z == 0 || x == c->g.cpd[0] - 1 || y == c->g.cpd[1] - 1 ||
               z == c->g.cpd[2] - 1;
    };
    for (long t = 0; t < tot; ++t) {
        int a = h[2 * t], b = h[2 * t + 1];
        if (a < 0) continue;
        v.push_back({std::min(a, b), std::max(a, b)});
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    *count = (int64_t)v.size();
    if (pairs) {
        int64_t k = 0;
        for (auto &p : v) {
            if (k >= max_pairs) break;
            int a = p.first, b = p.second;
            if (halo(a)) std::swap(a, b);
            pairs[2 * k] = a;
            pairs[2 * k + 1] = b;
            ++k;
        }
    }
    return MDX_OK;
}

int mdx_lj_tables(mdx_context *c, int n_ljc, double *eps24, double *sig2, double *shift6,
                  double *shift_rc) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, false));
    if (n_ljc != c->model.nlj) return fail(c, MDX_ERR_ARG, "n_ljc differs from the model");
    const size_t nn = (size_t)n_ljc * n_ljc;
    if (eps24) std::copy(c->eps24.begin(), c->eps24.begin() + nn, eps24);
    if (sig2) std::copy(c->sig2.begin(), c->sig2.begin() + nn, sig2);
    if (shift6) std::copy(c->shift6.begin(), c->shift6.begin() + nn, shift6);
    if (shift_rc) std::copy(c->shift_rc.begin(), c->shift_rc.begin() + nn, shift_rc);
    return MDX_OK;
}

int mdx_set_charges(mdx_context *c, int ncomp, const double *q) {
    if (!c || !q) return MDX_ERR_ARG;
    TRY(need(c, true, false, false));
    if (ncomp != c->model.ncomp) return fail(c, MDX_ERR_ARG, "ncomp differs from the model");
    for (int k = 0; k < ncomp; ++k)
        for (int s = 0; s < c->model.comp[k].n[T_C]; ++s)
            c->model.comp[k].param[T_C][s] = q[k * kMaxSite + s];
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_comps, c->model.comp, sizeof(Component) * ncomp,
                             hipMemcpyHostToDevice, c->stream));
    // the site arrays carry the charges: rebuild them for the current state
    // (re-binning an unchanged state keeps its order)
    if (c->have_state) TRY(launch_update(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDX_OK;
}

}  // extern "C"

namespace {

// Device buffers of one adjoint evaluation.
struct AdjBufs {
    double *site;  // 3 (LJ sites + charge sites): per-site VJP terms
    double *cot;   // 3 n: the cotangent in the current molecule order
    double *part;  // grid x (largest K * W): per-block parameter tables
    double *out;   // kAdjMaxLJ^2 * 4 + kMaxComp * kMaxSite * 2: reduced tables
};

constexpr size_t kAdjOutDoubles = (size_t)kAdjMaxLJ * kAdjMaxLJ * 4 + kMaxComp * kMaxSite * 2;

int adjoint_check(mdx_context *c, int64_t n, const double *cot, const double *vjp) {
    if (!c) return MDX_ERR_ARG;
    TRY(need(c, true, true, true));
    if (n != c->n) return fail(c, MDX_ERR_ARG, "n differs from the uploaded count");
    if (vjp && !cot) return fail(c, MDX_ERR_ARG, "vjp requested without a cotangent");
    if (c->nrep > 1) return fail(c, MDX_ERR_ARG, "not available for batched replicas");
    if (c->nsites[T_D] > 0 || c->nsites[T_Q] > 0)
        return fail(c, MDX_ERR_ARG, "the adjoint covers LJ and charge sites only");
    if (open_domain(c))
        return fail(c, MDX_ERR_ARG, "the adjoint serves domains without outflow walls");
    if (c->model.nlj > kAdjMaxLJ)
        return fail(c, MDX_ERR_CAPACITY, "the adjoint supports <= 16 LJ centres");
    return MDX_OK;
}

// a multiple of 8 from 8 cells up (the kernel's XCD-aware cell ranges)
int adjoint_grid(const mdx_context *c) {
    const int g = std::max(1, std::min(c->ncell, kAdjBlocks));
    return g >= 8 ? g & ~7 : g;
}

// Largest per-block table (keys x values) of the LJ and charge passes.
size_t adjoint_table_width(const mdx_context *c) {
    const size_t nl = c->model.nlj;
    return std::max(nl * 4 * nl, (size_t)c->model.ncomp * kMaxSite * 2);
}

// The adjoint on the given buffers (no allocation): the LJ and charge
// passes, the per-molecule VJP, the copies back and the symmetric sums.
int adjoint_run(mdx_context *c, const double *cot, double *dU_deps, double *dU_dsig,
                double *dW_deps, double *dW_dsig, double *dU_dq, double *dW_dq, double *vjp,
                const AdjBufs &b) {
    HIPCHK(c, hipSetDevice(c->device));
    TRY(ensure_binned(c));
    TRY(ensure_sites(c, 1));
    const int nl = c->model.nlj;
    const int nn = c->n, ncomp = c->model.ncomp;
    const int nsl = c->nsites[T_LJ], nsc = c->nsites[T_C];
    MolArrays &m = c->mol[c->cur];
    if (cot) {
        HIPCHK(c, hipMemcpyAsync(c->dl_tmp, cot, (size_t)nn * 3 * 8, hipMemcpyHostToDevice,
                                 c->stream));
        hipLaunchKernelGGL(gather_orig3_kernel(), dim3(nblk(nn)), dim3(kBlock), 0, c->stream, nn,
                           m.orig, c->dl_tmp, b.cot);
    }
    const int grid = adjoint_grid(c);
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    double *out_lj = b.out, *out_q = b.out + (size_t)kAdjMaxLJ * kAdjMaxLJ * 4;
    double *g_site[2] = {b.site, b.site + 3 * (size_t)nsl};
    const int ts[2] = {T_LJ, T_C};
    for (int u = 0; u < 2; ++u) {
        const int t = ts[u];
        if (c->nsites[t] == 0) continue;
        AdjArgs a;
        memset(&a, 0, sizeof(a));
        a.g = c->g;
        a.s = site_view(c, t);
        a.pp.eps24 = c->d_eps24;
        a.pp.sig2 = c->d_sig2;
        a.pp.shift6 = c->d_shift6;
        a.pp.nlj = nl;
        a.shift_rc = c->d_shift_rc;
        a.mcid = m.cid;
        if (cot) {
            a.cx = b.cot;
            a.cy = b.cot + nn;
            a.cz = b.cot + 2 * (size_t)nn;
            const size_t ns = c->nsites[t];
            a.gx = g_site[u];
            a.gy = g_site[u] + ns;
            a.gz = g_site[u] + 2 * ns;
        }
        a.part = b.part;
        a.nlj = nl;
        a.ncell = c->ncell;
        a.site_is_mol = c->point_lj && t == T_LJ;  // site_view: mol is the identity
        a.K = t == T_LJ ? nl : ncomp * kMaxSite;
        a.W = t == T_LJ ? 4 * nl : 2;
        // the kernels that keep their sums in registers (charges; LJ with at
        // most 4 ids) place the lane columns in their tile, not here
        const bool alias = adjoint_columns_in_tile(t, nl);
        const size_t lds = ((alias ? 0 : (size_t)a.W * 64) + (size_t)a.K * a.W) * 8 + 64 * 4 +
                           (t == T_LJ ? (size_t)nl * nl * kAdjPar * 8 : 0);
        hipLaunchKernelGGL(adjoint_kernel(t, nl, t == T_LJ && c->lj_at_origin), dim3(grid), dim3(64), lds, c->stream, a);
        const int width = a.K * a.W;
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->stream, b.part, grid,
                           width, grid, t == T_LJ ? out_lj : out_q);
    }
    HIPCHK(c, hipGetLastError());
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    // host tables on the stack (no allocation): at most kAdjMaxLJ LJ ids
    double hl[kAdjMaxLJ * kAdjMaxLJ * 4], hq[kMaxComp * kMaxSite * 2];
    const size_t nhl = (size_t)nl * nl * 4, nhq = (size_t)ncomp * kMaxSite * 2;
    std::fill(hl, hl + nhl, 0.0);
    std::fill(hq, hq + nhq, 0.0);
    if (nsl) HIPCHK(c, hipMemcpyAsync(hl, out_lj, nhl * 8, hipMemcpyDeviceToHost, c->stream));
    if (nsc) HIPCHK(c, hipMemcpyAsync(hq, out_q, nhq * 8, hipMemcpyDeviceToHost, c->stream));
    if (vjp) {
        const size_t s0 = nsl, s1 = nsc;
        hipLaunchKernelGGL(adjoint_mol_kernel(), dim3(nblk(nn)), dim3(kBlock), 0, c->stream, nn,
                           site_view(c, T_LJ).mstart, c->sites[T_C].mstart,
                           nsl ? g_site[0] : nullptr, g_site[0] + s0, g_site[0] + 2 * s0,
                           nsc ? g_site[1] : nullptr, g_site[1] + s1, g_site[1] + 2 * s1,
                           c->dl_tmp);
        hipLaunchKernelGGL(k_scatter_orig, dim3(nblk(nn)), dim3(kBlock), 0, c->stream, nn, m.orig,
                           c->dl_tmp, b.cot, nn, 3);
    }
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    if (vjp)
        HIPCHK(c, hipMemcpyAsync(vjp, b.cot, (size_t)nn * 3 * 8, hipMemcpyDeviceToHost,
                                 c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->timing) {
        // kernel time: the derivative passes with their reductions, and up to
        // the assembled VJP (transfers excluded)
        float a0 = 0.f, a1 = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a0, c->ev[0], c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&a1, c->ev[0], c->ev[3]));
        c->pair_ms = a0;
        c->total_ms = a1;
    }
    // each visit holds half of its pair's term under (a, b) = (own id, other
    // id): the symmetric parameter's derivative sums both orders
    double *outs[4] = {dU_deps, dU_dsig, dW_deps, dW_dsig};
    for (int g = 0; g < 4; ++g) {
        if (!outs[g]) continue;
        for (int x = 0; x < nl; ++x)
            for (int y = 0; y < nl; ++y) {
                double v = hl[((size_t)x * nl + y) * 4 + g];
                if (x != y) v += hl[((size_t)y * nl + x) * 4 + g];
                outs[g][(size_t)x * nl + y] = v;
            }
    }
    for (int k = 0; k < ncomp * kMaxSite; ++k) {
        if (dU_dq) dU_dq[k] = hq[2 * k];
        if (dW_dq) dW_dq[k] = hq[2 * k + 1];
    }
    return MDX_OK;
}

}  // namespace

extern "C" {

int mdx_adjoint(mdx_context *c, int64_t n, const double *cot, double *dU_deps, double *dU_dsig,
                double *dW_deps, double *dW_dsig, double *dU_dq, double *dW_dq, double *vjp) {
    TRY(adjoint_check(c, n, cot, vjp));
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->adj_site) {
        const size_t nsl = c->nsites[T_LJ], nsc = c->nsites[T_C];
        TRY(dalloc(c, &c->adj_site, 3 * (nsl + nsc), true));
        TRY(dalloc(c, &c->adj_cot, 3 * (size_t)c->n, true));
        TRY(dalloc(c, &c->adj_part, (size_t)kAdjBlocks * kAdjMaxLJ * kAdjMaxLJ * 4, true));
        TRY(dalloc(c, &c->adj_out, kAdjOutDoubles, true));
    }
    return adjoint_run(c, cot, dU_deps, dU_dsig, dW_deps, dW_dsig, dU_dq, dW_dq, vjp,
                       AdjBufs{c->adj_site, c->adj_cot, c->adj_part, c->adj_out});
}

int mdx_adjoint_noalloc(mdx_context *c, int64_t n, const double *cot, double *dU_deps,
                        double *dU_dsig, double *dW_deps, double *dW_dsig, double *dU_dq,
                        double *dW_dq, double *vjp) {
    TRY(adjoint_check(c, n, cot, vjp));
    if (!c->scratch)
        return fail(c, MDX_ERR_HIP, "the scratch array of this device could not be resolved");
    // [out | part | site | cot], each part 256-byte aligned (32 doubles)
    auto up = [](size_t x) { return (x + 31) & ~(size_t)31; };
    const size_t o_part = up(kAdjOutDoubles);
    const size_t o_site = o_part + up((size_t)adjoint_grid(c) * adjoint_table_width(c));
    const size_t nsite = cot ? 3 * ((size_t)c->nsites[T_LJ] + c->nsites[T_C]) : 0;
    const size_t o_cot = o_site + up(nsite);
    const size_t total = o_cot + (cot ? 3 * (size_t)c->n : 0);
    if (total > kScratchDoubles)
        return fail(c, MDX_ERR_CAPACITY,
                    "mdx_adjoint_noalloc: the state needs " + std::to_string(total * 8) +
                        " bytes of scratch, more than the " + std::to_string(kScratchDoubles * 8) +
                        " bytes the library holds per device (use mdx_adjoint)");
    double *s = c->scratch;
    return adjoint_run(c, cot, dU_deps, dU_dsig, dW_deps, dW_dsig, dU_dq, dW_dq, vjp,
                       AdjBufs{s + o_site, s + o_cot, s + o_part, s});
}

int mdx_set_counting(mdx_context *c, int enable) {
    if (!c) return MDX_ERR_ARG;
    c->counting = enable != 0;
    return MDX_OK;
}

int mdx_counters(mdx_context *c, int64_t *out) {
    if (!c || !out) return MDX_ERR_ARG;
    unsigned long long h[8 * kTypes];
    HIPCHK(c, hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // owner computes: every test and every pair was made from both sides
    for (int k = 0; k < 8; ++k) {
        unsigned long long sum = 0;
        for (int t = 0; t < kTypes; ++t) sum += h[8 * t + k];
        out[k] = (int64_t)(sum / 2);
    }
    return MDX_OK;
}

int mdx_kernel_stats(mdx_context *c, double *ms, int64_t *counters) {
    if (!c) return MDX_ERR_ARG;
    if (ms)
        for (int t = 0; t < kTypes; ++t) ms[t] = c->kernel_ms[t];
    if (counters) {
        unsigned long long h[8 * kTypes];
        HIPCHK(c, hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < 8 * kTypes; ++k) counters[k] = (int64_t)h[k];
    }
    return MDX_OK;
}

int mdx_set_timing(mdx_context *c, int enable) {
    if (!c) return MDX_ERR_ARG;
    c->timing = enable != 0;
    return MDX_OK;
}

int mdx_last_kernel_ms(mdx_context *c, double *pair_ms, double *total_ms) {
    if (!c) return MDX_ERR_ARG;
    if (pair_ms) *pair_ms = c->pair_ms;
    if (total_ms) *total_ms = c->total_ms;
    return MDX_OK;
}

}  // extern "C"
